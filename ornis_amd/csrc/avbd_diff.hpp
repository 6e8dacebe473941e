// avbd_diff.hpp — reverse pass of the AVBD step (the adjoint kernels).
//
// The backward pass walks a recorded step backwards: the finish (BDF1 and
// sleep), then for each iteration from the last the dual update and the
// colour sweeps from the last colour, then the step's prelude (inertial
// targets and the warm start of the duals).  The recorded step (`Tape`)
// holds the step's fixed active set (pairs, points with anchors, colour
// lists, sleep states) and the solver records before the sweep and after
// every iteration, so every body solve and dual update is re-evaluated at
// exactly the inputs the forward step gave it.  Derivative formulas:
// avbd_adjoint.hpp.  Adjoints of values read by several bodies in one launch
// are accumulated with atomics.
#pragma once

#include "avbd_adjoint.hpp"
#include "avbd_kernels.hpp"

using AV = adj::V3<float>;
using AQ = adj::Q4<float>;

__device__ __forceinline__ AV av(float4 v) { return AV{v.x, v.y, v.z}; }
__device__ __forceinline__ AQ aq(float4 v) { return AQ{v.x, v.y, v.z, v.w}; }
__device__ __forceinline__ void atomic_add3(float4* p, AV v) {
  atomicAdd(&p->x, v.x);
  atomicAdd(&p->y, v.y);
  atomicAdd(&p->z, v.z);
}
__device__ __forceinline__ void atomic_add4(float4* p, AQ v) {
  atomicAdd(&p->x, v.x);
  atomicAdd(&p->y, v.y);
  atomicAdd(&p->z, v.z);
  atomicAdd(&p->w, v.w);
}

// Device adjoint state of a backward pass.
struct Adj {
  float4* xb;       // n x 2: adjoints of the body pose (position, rotation)
  float4* vb;       // n x 2: adjoints of the body velocity, angular velocity
  float4* x0b;      // n x 2: adjoints of the step-start pose (this step)
  float4* yb;       // n x 2: adjoints of the inertial targets (this step)
  float4* lb;       // cap x 2: adjoints of the current duals (lam, pen)
  float4* lb_prev;  // cap x 2: adjoints of the previous step's final duals
  float4* pb;       // pair cap: adjoints of (mu1, mu2, damper stiffness)
  float* mass;      // n: d/d mass
  float* fric;      // n: d/d friction
};

// One recorded step (device pointers).
struct Tape {
  float4* sb;        // (iters + 1) x n x 4 solver records: before the sweep, after iteration it
  float4* qrec;      // iters x qcap x 4 point records: at the start of iteration it
  size_t qcap;
  uint2* ab;
  float4* prec;
  uint2* ppts;
  uint32_t* qpair;
  int* qsrc;
  uint4* inc;        // incidence items of the step
  uint4* clist;      // colour lists
  int* colour;       // per body (-1: not swept)
  float4* rot0;      // step-start rotation
  float4* vel;       // step-start velocity
  float4* angvel;    // step-start angular velocity
  float4* inertial;  // inertial targets
  float4* inertial_rot;
  float4* inertia;   // body-frame inertia, w = mass
  uint8_t* slept;    // fell asleep at the end of the step
  uint32_t n, npairs, npts, iters;
};

__device__ __forceinline__ const float4* tape_sb(const Tape& t, uint32_t it) { return t.sb + (size_t)it * t.n * 4; }
__device__ __forceinline__ const float4* tape_q(const Tape& t, uint32_t it) { return t.qrec + (size_t)it * t.qcap * 4; }

// Contact of point q at solver records sa / sb (pose, rotation, start pose,
// rotation delta) and point record qr; rotations at the step start from
// the tape.
__device__ __forceinline__ adj::Contact<float> make_contact(const Tape& t, uint32_t p, const float4* qr,
                                                            const float4* sa, const float4* sbb, uint2 ab) {
  adj::Contact<float> c;
  const float4 n4 = t.prec[4 * (size_t)p], t14 = t.prec[4 * (size_t)p + 1], t24 = t.prec[4 * (size_t)p + 2],
               mu = t.prec[4 * (size_t)p + 3];
  c.n = av(n4);
  c.t1 = av(t14);
  c.t2 = av(t24);
  c.mu1 = mu.x;
  c.mu2 = mu.y;
  c.damp = t14.w;
  c.touching = n4.w <= 0.0f;
  c.ra = av(qr[0]);
  c.rb = av(qr[1]);
  for (int k = 0; k < 3; ++k) {
    c.lam[k] = k == 0 ? qr[2].x : (k == 1 ? qr[2].y : qr[2].z);
    c.pen[k] = k == 0 ? qr[3].x : (k == 1 ? qr[3].y : qr[3].z);
  }
  c.rad_a = sa[2].w;
  c.rad_b = sbb[2].w;
  c.xa = av(sa[0]);
  c.qa = aq(sa[1]);
  c.xa0 = av(sa[2]);
  c.qa0 = aq(t.rot0[ab.x]);
  c.xb = av(sbb[0]);
  c.qb = aq(sbb[1]);
  c.xb0 = av(sbb[2]);
  c.qb0 = aq(t.rot0[ab.y]);
  return c;
}

// Reverse of BDF1 and sleep (mod.rs:1062-1122): the adjoints of the step's
// outputs (pose, velocities) become adjoints of the last sweep's pose and
// of the step-start pose.
__global__ void k_adj_finish(Tape t, Adj a) {
  const uint32_t h = gid();
  if (h >= t.n || t.colour[h] < 0) return;
  const float4* r = tape_sb(t, t.iters) + 4 * (size_t)h;
  const AV x0 = av(r[2]);
  const AQ q = aq(r[1]), q0 = aq(t.rot0[h]);
  const bool slept = t.slept[h] != 0;
  const AV vbar = av(a.vb[2 * h]), wbar = av(a.vb[2 * h + 1]);
  AV xbar = av(a.xb[2 * h]);
  AQ qbar = adj::qzero<float>(), q0bar = adj::qzero<float>();
  adj::qnormalize_vjp(q, aq(a.xb[2 * h + 1]), &qbar);
  AV x0bar = adj::zero<float>();
  if (!slept) {
    xbar += vbar * (1.0f / K_DT);
    x0bar -= vbar * (1.0f / K_DT);
    const AV spin = adj::quat_diff_vec(q, q0);
    if (!(adj::len(spin) < K_SPIN_REST)) adj::quat_diff_vec_vjp(q, q0, wbar * (1.0f / K_DT), &qbar, &q0bar);
  }
  a.xb[2 * h] = make_float4(xbar.x, xbar.y, xbar.z, 0.0f);
  a.xb[2 * h + 1] = make_float4(qbar.x, qbar.y, qbar.z, qbar.w);
  float4 o0 = a.x0b[2 * h], o1 = a.x0b[2 * h + 1];
  a.x0b[2 * h] = make_float4(o0.x + x0bar.x, o0.y + x0bar.y, o0.z + x0bar.z, 0.0f);
  a.x0b[2 * h + 1] = make_float4(o1.x + q0bar.x, o1.y + q0bar.y, o1.z + q0bar.z, o1.w + q0bar.w);
  a.vb[2 * h] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  a.vb[2 * h + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ bool adj_target(const Dev& d, uint32_t h) { return d.b.type[h] == BT_DYNAMIC; }

// Scatter a Contact's body adjoints (side A / B) to the bodies.
__device__ __forceinline__ void scatter_side(const Dev& d, Adj& a, uint32_t h, AV x, AQ q, AV x0, AQ q0) {
  if (!adj_target(d, h)) return;
  atomic_add3(&a.xb[2 * h], x);
  atomic_add4(&a.xb[2 * h + 1], q);
  atomic_add3(&a.x0b[2 * h], x0);
  atomic_add4(&a.x0b[2 * h + 1], q0);
}

// Reverse of iteration it's dual update (joints_ext.rs:464-573), one
// thread per point.
__global__ void k_adj_dual(Dev d, Tape t, Adj a, uint32_t it) {
  const uint32_t q = gid();
  if (q >= t.npts) return;
  const uint32_t p = t.qpair[q];
  const uint2 ab = t.ab[p];
  const float4* rec = tape_sb(t, it + 1);
  const adj::Contact<float> c =
      make_contact(t, p, tape_q(t, it) + 4 * (size_t)q, rec + 4 * (size_t)ab.x, rec + 4 * (size_t)ab.y, ab);
  const float4 l0 = a.lb[2 * q], l1 = a.lb[2 * q + 1];
  const float lb_o[3] = {l0.x, l0.y, l0.z}, pb_o[3] = {l1.x, l1.y, l1.z};
  adj::ContactBar<float> o;
  adj::clear(o);
  adj::dual_vjp(c, lb_o, pb_o, o);
  a.lb[2 * q] = make_float4(o.lam[0], o.lam[1], o.lam[2], 0.0f);
  a.lb[2 * q + 1] = make_float4(o.pen[0], o.pen[1], o.pen[2], 0.0f);
  scatter_side(d, a, ab.x, o.xa, o.qa, o.xa0, o.qa0);
  scatter_side(d, a, ab.y, o.xb, o.qb, o.xb0, o.qb0);
  if (o.mu1 != 0.0f || o.mu2 != 0.0f) {
    atomicAdd(&a.pb[p].x, o.mu1);
    atomicAdd(&a.pb[p].y, o.mu2);
  }
}

// Partner record at the moment a body of colour c is solved in iteration
// it: bodies of earlier colours were already solved in this iteration.
struct TapePartner {
  const float4* cur;
  const float4* next;
  const int* colour;
  int c;
  __device__ __forceinline__ const float4* operator()(uint32_t h) const {
    const int k = colour[h];
    return (k >= 0 && k < c ? next : cur) + 4 * (size_t)h;
  }
};

// Reverse of the body solves of colour c in iteration it (solve_body,
// assembly.rs:73-903), sixteen lanes per body as in k_primal8: the forward
// solve is re-evaluated (same rows, same sums), then with s_bar the adjoint
// of the solve's step (from the pose adjoint through quat_integrate) and
// w = H^-1 s_bar, every point-side adds -d(w^T g + w^T H dx)/d(inputs).
template <int W>
__global__ void __launch_bounds__(WAVE* W) k_adj_colour(Dev d, Tape t, Adj a, uint32_t it, int c,
                                                      const uint4* __restrict__ list, uint32_t count) {
  constexpr int NB = (WAVE / BODY_LANES) * W;
  const int tid = threadIdx.x;
  const int slot = tid % BODY_LANES;
  const uint32_t idx = blockIdx.x * NB + tid / BODY_LANES;
  const bool valid = idx < count;
  uint32_t h = 0, i0 = 0, np = 0;
  if (valid) {
    const uint4 cr = list[idx];
    h = cr.x & 0x7fffffffu;
    i0 = cr.y;
    np = cr.z;
  }
  Dev dd = d;
  dd.P.rec = t.prec;
  dd.Q.rec = const_cast<float4*>(tape_q(t, it));
  const float4* cur = tape_sb(t, it);
  const float4* next = tape_sb(t, it + 1);
  const TapePartner partner{cur, next, t.colour, c};
  float4 own[4];
  {
    const float4* rec = cur + 4 * (size_t)h;
    own[0] = rec[0];
    own[1] = rec[1];
    own[2] = rec[2];
    own[3] = rec[3];
  }
  // forward re-evaluation of the solve
  float acc[27];
#pragma unroll
  for (int e = 0; e < 27; ++e) acc[e] = 0.0f;
  auto add_row = [&](int kind, const Row& row) {
    if (kind == 0)
      acc_stamp(acc, row);
    else if (kind == 1)
      acc_couple(acc, row);
    else
      acc_diag(acc, row);
  };
  for (uint32_t k = slot; k < np; k += BODY_LANES) point_rows<4>(dd, t.inc[i0 + k], own, add_row, partner);
  InertialIn in;
  in.p = own[0];
  in.q = own[1];
  in.inertia = t.inertia[h];
  in.target = t.inertial[h];
  in.target_rot = t.inertial_rot[h];
  float blk[28];
  inertial_block(in, blk);
  float L21[21], rhs[6];
#pragma unroll
  for (int e = 0; e < 21; ++e) L21[e] = blk[e] + tree16(acc[e]);
#pragma unroll
  for (int k = 0; k < 6; ++k) rhs[k] = -(blk[21 + k] + tree16(acc[21 + k]));
  float dx[6];
  const bool solved = valid && solve_6x6_dev(L21, rhs, dx);
  if (!solved) return;  // a breakdown kept the pose: its adjoints pass through
  // adjoint of the step dx from the pose adjoint
  const AQ q = aq(own[1]);
  const AV xbar_out = av(a.xb[2 * h]);
  const AQ qbar_out = aq(a.xb[2 * h + 1]);
  AQ qbar_dir = adj::qzero<float>();
  AV sang = adj::zero<float>();
  adj::quat_integrate_vjp(q, AV{dx[3], dx[4], dx[5]}, qbar_out, &qbar_dir, &sang);
  const float sbar[6] = {xbar_out.x, xbar_out.y, xbar_out.z, sang.x, sang.y, sang.z};
  float w[6];
  if (!adj::solve6(L21, sbar, w)) return;
  // point-sides of this lane
  AV ox = adj::zero<float>(), ox0 = adj::zero<float>();
  AQ oq = adj::qzero<float>(), oq0 = adj::qzero<float>();
  for (uint32_t k = slot; k < np; k += BODY_LANES) {
    const uint4 itm = t.inc[i0 + k];
    const uint32_t qi = itm.x & 0x7fffffffu;
    const bool is_a = (itm.x >> 31) == 0u;
    const uint32_t p = itm.y;
    const uint2 ab = make_uint2(itm.z, itm.w);
    const float4* ra = is_a ? own : partner(ab.x);
    const float4* rb = is_a ? partner(ab.y) : own;
    const adj::Contact<float> cc = make_contact(t, p, tape_q(t, it) + 4 * (size_t)qi, ra, rb, ab);
    adj::ContactBar<float> o;
    adj::clear(o);
    adj::point_side_vjp(cc, is_a, w, dx, o);
    if (is_a) {
      ox += o.xa;
      oq += o.qa;
      ox0 += o.xa0;
      oq0 += o.qa0;
      scatter_side(d, a, ab.y, o.xb, o.qb, o.xb0, o.qb0);
    } else {
      ox += o.xb;
      oq += o.qb;
      ox0 += o.xb0;
      oq0 += o.qb0;
      scatter_side(d, a, ab.x, o.xa, o.qa, o.xa0, o.qa0);
    }
    float4 l0 = a.lb[2 * qi], l1 = a.lb[2 * qi + 1];
    a.lb[2 * qi] = make_float4(l0.x + o.lam[0], l0.y + o.lam[1], l0.z + o.lam[2], 0.0f);
    a.lb[2 * qi + 1] = make_float4(l1.x + o.pen[0], l1.y + o.pen[1], l1.z + o.pen[2], 0.0f);
    if (o.mu1 != 0.0f || o.mu2 != 0.0f || o.damp != 0.0f) {
      atomicAdd(&a.pb[p].x, o.mu1);
      atomicAdd(&a.pb[p].y, o.mu2);
      atomicAdd(&a.pb[p].z, o.damp);
    }
  }
  // the body's own terms, summed over its lanes
  float v[14] = {ox.x, ox.y, ox.z, oq.x, oq.y, oq.z, oq.w, ox0.x, ox0.y, ox0.z, oq0.x, oq0.y, oq0.z, oq0.w};
#pragma unroll
  for (int e = 0; e < 14; ++e) v[e] = tree16(v[e]);
  if (slot != 0) return;
  // inertial terms
  const float m = in.inertia.w;
  const AV dunit{in.inertia.x / m, in.inertia.y / m, in.inertia.z / m};
  AV xb_in = adj::zero<float>(), yb = adj::zero<float>();
  AQ qb_in = adj::qzero<float>(), qyb = adj::qzero<float>();
  float mb = 0.0f;
  adj::inertial_vjp(av(own[0]), q, av(in.target), aq(in.target_rot), m, dunit, w, dx, K_DT, &xb_in, &qb_in, &yb, &qyb,
                    &mb);
  a.xb[2 * h] = make_float4(xbar_out.x + v[0] + xb_in.x, xbar_out.y + v[1] + xb_in.y, xbar_out.z + v[2] + xb_in.z, 0.0f);
  a.xb[2 * h + 1] = make_float4(qbar_dir.x + v[3] + qb_in.x, qbar_dir.y + v[4] + qb_in.y, qbar_dir.z + v[5] + qb_in.z,
                                qbar_dir.w + v[6] + qb_in.w);
  float4 o0 = a.x0b[2 * h], o1 = a.x0b[2 * h + 1];
  a.x0b[2 * h] = make_float4(o0.x + v[7], o0.y + v[8], o0.z + v[9], 0.0f);
  a.x0b[2 * h + 1] = make_float4(o1.x + v[10], o1.y + v[11], o1.z + v[12], o1.w + v[13]);
  float4 y0 = a.yb[2 * h], y1 = a.yb[2 * h + 1];
  a.yb[2 * h] = make_float4(y0.x + yb.x, y0.y + yb.y, y0.z + yb.z, 0.0f);
  a.yb[2 * h + 1] = make_float4(y1.x + qyb.x, y1.y + qyb.y, y1.z + qyb.z, y1.w + qyb.w);
  a.mass[h] += mb;
}

// Reverse of the step's prelude (mod.rs:993-1049): the pose adjoints at the
// start of the sweep are adjoints of the inertial targets y = x0 + v dt +
// g dt^2, q_y = quat_integrate(q0, w dt); with the step-start adjoints they
// become the adjoints of the previous step's outputs.  Bodies not swept
// keep their state through the step.
__global__ void k_adj_prelude(Dev d, Tape t, Adj a) {
  const uint32_t h = gid();
  if (h >= t.n || !adj_target(d, h)) return;
  const float4 x0b = a.x0b[2 * h], q0b4 = a.x0b[2 * h + 1];
  if (t.colour[h] < 0) {
    const float4 xb = a.xb[2 * h], qb = a.xb[2 * h + 1];
    a.xb[2 * h] = make_float4(xb.x + x0b.x, xb.y + x0b.y, xb.z + x0b.z, 0.0f);
    a.xb[2 * h + 1] = make_float4(qb.x + q0b4.x, qb.y + q0b4.y, qb.z + q0b4.z, qb.w + q0b4.w);
  } else {
    const float4 xb = a.xb[2 * h], qb = a.xb[2 * h + 1], yb = a.yb[2 * h], qyb = a.yb[2 * h + 1];
    const AV ybar{xb.x + yb.x, xb.y + yb.y, xb.z + yb.z};
    const AQ qybar{qb.x + qyb.x, qb.y + qyb.y, qb.z + qyb.z, qb.w + qyb.w};
    const AQ q0 = aq(t.rot0[h]);
    const AV wdt = av(t.angvel[h]) * K_DT;
    AQ q0bar{q0b4.x, q0b4.y, q0b4.z, q0b4.w};
    AV wdtbar = adj::zero<float>();
    adj::quat_integrate_vjp(q0, wdt, qybar, &q0bar, &wdtbar);
    a.xb[2 * h] = make_float4(x0b.x + ybar.x, x0b.y + ybar.y, x0b.z + ybar.z, 0.0f);
    a.xb[2 * h + 1] = make_float4(q0bar.x, q0bar.y, q0bar.z, q0bar.w);
    a.vb[2 * h] = make_float4(ybar.x * K_DT, ybar.y * K_DT, ybar.z * K_DT, 0.0f);
    a.vb[2 * h + 1] = make_float4(wdtbar.x * K_DT, wdtbar.y * K_DT, wdtbar.z * K_DT, 0.0f);
  }
  a.x0b[2 * h] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  a.x0b[2 * h + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  a.yb[2 * h] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  a.yb[2 * h + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Reverse of the warm start (rows.rs:810-904 with the decay of mod.rs:
// 1021-1049): lam0 = alpha gamma lam_prev[src], pen0 = clamp(gamma
// pen_prev[src], PEN_MIN, PEN_MAX).  Two points may inherit from one.
__global__ void k_adj_warm(Tape t, Adj a) {
  const uint32_t q = gid();
  if (q >= t.npts) return;
  const int src = t.qsrc[q];
  if (src < 0) return;
  const float4 l0 = a.lb[2 * q], l1 = a.lb[2 * q + 1];
  const float4 pen0 = tape_q(t, 0)[4 * (size_t)q + 3];
  const float AG = K_ALPHA * K_GAMMA;
  atomic_add3(&a.lb_prev[2 * (size_t)src], AV{l0.x * AG, l0.y * AG, l0.z * AG});
  const float pv[3] = {pen0.x, pen0.y, pen0.z}, pb[3] = {l1.x, l1.y, l1.z};
  float o[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) o[k] = (pv[k] > K_PEN_MIN && pv[k] < K_PEN_MAX) ? pb[k] * K_GAMMA : 0.0f;
  atomic_add3(&a.lb_prev[2 * (size_t)src + 1], AV{o[0], o[1], o[2]});
}

// Pair parameters: friction (mu = max of the bodies' coefficients, the
// larger one's body takes the adjoint; ties the second, as fmax_ picks) and
// the separated-pair damper stiffness min(2 m_red / dt^2 / npts, PEN_MAX),
// m_red = 1 / (k_a + k_b), k = 1 / mass.
__global__ void k_adj_pairs(Dev d, Tape t, Adj a) {
  const uint32_t p = gid();
  if (p >= t.npairs) return;
  const float4 pb = a.pb[p];
  a.pb[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const uint2 ab = t.ab[p];
  if (pb.x != 0.0f || pb.y != 0.0f) {
    const float fa = d.b.mat[ab.x].x, fb = d.b.mat[ab.y].x;
    atomicAdd(&a.fric[fa > fb ? ab.x : ab.y], pb.x + pb.y);
  }
  const float damp = t.prec[4 * (size_t)p + 1].w;
  if (pb.z != 0.0f && damp > 0.0f && damp < K_PEN_MAX) {
    const float4* r = tape_sb(t, 0);
    const float ka = d.b.type[ab.x] == BT_DYNAMIC ? r[4 * (size_t)ab.x].w : 0.0f;
    const float kb = d.b.type[ab.y] == BT_DYNAMIC ? r[4 * (size_t)ab.y].w : 0.0f;
    const uint2 pts = t.ppts[p];
    const float npts = (float)(pts.y > 1u ? pts.y : 1u);
    const float mred = 1.0f / (ka + kb);
    // d damp / d m = 2 m_red^2 k^2 / (dt^2 npts) for k = 1 / m > 0
    const float g = pb.z * 2.0f * mred * mred / (K_DT * K_DT) / npts;
    if (ka > 0.0f) atomicAdd(&a.mass[ab.x], g * ka * ka);
    if (kb > 0.0f) atomicAdd(&a.mass[ab.y], g * kb * kb);
  }
}
