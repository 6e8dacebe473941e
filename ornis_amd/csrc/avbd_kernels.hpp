// avbd_kernels.hpp — the HIP kernels of one AVBD step on gfx950.
//
// Pipeline of a fixed 1/60 s substep (step_inner, crates/physics/src/avbd/
// mod.rs:846-1143), every phase on the device:
//   k_begin            torque -> angular velocity, TOI mover flags
//   k_grid_*           hash grid of the bodies' bounding spheres
//   k_toi_*           linear TOI clamp against the step-top poses, one
//                      thread per (mover, target) cast
//   k_bp_*             candidate pairs, canonically ordered by (a, b)
//   k_narrow           discover_pair: prefilters, signed gap, normal,
//                      manifold points, wake vote (rows.rs:432-756)
//   k_merge            persistent-dual merge into the new pair table
//                      (rows.rs:795-931) fused with the warm-start decay
//   k_inertial etc.    inertial targets and warm-started positions
//   k_inc_* / k_jp_*   per-body incidence lists and the greedy colouring
//   k_primal<L>        L lanes per body: rows staged in LDS, summed
//                      per matrix element in row order, 6x6 LDL^T in FP32
//   k_dual / joints    dual / penalty update, one thread per point
//   k_bdf1 / sleep     velocities, renormalisation, sleep bookkeeping
#pragma once

#include <type_traits>

#include "avbd_geom.hpp"
#include "avbd_group.hpp"
#include "avbd_state.hpp"

#define TPB 256

__device__ __forceinline__ uint32_t gid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Adds the block's sum of v to *ctr with one atomic (every thread of the
// TPB-thread block must call it).  Counters hit by a whole launch use this:
// atomics on one address from every wavefront serialise.
__device__ __forceinline__ void block_add(uint32_t* ctr, uint32_t v) {
  __shared__ uint32_t part[TPB / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (__lane_id() == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < TPB / 64; ++w) t += part[w];
    if (t) atomicAdd(ctr, t);
  }
}

// ---------------------------------------------------------------------------
// counts the device derives (a step never needs the host to read them)
// ---------------------------------------------------------------------------
// TOI casts and broadphase candidates: their scan totals, or zero when the
// total exceeds the buffer's capacity (the fill kernels then record the
// overflow and write nothing, so every later kernel stays in bounds).
__device__ __forceinline__ uint32_t dev_ncast(const Dev& d) {
  const uint32_t t = d.toi_off[d.n];
  return t <= d.cap_toi ? t : 0u;
}
__device__ __forceinline__ uint32_t dev_ncand(const Dev& d) {
  const uint32_t t = d.cand_off[d.n];
  return t <= d.cap_cand ? t : 0u;
}
// This step's pairs and points (written by k_pair_offsets).
__device__ __forceinline__ uint32_t dev_npairs(const Dev& d) { return d.dc[DC_NPAIRS]; }
__device__ __forceinline__ uint32_t dev_npts(const Dev& d) { return d.dc[DC_NPTS]; }
__device__ __forceinline__ void note_overflow(const Dev& d, uint32_t word, uint32_t need) {
  atomicOr(d.ovf + OVF_FLAGS, 1u << word);
  atomicMax(d.ovf + word, need);
}

// ---------------------------------------------------------------------------
// body helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool has_axis(float4 fd) { return fd.x != 0.0f || fd.y != 0.0f || fd.z != 0.0f; }

__device__ __forceinline__ f3 shape_inertia_dev(int kind, f3 p, float mass) {
  if (!(mass > 0.0f) || !isfinite(mass)) return zero3();
  if (kind == SH_SPHERE) {
    float v = 0.4f * mass * p.x * p.x;
    return f3{v, v, v};
  }
  if (kind == SH_BOX) {
    f3 s = mul(p, 2.0f);
    float x = s.x * s.x, y = s.y * s.y, z = s.z * s.z;
    float k = mass / 12.0f;
    return f3{k * (y + z), k * (z + x), k * (x + y)};
  }
  if (kind == SH_CAPSULE) return capsule_inertia(p.x, p.y, mass);
  return zero3();
}

// wake_body (crates/physics/src/avbd/mod.rs:833-842; body.rs:410-418).
__device__ __forceinline__ void wake_body_dev(const Dev& d, uint32_t h) {
  if (!d.b.asleep[h]) return;
  d.b.asleep[h] = 0;
  d.b.timer[h] = 0.0f;
  if (d.b.type[h] == BT_DYNAMIC) {
    float4 in = d.b.inertia[h];
    float4 sh = d.b.shape[h];
    f3 I = shape_inertia_dev(d.b.kind[h], xyz(sh), in.w);
    d.b.inertia[h] = pk(I, in.w);
    float4 p = d.b.pos[h];
    p.w = 1.0f / in.w;
    d.b.pos[h] = p;
  }
}

__device__ __forceinline__ shp body_shape(const Dev& d, uint32_t h) {
  float4 sh = d.b.shape[h];
  return shp{d.b.kind[h], xyz(sh), xyz(d.b.pos[h]), asq(d.b.rot[h])};
}

__device__ __forceinline__ bool pair_allowed_dev(const Dev& d, uint32_t a, uint32_t b) {
  return (d.b.layer[a] & d.b.mask[b]) != 0u && (d.b.layer[b] & d.b.mask[a]) != 0u;
}

__device__ __forceinline__ bool jointed_dev(const Dev& d, uint32_t a, uint32_t b) {
  for (uint32_t k = d.b.jp_off[a]; k < d.b.jp_off[a + 1]; ++k)
    if (d.b.jp[k] == b) return true;
  return false;
}

// ---------------------------------------------------------------------------
// phase 1: torque and TOI mover flags (mod.rs:855-871, :882-895)
// ---------------------------------------------------------------------------
// The TOI gate (mod.rs:882-895) of a solvable body: awake, not a trigger,
// and its step displacement above K_CCD_GATE of its smallest dimension.
__device__ __forceinline__ uint8_t fast_mover(const Dev& d, uint32_t h) {
  if (d.b.asleep[h] || d.b.trigger[h]) return 0;
  f3 disp = mul(xyz(d.b.vel[h]), K_DT);
  float4 sh = d.b.shape[h];
  return !(len3(disp) <= K_CCD_GATE * shape_min_dim(d.b.kind[h], xyz(sh))) ? 1 : 0;
}

__device__ __forceinline__ uint8_t begin_body(const Dev& d, uint32_t h) {
  float4 p = d.b.pos[h];
  bool solvable = d.b.type[h] == BT_DYNAMIC && p.w > 0.0f;
  uint8_t fast = 0;
  if (solvable) {
    float4 in = d.b.inertia[h];
    f3 I = xyz(in);
    qf q = asq(d.b.rot[h]);
    float iw[3][3], iwi[3][3];
    world_inertia(I, q, iw);
    f3 dg{I.x > 0.0f ? I.x : 0.0f, I.y > 0.0f ? I.y : 0.0f, I.z > 0.0f ? I.z : 0.0f};
    inverse_symmetric(iw, dg, iwi);
    float4 tq = d.b.torque[h];
    d.b.torque[h] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 w = d.b.angvel[h];
    f3 nw = add(xyz(w), mul(mat3_vec(iwi, xyz(tq)), K_DT));
    d.b.angvel[h] = pk(nw, w.w);
    fast = fast_mover(d, h);
  }
  d.fast[h] = fast;
  return fast;
}
__global__ void __launch_bounds__(TPB) k_begin(Dev d) {
  uint32_t h = gid();
  uint8_t fast = 0;
  if (h < d.n) fast = begin_body(d, h);
  block_add(&d.counters[CTR_FAST], fast);
}

// The inverse world inertia of body h applied to w (ang_inv_wa,
// joints_ext.rs:423-438); zero unless the body is solvable.
__device__ __forceinline__ f3 ang_inv_wa_dev(const Dev& d, uint32_t h, f3 w) {
  if (!(d.b.type[h] == BT_DYNAMIC && d.b.pos[h].w > 0.0f)) return zero3();
  const f3 I = xyz(d.b.inertia[h]);
  float iw[3][3], iwi[3][3];
  world_inertia(I, asq(d.b.rot[h]), iw);
  inverse_symmetric(iw, f3{fmax_(I.x, 0.0f), fmax_(I.y, 0.0f), fmax_(I.z, 0.0f)}, iwi);
  return mat3_vec(iwi, w);
}

// A Revolute / Prismatic joint's live angle / slide from its assembly value
// (rows.rs:149-151; assembly.rs:811-849).
__device__ __forceinline__ float limit_value(const Dev& d, uint32_t j, int kind, qf qa, qf qb, f3 pa, f3 pb, f3 wa) {
  if (kind == JK_REVOLUTE) return wrap_pi(hinge_twist(qa, qb, xyz(d.J.ax[2 * j])) - d.J.dref[j].x);
  return dot3(sub(pb, pa), wa) - d.J.dref[j].x;
}
// limit_state / warm_limit_state (rows.rs:156-182): 1 lower, 0 upper, -1 free.
__device__ __forceinline__ int limit_state(float v, float lo, float hi, float slop) {
  if (v <= lo + slop) return 1;
  if (v >= hi - slop) return 0;
  return -1;
}
__device__ __forceinline__ int warm_limit_state(float v, float lo, float hi, float slop, float force) {
  const int st = limit_state(v, lo, hi, slop);
  if (st >= 0) return st;
  return force < 0.0f ? 1 : (force > 0.0f ? 0 : -1);
}
// regularized_limit (rows.rs:186-191).
__device__ __forceinline__ float regularized_limit(float value, float initial, float bound, bool lower) {
  const float old = initial - bound;
  const float violation = lower ? fmin_(old, 0.0f) : fmax_(old, 0.0f);
  return value - bound - K_ALPHA * violation;
}

// motor_impulse (joints_ext.rs:237-356) of one Revolute / Prismatic joint
// with a velocity motor: the budget-clamped deadbeat impulse driving the
// axis rate to its target (JointMotor::velocity_impulse, joint.rs:
// 1172-1177), paused while the limit is violated (joints_ext.rs:440-462).
__device__ void motor_impulse_one(const Dev& d, uint32_t j) {
  const int kind = d.J.kind[j];
  const uint2 ab = d.J.ab[j];
  const float4 pr = d.J.dref[j], mo = d.J.qref[j];
  if (!(mo.y > 0.0f)) return;
  const float4 pa4 = d.b.pos[ab.x], pb4 = d.b.pos[ab.y];
  const qf qa = asq(d.b.rot[ab.x]), qb = asq(d.b.rot[ab.y]);
  const f3 wa = normalize_or(qrot(qa, xyz(d.J.ax[2 * j])), f3{0.0f, 0.0f, 1.0f});
  if ((uint32_t)pr.w & HINGE_LIMIT) {
    const f3 pa = add(xyz(pa4), qrot(qa, xyz(d.J.la[j]))), pb = add(xyz(pb4), qrot(qb, xyz(d.J.lb[j])));
    const float v = limit_value(d, j, kind, qa, qb, pa, pb, wa);
    if (limit_state(v, pr.y, pr.z, kind == JK_REVOLUTE ? K_LIMIT_SLOP_ANG : K_LIMIT_SLOP_LIN) >= 0) return;
  }
  const bool sa = d.b.type[ab.x] == BT_DYNAMIC && pa4.w > 0.0f, sbv = d.b.type[ab.y] == BT_DYNAMIC && pb4.w > 0.0f;
  const float cap = mo.y * K_DT;
  if (kind == JK_REVOLUTE || kind == JK_WHEEL) {  // a wheel turns about its axle (ax holds it)
    const float4 wa4 = d.b.angvel[ab.x], wb4 = d.b.angvel[ab.y];
    const float w = dot3(sub(xyz(wb4), xyz(wa4)), wa);
    const f3 ia = ang_inv_wa_dev(d, ab.x, wa), ib = ang_inv_wa_dev(d, ab.y, wa);
    const float k = dot3(ia, wa) + dot3(ib, wa);
    if (k < K_MIN_EFF_MASS) return;
    const float dj = clampf_((mo.x - w) / k, -cap, cap);
    if (sa) d.b.angvel[ab.x] = pk(add(xyz(wa4), mul(ia, -dj)), wa4.w);
    if (sbv) d.b.angvel[ab.y] = pk(add(xyz(wb4), mul(ib, dj)), wb4.w);
  } else {
    const float4 va4 = d.b.vel[ab.x], vb4 = d.b.vel[ab.y];
    const float v = dot3(sub(xyz(vb4), xyz(va4)), wa);
    const float ka = d.b.type[ab.x] == BT_DYNAMIC ? pa4.w : 0.0f;
    const float kb = d.b.type[ab.y] == BT_DYNAMIC ? pb4.w : 0.0f;
    const float k = ka + kb;
    if (k < K_MIN_EFF_MASS) return;
    const float dj = clampf_((mo.x - v) / k, -cap, cap);
    if (sa) d.b.vel[ab.x] = pk(add(xyz(va4), mul(wa, -dj * ka)), va4.w);
    if (sbv) d.b.vel[ab.y] = pk(add(xyz(vb4), mul(wa, dj * kb)), vb4.w);
  }
}

// free_motor_impulse (joints_ext.rs:368-421) of a Motor joint: per world
// axis, the budget-clamped deadbeat linear then angular impulse through the
// pair's effective mass at the centres (MotorDrive::linear_impulse /
// angular_impulse, joint.rs:485-499).
__device__ void free_motor_impulse(const Dev& d, uint32_t j) {
  const uint2 ab = d.J.ab[j];
  const float4 lt = d.J.ax[2 * j], at = d.J.ax[2 * j + 1];
  for (int e = 0; e < 3; ++e) {
    f3 axis = zero3();
    setc(axis, e, 1.0f);
    const float4 pa4 = d.b.pos[ab.x], pb4 = d.b.pos[ab.y];
    const bool sa = d.b.type[ab.x] == BT_DYNAMIC && pa4.w > 0.0f, sbv = d.b.type[ab.y] == BT_DYNAMIC && pb4.w > 0.0f;
    const float4 va4 = d.b.vel[ab.x], vb4 = d.b.vel[ab.y];
    const float v = dot3(sub(xyz(vb4), xyz(va4)), axis);
    const float ka_lin = d.b.type[ab.x] == BT_DYNAMIC ? pa4.w : 0.0f;
    const float kb_lin = d.b.type[ab.y] == BT_DYNAMIC ? pb4.w : 0.0f;
    const float kl = ka_lin + kb_lin;
    float dj = 0.0f;
    if (lt.w > 0.0f && kl > 0.0f) dj = clampf_((dot3(xyz(lt), axis) - v) / kl, -lt.w * K_DT, lt.w * K_DT);
    if (dj != 0.0f) {
      if (sa) d.b.vel[ab.x] = pk(add(xyz(va4), mul(axis, -dj * ka_lin)), va4.w);
      if (sbv) d.b.vel[ab.y] = pk(add(xyz(vb4), mul(axis, dj * kb_lin)), vb4.w);
    }
    const float4 wa4 = d.b.angvel[ab.x], wb4 = d.b.angvel[ab.y];
    const float w = dot3(sub(xyz(wb4), xyz(wa4)), axis);
    const f3 ia = ang_inv_wa_dev(d, ab.x, axis), ib = ang_inv_wa_dev(d, ab.y, axis);
    const float ka = dot3(ia, axis) + dot3(ib, axis);
    float dw = 0.0f;
    if (at.w > 0.0f && ka > 0.0f) dw = clampf_((dot3(xyz(at), axis) - w) / ka, -at.w * K_DT, at.w * K_DT);
    if (dw != 0.0f) {
      if (sa) d.b.angvel[ab.x] = pk(add(xyz(wa4), mul(ia, -dw)), wa4.w);
      if (sbv) d.b.angvel[ab.y] = pk(add(xyz(wb4), mul(ib, dw)), wb4.w);
    }
  }
}

// rope_impulse (joints_ext.rs:173-223): once per step, after the torques,
// every stretched rope whose ends separate has the separating anchor
// velocity removed through the pair's effective mass.  Ropes sharing a body
// see each other's updates, so one lane walks them in joint order.
__global__ void k_joint_impulses(Dev d) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (uint32_t m = 0; m < d.nmotors; ++m) {  // all motors, then all ropes
    const uint32_t j = d.motors[m];
    if (d.J.kind[j] == JK_MOTOR)
      free_motor_impulse(d, j);
    else
      motor_impulse_one(d, j);
  }
  for (uint32_t r = 0; r < d.nropes; ++r) {
    const uint32_t j = d.ropes[r];
    const uint2 ab = d.J.ab[j];
    const float max = d.J.dref[j].x;
    const float4 pa4 = d.b.pos[ab.x], pb4 = d.b.pos[ab.y];
    const f3 pa = add(xyz(pa4), qrot(asq(d.b.rot[ab.x]), xyz(d.J.la[j])));
    const f3 pb = add(xyz(pb4), qrot(asq(d.b.rot[ab.y]), xyz(d.J.lb[j])));
    const f3 delta = sub(pb, pa);
    const float len = len3(delta);
    if (len <= max) continue;
    const f3 nrm = divs(delta, len);
    const f3 ra = sub(pa, xyz(pa4)), rb = sub(pb, xyz(pb4));
    const float ka = d.b.type[ab.x] == BT_DYNAMIC ? pa4.w : 0.0f;
    const float kb = d.b.type[ab.y] == BT_DYNAMIC ? pb4.w : 0.0f;
    const f3 ta = cross3(ra, nrm), tb = cross3(rb, nrm);
    const float k = ka + kb + dot3(ang_inv_wa_dev(d, ab.x, ta), ta) + dot3(ang_inv_wa_dev(d, ab.y, tb), tb);
    if (k < K_MIN_EFF_MASS) continue;
    const float4 va4 = d.b.vel[ab.x], vb4 = d.b.vel[ab.y], wa4 = d.b.angvel[ab.x], wb4 = d.b.angvel[ab.y];
    const f3 va = add(xyz(va4), cross3(xyz(wa4), ra));
    const f3 vb = add(xyz(vb4), cross3(xyz(wb4), rb));
    const float vsep = dot3(sub(vb, va), nrm);
    if (vsep <= 0.0f) continue;
    const float dj = -vsep / k;
    if (d.b.type[ab.x] == BT_DYNAMIC && pa4.w > 0.0f) {
      const f3 ia = ang_inv_wa_dev(d, ab.x, ta);
      d.b.vel[ab.x] = pk(sub(xyz(va4), mul(nrm, dj * ka)), va4.w);
      d.b.angvel[ab.x] = pk(sub(xyz(wa4), mul(ia, dj)), wa4.w);
    }
    if (d.b.type[ab.y] == BT_DYNAMIC && pb4.w > 0.0f) {
      const f3 ib = ang_inv_wa_dev(d, ab.y, tb);
      const float4 vb2 = d.b.vel[ab.y], wb2 = d.b.angvel[ab.y];  // body a's update above may be this body's
      d.b.vel[ab.y] = pk(add(xyz(vb2), mul(nrm, dj * kb)), vb2.w);
      d.b.angvel[ab.y] = pk(add(xyz(wb2), mul(ib, dj)), wb2.w);
    }
  }
}

// The TOI gate again after the rope catch changed velocities.
__global__ void __launch_bounds__(TPB) k_refast(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  const bool solvable = d.b.type[h] == BT_DYNAMIC && d.b.pos[h].w > 0.0f;
  d.fast[h] = solvable ? fast_mover(d, h) : 0;
}

// ---------------------------------------------------------------------------
// hash grid over finite, not-large bodies (acceleration structure only:
// every candidate it yields is re-tested exactly)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int cell_of(float x, float cell) { return (int)floorf(x / cell); }
// Cells are per scene: bodies of different scenes never pair, so a scene's
// cells only hold (up to hash collisions) its own bodies, however many
// scenes share the same space.  The key packs cell coordinates and scene,
// 16 bits each (cells 2^16 apart, or scenes 2^16 apart, share a key; a
// search never spans that far, and every candidate is re-tested exactly,
// scene included).
__device__ __forceinline__ unsigned long long cell_key(int x, int y, int z, uint32_t sc) {
  return ((unsigned long long)((uint32_t)x & 0xffffu)) | ((unsigned long long)((uint32_t)y & 0xffffu) << 16) |
         ((unsigned long long)((uint32_t)z & 0xffffu) << 32) | ((unsigned long long)(sc & 0xffffu) << 48);
}
__device__ __forceinline__ uint32_t bucket_of(int x, int y, int z, uint32_t sc, uint32_t nb) {
  uint32_t h = (uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^ (uint32_t)z * 83492791u ^ sc * 2654435761u;
  h ^= h >> 13;
  h *= 0x5bd1e995u;
  h ^= h >> 15;
  return h & (nb - 1u);
}

// `after_toi`: the rebuild after the TOI clamps, which runs only when
// there were casts.
__global__ void k_grid_clear(Dev d, uint32_t after_toi) {
  uint32_t i = gid();
  if (i >= d.g.nbuckets || (after_toi && dev_ncast(d) == 0u)) return;
  d.g.cnt[i] = 0u;
  d.g.cursor[i] = 0u;
}

__global__ void k_grid_count(Dev d, uint32_t after_toi) {
  uint32_t h = gid();
  if (h >= d.n || (after_toi && dev_ncast(d) == 0u)) return;
  if (d.g.is_large[h]) {
    d.g.bucket[h] = 0xffffffffu;
    return;
  }
  float4 p = d.b.pos[h];
  int cx = cell_of(p.x, d.g.cell), cy = cell_of(p.y, d.g.cell), cz = cell_of(p.z, d.g.cell);
  const uint32_t sc = d.b.scene[h];
  uint32_t bk = bucket_of(cx, cy, cz, sc, d.g.nbuckets);
  d.g.hkey[h] = cell_key(cx, cy, cz, sc);
  d.g.bucket[h] = bk;
  atomicAdd(&d.g.cnt[bk], 1u);
}

__global__ void k_grid_fill(Dev d, uint32_t after_toi) {
  uint32_t h = gid();
  if (h >= d.n || (after_toi && dev_ncast(d) == 0u)) return;
  uint32_t bk = d.g.bucket[h];
  if (bk == 0xffffffffu) return;
  uint32_t at = d.g.start[bk] + atomicAdd(&d.g.cursor[bk], 1u);
  d.g.items[at] = h;
  d.g.ipos[at] = pk(xyz(d.b.pos[h]), d.b.shape[h].w);
  d.g.ikey[at] = d.g.hkey[h];
}

// Distinct buckets of the cells in [lo, hi] (at most cap); returns count.
__device__ __forceinline__ int gather_buckets(const Dev& d, uint32_t sc, int lx, int ly, int lz, int hx, int hy,
                                              int hz, uint32_t* out, int cap) {
  int m = 0;
  for (int x = lx; x <= hx; ++x)
    for (int y = ly; y <= hy; ++y)
      for (int z = lz; z <= hz; ++z) {
        uint32_t bk = bucket_of(x, y, z, sc, d.g.nbuckets);
        bool dup = false;
        for (int i = 0; i < m; ++i)
          if (out[i] == bk) dup = true;
        if (!dup && m < cap) out[m++] = bk;
      }
  return m;
}

// ---------------------------------------------------------------------------
// TOI clamp (mod.rs:882-978; cast_shape, distance.rs:841-891).  Every cast
// reads the step-top poses; clamps are applied afterwards.
// ---------------------------------------------------------------------------
// Bodies joined to h by a joint of any kind (the TOI step excludes every
// joined partner, Fixed ones included: mod.rs:900-905), through h's
// incidence list.
__device__ __forceinline__ bool joined_dev(const Dev& d, uint32_t h, uint32_t o) {
  for (uint32_t k = d.b.ji_off[h]; k < d.b.ji_off[h + 1]; ++k) {
    const uint32_t e = d.b.ji[k];
    if (d.J.kind[e >> 1] == JK_GEAR) continue;  // mod.rs:903-905
    const uint2 ab = d.J.ab[e >> 1];
    if ((e & 1u) == 0u ? ab.y == o : ab.x == o) return true;
  }
  return false;
}
__device__ __forceinline__ bool toi_target_ok(const Dev& d, uint32_t h, uint32_t o) {
  if (o == h || d.b.trigger[o] || d.b.scene[o] != d.b.scene[h] || !pair_allowed_dev(d, h, o)) return false;
  return !joined_dev(d, h, o);
}

// Conservative advancement of mover against target (cast_shape,
// distance.rs:841-891); t < 0 if no hit.
//
// Box against box: the SAT gap is a lower bound of the distance (the
// separation along any unit axis is), so once the gap lowered by `slack`
// (a bound on the rounding of both computations, proportional to the
// coordinates' magnitude) already exceeds the touch distance and the
// remaining sweep, the loop would end without a hit whatever the exact
// distance: the expensive feature-distance pass is skipped and the outcome
// is unchanged.
template <bool BOXES = true>
HD float cast_one(const shp& mover, f3 dir, float len, const shp& target, f3* normal, f3* point) {
  const bool bb = BOXES && mover.kind == SH_BOX && target.kind == SH_BOX;
  float slack = 0.0f;
  if (bb) {
    float s = 1.0f + len + len3(mover.p) + len3(target.p);
    s += fmaxf(fmaxf(fabsf(mover.pos.x), fabsf(mover.pos.y)), fabsf(mover.pos.z));
    s += fmaxf(fmaxf(fabsf(target.pos.x), fabsf(target.pos.y)), fabsf(target.pos.z));
    slack = 4e-5f * s;
  }
  float t = 0.0f;
  for (int it = 0; it < 24; ++it) {
    shp m = mover;
    m.pos = add(mover.pos, mul(dir, t));
    dist_t r;
    if (bb) {
      f3 sn;
      float gap = separating_axis(m, m.p, target, target.p, &sn);
      float lo = gap - slack;
      if (lo > K_SHAPE_TOUCH && t + (lo - K_SHAPE_TOUCH * 0.5f) >= len) return -1.0f;
      r = box_box_from_sat(m, m.p, target, target.p, gap, sn);
    } else {
      float sat;
      r = shape_distance<BOXES>(m, target, &sat);
    }
    if (r.dist <= K_SHAPE_TOUCH) {
      if (t > 0.0f) {
        *normal = normalize_or(sub(r.pa, r.pb), neg(dir));
        *point = r.pb;
        return t;
      }
      return -1.0f;
    }
    t += r.dist - K_SHAPE_TOUCH * 0.5f;
    if (t >= len) break;
  }
  return -1.0f;
}

// Can the bounding spheres come within the touch distance along the sweep?
__device__ __forceinline__ bool sweep_may_touch(f3 p0, f3 dir, float len, float rm, f3 po, float ro) {
  f3 w = sub(po, p0);
  float s = dot3(w, dir);
  s = s < 0.0f ? 0.0f : (s > len ? len : s);
  f3 c = sub(w, mul(dir, s));
  float lim = (rm + ro + K_SHAPE_TOUCH) * 1.001f + 1e-4f;
  return len2(c) <= lim * lim * 1.0001f;
}

// Candidate targets of fast mover h: grid members near its swept bounding
// sphere, then the large bodies, after a bounding-sphere sweep test (a
// target outside it cannot come within the touch distance, so it cannot
// produce a hit) and the target filters.  Order is irrelevant: the hit is
// the minimum of (t, target index) over all targets.  emit(o) receives
// every target; returns the count.
template <typename Emit>
__device__ uint32_t toi_targets(const Dev& d, uint32_t h, Emit emit) {
  f3 delta = mul(xyz(d.b.vel[h]), K_DT);
  float len = len3(delta);
  if (len < K_NEAR_ZERO) return 0;
  f3 dir = divs(delta, len);
  f3 p0 = xyz(d.b.pos[h]);
  float rm = d.b.shape[h].w;
  uint32_t m = 0;
  auto consider = [&](uint32_t o, float4 po) {
    if (isfinite(po.w) && !sweep_may_touch(p0, dir, len, rm, xyz(po), po.w)) return;
    if (!toi_target_ok(d, h, o)) return;
    emit(m, o);
    ++m;
  };
  const uint32_t sch = d.b.scene[h];
  float r = rm + d.g.cell * 0.5f + 0.01f;
  f3 p1 = add(p0, delta);
  int lx = cell_of(fminf(p0.x, p1.x) - r, d.g.cell), hx = cell_of(fmaxf(p0.x, p1.x) + r, d.g.cell);
  int ly = cell_of(fminf(p0.y, p1.y) - r, d.g.cell), hy = cell_of(fmaxf(p0.y, p1.y) + r, d.g.cell);
  int lz = cell_of(fminf(p0.z, p1.z) - r, d.g.cell), hz = cell_of(fmaxf(p0.z, p1.z) + r, d.g.cell);
  long span = (long)(hx - lx + 1) * (long)(hy - ly + 1) * (long)(hz - lz + 1);
  if (span > 0 && span <= 64) {
    // each cell's members, found in its bucket by their cell key (buckets
    // shared by several visited cells are then scanned once per cell)
    for (int x = lx; x <= hx; ++x)
      for (int y = ly; y <= hy; ++y)
        for (int z = lz; z <= hz; ++z) {
          uint32_t bk = bucket_of(x, y, z, sch, d.g.nbuckets);
          unsigned long long key = cell_key(x, y, z, sch);
          uint32_t s0 = d.g.start[bk], c0 = d.g.cnt[bk];
          for (uint32_t k = 0; k < c0; ++k)
            if (d.g.ikey[s0 + k] == key) consider(d.g.items[s0 + k], d.g.ipos[s0 + k]);
        }
  } else {
    for (uint32_t k = d.g.soff[sch]; k < d.g.soff[sch + 1]; ++k) {  // every small body of the scene
      uint32_t o = d.g.sitems[k];
      if (!d.g.is_large[o]) consider(o, pk(xyz(d.b.pos[o]), d.b.shape[o].w));
    }
  }
  for (uint32_t k = d.g.lsc_off[sch]; k < d.g.lsc_off[sch + 1]; ++k) {
    uint32_t o = d.g.lsc[k];
    consider(o, pk(xyz(d.b.pos[o]), d.b.shape[o].w));
  }
  return m;
}

// Counting pass; the first TOI_CACHE targets are kept for the fill pass.
// Resets every body's clamp and counts the targets of fast bodies,
// BP_LANES lanes per body: lane s takes the swept cells s, s + BP_LANES, ...
// (or every BP_LANES-th small body of the scene when the sweep spans more
// than 64 cells) and every BP_LANES-th large body, the same targets as
// toi_targets.  The first TOI_CACHE targets found, in any order (every cast
// is independent), are kept for the fill pass.
__global__ void __launch_bounds__(TPB) k_toi_count(Dev d) {
  __shared__ uint32_t s_cnt[TPB / BP_LANES];
  const uint32_t h = gid() / BP_LANES;
  const uint32_t s = threadIdx.x % BP_LANES, lb = threadIdx.x / BP_LANES;
  if (s == 0) s_cnt[lb] = 0u;
  __syncthreads();
  if (h < d.n && s == 0) {
    d.toi_hit[h] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);  // w = travel to the hit; < 0: none
    d.toi_best[h] = ~0ull;
    if (d.ev_on) d.pre_vel[h] = d.b.vel[h];  // mod.rs:874-876: after the torque, before the clamp
  }
  bool moving = false;
  if (h < d.n && d.fast[h]) {
    const f3 delta = mul(xyz(d.b.vel[h]), K_DT);
    const float len = len3(delta);
    moving = !(len < K_NEAR_ZERO);
    if (moving) {
      uint32_t* cache = d.toi_cache + (size_t)h * TOI_CACHE;
      const f3 dir = divs(delta, len);
      const f3 p0 = xyz(d.b.pos[h]);
      const float rm = d.b.shape[h].w;
      auto consider = [&](uint32_t o, float4 po) {
        if (isfinite(po.w) && !sweep_may_touch(p0, dir, len, rm, xyz(po), po.w)) return;
        if (!toi_target_ok(d, h, o)) return;
        const uint32_t slot = atomicAdd(&s_cnt[lb], 1u);
        if (slot < TOI_CACHE) cache[slot] = o;
      };
      const uint32_t sch = d.b.scene[h];
      const float r = rm + d.g.cell * 0.5f + 0.01f;
      const f3 p1 = add(p0, delta);
      const int lx = cell_of(fminf(p0.x, p1.x) - r, d.g.cell), hx = cell_of(fmaxf(p0.x, p1.x) + r, d.g.cell);
      const int ly = cell_of(fminf(p0.y, p1.y) - r, d.g.cell), hy = cell_of(fmaxf(p0.y, p1.y) + r, d.g.cell);
      const int lz = cell_of(fminf(p0.z, p1.z) - r, d.g.cell), hz = cell_of(fmaxf(p0.z, p1.z) + r, d.g.cell);
      const long nx = hx - lx + 1, ny = hy - ly + 1, nz = hz - lz + 1;
      const long span = nx * ny * nz;
      if (span > 0 && span <= 64) {
        for (long ci = s; ci < span; ci += BP_LANES) {
          const int x = lx + (int)(ci / (ny * nz)), y = ly + (int)(ci / nz % ny), z = lz + (int)(ci % nz);
          const uint32_t bk = bucket_of(x, y, z, sch, d.g.nbuckets);
          const unsigned long long key = cell_key(x, y, z, sch);
          const uint32_t s0 = d.g.start[bk], c0 = d.g.cnt[bk];
          for (uint32_t k = 0; k < c0; ++k)
            if (d.g.ikey[s0 + k] == key) consider(d.g.items[s0 + k], d.g.ipos[s0 + k]);
        }
      } else {
        for (uint32_t k = d.g.soff[sch] + s; k < d.g.soff[sch + 1]; k += BP_LANES) {
          const uint32_t o = d.g.sitems[k];
          if (!d.g.is_large[o]) consider(o, pk(xyz(d.b.pos[o]), d.b.shape[o].w));
        }
      }
      for (uint32_t k = d.g.lsc_off[sch] + s; k < d.g.lsc_off[sch + 1]; k += BP_LANES) {
        const uint32_t o = d.g.lsc[k];
        consider(o, pk(xyz(d.b.pos[o]), d.b.shape[o].w));
      }
    }
  }
  __syncthreads();
  if (h < d.n && s == 0) d.toi_cnt[h] = moving ? s_cnt[lb] : 0u;
}

__global__ void k_toi_fill(Dev d) {
  uint32_t h = gid();
  const uint32_t total = d.toi_off[d.n];
  if (total > d.cap_toi) {  // recorded; the casts read as none
    if (h == 0) note_overflow(d, OVF_TOI, total);
    return;
  }
  if (h >= d.n || !d.fast[h]) return;
  uint2* out = d.toi_pairs + d.toi_off[h];
  uint32_t m = d.toi_cnt[h];
  if (m <= TOI_CACHE) {
    const uint32_t* cache = d.toi_cache + (size_t)h * TOI_CACHE;
    for (uint32_t i = 0; i < m; ++i) out[i] = make_uint2(h, cache[i]);
  } else {
    toi_targets(d, h, [&](uint32_t i, uint32_t o) { out[i] = make_uint2(h, o); });
  }
}

__device__ __forceinline__ unsigned long long toi_key(float t, uint32_t o) {
  return ((unsigned long long)__float_as_uint(t) << 32) | o;
}

// cast_one for a group of G lanes: box-box distances are evaluated by the
// whole group (same results on every lane); other shape pairs are cheap
// closed forms that every lane evaluates alike.
// Rounding allowance of the swept separating-axis bound of a box-box cast.
__device__ __forceinline__ float cast_slack(const shp& mover, float len, const shp& target) {
  float sc = 1.0f + len + len3(mover.p) + len3(target.p);
  sc += fmaxf(fmaxf(fabsf(mover.pos.x), fabsf(mover.pos.y)), fabsf(mover.pos.z));
  sc += fmaxf(fmaxf(fabsf(target.pos.x), fabsf(target.pos.y)), fabsf(target.pos.z));
  return 4e-5f * sc;
}

// The box-box cast's iterate at travel t: the mover translated to t, its
// separating-axis gap and normal.  Returns true when the cast misses from
// here: an axis keeps the rest of the sweep beyond the touch distance, or
// the gap cannot be closed within the travel left.
template <int G>
__device__ __forceinline__ bool cast_bound(const shp& mover, f3 dir, float len, const shp& target, float t,
                                           float slack, int s, shp* m, f3* sn, float* gap) {
  *m = mover;
  m->pos = add(mover.pos, mul(dir, t));
  bool clear;
  *gap = grp::sat<G>(*m, m->p, target, target.p, sn, s, &clear, dir, len - t, K_SHAPE_TOUCH, slack);
  if (clear) return true;  // no iterate can come within the touch distance
  float lo = *gap - slack;
  return lo > K_SHAPE_TOUCH && t + (lo - K_SHAPE_TOUCH * 0.5f) >= len;
}

template <int G>
__device__ float cast_group(const shp& mover, f3 dir, float len, const shp& target, f3* normal, f3* point, int s) {
  if (!(mover.kind == SH_BOX && target.kind == SH_BOX)) return cast_one<false>(mover, dir, len, target, normal, point);
  const float slack = cast_slack(mover, len, target);
  float t = 0.0f;
  for (int it = 0; it < 24; ++it) {
    shp m;
    f3 sn;
    float gap;
    if (cast_bound<G>(mover, dir, len, target, t, slack, s, &m, &sn, &gap)) return -1.0f;
    dist_t r = grp::box_box<G>(m, m.p, target, target.p, gap, sn, s);
    if (r.dist <= K_SHAPE_TOUCH) {
      if (t > 0.0f) {
        *normal = normalize_or(sub(r.pa, r.pb), neg(dir));
        *point = r.pb;
        return t;
      }
      return -1.0f;
    }
    t += r.dist - K_SHAPE_TOUCH * 0.5f;
    if (t >= len) break;
  }
  return -1.0f;
}

// A cast's outcome: its (normal, t) is kept and the mover keeps the
// smallest (t, target) key: t > 0, so its IEEE bits order like the value.
__device__ __forceinline__ void toi_result(const Dev& d, uint32_t i, uint2 ho, f3 nrm, f3 pt, float t) {
  d.toi_res[i] = pk(nrm, t);
  d.toi_pt[i] = pk(pt, 0.0f);
  if (t < 0.0f) return;
  atomicMin(&d.toi_best[ho.x], toi_key(t, ho.y));
}

// First look at every cast, G lanes per cast.  Casts that are not box-box
// are cheap closed forms and finish here; a box-box cast finishes here when
// cast_group's first iterate (travel 0, the same arithmetic) already misses
// — most casts, since movers mostly pass their neighbours.  The rest are
// listed for k_toi_pair.
template <int G>
__global__ void __launch_bounds__(TPB) k_toi_gate(Dev d) {
  __shared__ uint32_t s_n, s_base;
  const uint32_t npairs = dev_ncast(d);
  if (threadIdx.x == 0) s_n = 0u;
  __syncthreads();
  const uint32_t i = gid() / G;
  const int s = (int)(threadIdx.x % G);
  bool live = false;
  if (i < npairs) {
    const uint2 ho = d.toi_pairs[i];
    const f3 delta = mul(xyz(d.b.vel[ho.x]), K_DT);
    const float len = len3(delta);
    const f3 dir = divs(delta, len);
    const shp mover = body_shape(d, ho.x), target = body_shape(d, ho.y);
    f3 nrm = zero3(), pt = zero3();
    float t = -1.0f;
    if (!(mover.kind == SH_BOX && target.kind == SH_BOX)) {
      t = cast_one<false>(mover, dir, len, target, &nrm, &pt);
    } else {
      shp m;
      f3 sn;
      float gap;
      live = !cast_bound<G>(mover, dir, len, target, 0.0f, cast_slack(mover, len, target), s, &m, &sn, &gap);
    }
    if (s == 0 && !live) toi_result(d, i, ho, nrm, pt, t);
  }
  const uint32_t slot = (live && s == 0) ? atomicAdd(&s_n, 1u) : 0u;
  __syncthreads();
  if (threadIdx.x == 0 && s_n) s_base = atomicAdd(&d.counters[CTR_TOI_LIVE], s_n);
  __syncthreads();
  if (live && s == 0) d.toi_live[s_base + slot] = i;
}

// The listed box-box casts in full, one per group of G lanes, over a
// grid-stride loop (the list length is only known on the device).
template <int G>
__global__ void __launch_bounds__(TPB) k_toi_pair(Dev d) {
  const uint32_t live = d.counters[CTR_TOI_LIVE];
  const uint32_t groups = gridDim.x * (TPB / G);
  const int s = (int)(threadIdx.x % G);
  for (uint32_t q = gid() / G; q < live; q += groups) {
    const uint32_t i = d.toi_live[q];
    const uint2 ho = d.toi_pairs[i];
    const f3 delta = mul(xyz(d.b.vel[ho.x]), K_DT);
    const float len = len3(delta);
    const f3 dir = divs(delta, len);
    f3 nrm = zero3(), pt = zero3();
    const float t = cast_group<G>(body_shape(d, ho.x), dir, len, body_shape(d, ho.y), &nrm, &pt, s);
    if (s == 0) toi_result(d, i, ho, nrm, pt, t);
  }
}

// The mover's hit is the cast holding the winning key (a repeated
// (mover, target) entry carries identical values).
__global__ void __launch_bounds__(TPB) k_toi_win(Dev d) {
  uint32_t i = gid();
  if (i >= dev_ncast(d)) return;
  float4 r = d.toi_res[i];
  if (!(r.w >= 0.0f)) return;
  uint2 ho = d.toi_pairs[i];
  if (d.toi_best[ho.x] != toi_key(r.w, ho.y)) return;
  d.toi_hit[ho.x] = r;
  d.toi_tgt[ho.x] = ho.y;
  d.toi_hpt[ho.x] = d.toi_pt[i];
}

__device__ __forceinline__ bool toi_apply_body(const Dev& d, uint32_t h) {
  float4 hit = d.toi_hit[h];
  if (!(hit.w >= 0.0f) || !d.fast[h]) return false;
  uint32_t o = d.toi_tgt[h];
  f3 nrm = xyz(hit);
  float4 v4 = d.b.vel[h];
  f3 vel = xyz(v4);
  f3 disp = mul(vel, K_DT);
  float len = fmax_(len3(disp), K_MIN_TOI_LEN);
  float frac = clampf_(hit.w / len, 0.0f, 1.0f);
  float e = fmin_(d.b.mat[h].y, d.b.mat[o].y);
  float4 p = d.b.pos[h];
  d.b.pos[h] = pk(add(xyz(p), mul(disp, frac)), p.w);
  float vn = dot3(vel, nrm);
  if (vn < 0.0f) {
    float bounce = vn < -K_CCD_BOUNCE ? 1.0f + e : 1.0f;
    d.b.vel[h] = pk(sub(vel, mul(nrm, bounce * vn)), v4.w);
  }
  d.b.timer[h] = 0.0f;
  d.b.flag_a[o] = 1;  // wake the target
  return true;
}
// The clamp's impact event of mover h (mod.rs:947-960): the closing speed
// from the velocities before the clamp; (a, b) in handle order, the normal
// as the clamp found it when the mover is b, negated when it is a.
__device__ __forceinline__ void toi_impact_event(const Dev& d, uint32_t h) {
  uint32_t flag = 0u;
  const float4 hit = d.toi_hit[h];
  if (hit.w >= 0.0f && d.fast[h]) {
    const uint32_t o = d.toi_tgt[h];
    const f3 nrm = xyz(hit);
    const float approach = -dot3(sub(xyz(d.pre_vel[h]), xyz(d.pre_vel[o])), nrm);
    if (approach > 1.0f) {  // CONTACT_HIT_THRESHOLD (trigger.rs:63)
      const bool lo = h < o;
      const f3 n = lo ? neg(nrm) : nrm;
      const float4 pt = d.toi_hpt[h];
      d.ev_crec[3 * (size_t)h] = make_uint4(EVK_HIT, lo ? h : o, lo ? o : h, 0u);
      d.ev_crec[3 * (size_t)h + 1] =
          make_uint4(__float_as_uint(pt.x), __float_as_uint(pt.y), __float_as_uint(pt.z), __float_as_uint(approach));
      d.ev_crec[3 * (size_t)h + 2] = make_uint4(__float_as_uint(n.x), __float_as_uint(n.y), __float_as_uint(n.z), 0u);
      flag = 1u;
    }
  }
  d.ev_cflag[h] = flag;
}
__global__ void __launch_bounds__(TPB) k_toi_apply(Dev d) {
  if (dev_ncast(d) == 0u) return;  // whole launch
  uint32_t h = gid();
  bool hit = false;
  if (h < d.n && d.ev_on) toi_impact_event(d, h);  // before the clamp changes the velocity
  if (h < d.n) hit = toi_apply_body(d, h);
  block_add(&d.counters[CTR_TOIHITS], hit ? 1u : 0u);
}

// Apply pending wakes (flag_a) and clear them.
__device__ __forceinline__ void apply_wake_body(const Dev& d, uint32_t h) {
  if (d.b.flag_a[h]) {
    d.b.flag_a[h] = 0;
    wake_body_dev(d, h);
  }
}
__global__ void k_apply_wake(Dev d, uint32_t after_toi) {
  uint32_t h = gid();
  if (h >= d.n || (after_toi && dev_ncast(d) == 0u)) return;
  apply_wake_body(d, h);
}

// ---------------------------------------------------------------------------
// broadphase: candidates (a < b) passing the reference prefilters
// (rows.rs:432-479): scene, layer/mask, pin-joint no-collide, bounding
// spheres within GEN_MARGIN.  Output is grouped by a and sorted by b.
// ---------------------------------------------------------------------------
// The non-geometric prefilters (scene, layer/mask, pin-joint no-collide).
__device__ __forceinline__ bool prefilter_ids(const Dev& d, uint32_t a, uint32_t b) {
  if (d.b.scene[a] != d.b.scene[b]) return false;
  if (!pair_allowed_dev(d, a, b)) return false;
  return !(!d.b.trigger[a] && !d.b.trigger[b] && d.b.jp_off[a] != d.b.jp_off[a + 1] && jointed_dev(d, a, b));
}
// Bounding spheres within GEN_MARGIN: |pa - pb| <= ra + rb + GEN_MARGIN as
// the reference evaluates it.  Squared distances clearly inside or outside
// decide without the square root; the narrow band around the limit takes
// the exact form.
__device__ __forceinline__ bool spheres_near(f3 pa, float ra, f3 pb, float rb) {
  const f3 dv = sub(pa, pb);
  const float r = ra + rb + K_GEN_MARGIN;
  const float d2 = dot3(dv, dv), r2 = r * r;
  if (d2 > r2 * 1.0001f) return false;
  if (d2 < r2 * 0.9999f) return true;
  return !(sqrtf(d2) > r);
}
__device__ __forceinline__ bool prefilter(const Dev& d, uint32_t a, uint32_t b) {
  return spheres_near(xyz(d.b.pos[a]), d.b.shape[a].w, xyz(d.b.pos[b]), d.b.shape[b].w) && prefilter_ids(d, a, b);
}

// Candidates (a, o > a) of small body a: grid members of the 27 cells
// around it (sphere test on the packed item data first), then the large
// bodies.  emit(i, o) receives the i-th candidate; returns the count.
template <typename Emit>
__device__ __forceinline__ uint32_t bp_small(const Dev& d, uint32_t a, Emit emit) {
  float4 p = d.b.pos[a];
  const f3 pa = xyz(p);
  const float ra = d.b.shape[a].w;
  int cx = cell_of(p.x, d.g.cell), cy = cell_of(p.y, d.g.cell), cz = cell_of(p.z, d.g.cell);
  uint32_t bks[27];
  int nbk = gather_buckets(d, d.b.scene[a], cx - 1, cy - 1, cz - 1, cx + 1, cy + 1, cz + 1, bks, 27);
  uint32_t m = 0;
  for (int i = 0; i < nbk; ++i) {
    uint32_t s0 = d.g.start[bks[i]], c0 = d.g.cnt[bks[i]];
    for (uint32_t k = 0; k < c0; ++k) {
      float4 q = d.g.ipos[s0 + k];
      if (!spheres_near(pa, ra, xyz(q), q.w)) continue;
      uint32_t o = d.g.items[s0 + k];
      if (o > a && prefilter_ids(d, a, o)) emit(m++, o);
    }
  }
  const uint32_t sc = d.b.scene[a];
  for (uint32_t k = d.g.lsc_off[sc]; k < d.g.lsc_off[sc + 1]; ++k) {
    uint32_t o = d.g.lsc[k];
    if (o > a && prefilter(d, a, o)) emit(m++, o);
  }
  return m;
}

// Counting pass, BP_LANES lanes per body: lane s scans the body's cells s,
// s + BP_LANES, ... (a bucket shared with an earlier cell of the 27 is
// skipped, as bp_small skips it) and every BP_LANES-th large body.  The first
// BP_CACHE candidates found, in any order (the fill pass sorts them), are
// kept for the fill pass; slots come from a per-body LDS counter.
__global__ void __launch_bounds__(TPB) k_bp_count(Dev d, uint32_t* cnt) {
  __shared__ uint32_t s_cnt[TPB / BP_LANES];
  const uint32_t a = gid() / BP_LANES;
  const uint32_t s = threadIdx.x % BP_LANES, lb = threadIdx.x / BP_LANES;
  if (s == 0) s_cnt[lb] = 0u;
  __syncthreads();
  if (a < d.n && !d.g.is_large[a]) {
    uint32_t* cache = d.bp_cache + (size_t)a * BP_CACHE;
    auto emit = [&](uint32_t o) {
      const uint32_t slot = atomicAdd(&s_cnt[lb], 1u);
      if (slot < BP_CACHE) cache[slot] = o;
    };
    const float4 p = d.b.pos[a];
    const f3 pa = xyz(p);
    const float ra = d.b.shape[a].w;
    const uint32_t sc = d.b.scene[a];
    const int cx = cell_of(p.x, d.g.cell), cy = cell_of(p.y, d.g.cell), cz = cell_of(p.z, d.g.cell);
    // lane s holds the buckets of cells s + BP_LANES m; the group's lanes
    // exchange them once to find the cells whose bucket an earlier cell of
    // the 27 already scans
    constexpr int NS = (27 + BP_LANES - 1) / BP_LANES;
    uint32_t bkm[NS];
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const uint32_t ci = s + BP_LANES * m;
      bkm[m] = ci < 27u ? bucket_of(cx - 1 + (int)(ci / 9), cy - 1 + (int)(ci / 3 % 3), cz - 1 + (int)(ci % 3), sc,
                                    d.g.nbuckets)
                        : 0xffffffffu;
    }
    bool dupm[NS];
#pragma unroll
    for (int m = 0; m < NS; ++m) dupm[m] = false;
#pragma unroll
    for (int t = 0; t < BP_LANES; ++t)
#pragma unroll
      for (int m2 = 0; m2 < NS; ++m2) {
        const uint32_t o = __shfl(bkm[m2], t, BP_LANES);  // bucket of cell t + BP_LANES m2
        const uint32_t cj = (uint32_t)t + BP_LANES * m2;
#pragma unroll
        for (int m = 0; m < NS; ++m) dupm[m] = dupm[m] || (cj < s + BP_LANES * m && o == bkm[m]);
      }
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const uint32_t ci = s + BP_LANES * m;
      if (ci >= 27u || dupm[m]) continue;
      const uint32_t bk = bkm[m];
      const uint32_t s0 = d.g.start[bk], c0 = d.g.cnt[bk];
      for (uint32_t k = 0; k < c0; ++k) {
        const float4 q = d.g.ipos[s0 + k];
        if (!spheres_near(pa, ra, xyz(q), q.w)) continue;
        const uint32_t o = d.g.items[s0 + k];
        if (o > a && prefilter_ids(d, a, o)) emit(o);
      }
    }
    for (uint32_t k = d.g.lsc_off[sc] + s; k < d.g.lsc_off[sc + 1]; k += BP_LANES) {
      const uint32_t o = d.g.lsc[k];
      if (o > a && prefilter(d, a, o)) emit(o);
    }
  }
  __syncthreads();
  if (a < d.n && s == 0) cnt[a] = d.g.is_large[a] ? 0u : s_cnt[lb];
}

// Large bodies pair with every later body of their scene passing the
// prefilter.  Flag row k covers large body k's scene (bodies in index
// order); t indexes all rows back to back.
__device__ __forceinline__ uint32_t large_row(const Dev& d, uint32_t t) {
  uint32_t lo = 0, hi = d.g.nlarge;  // last row whose base is <= t
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (d.g.lrange[mid].z <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}
__global__ void k_bp_large_flag(Dev d) {
  uint32_t t = gid();
  if (t >= d.g.lflat) return;
  uint4 r = d.g.lrange[large_row(d, t)];
  uint32_t b = d.g.sitems[r.x + (t - r.z)];
  d.lflag[t] = (r.w < b && prefilter(d, r.w, b)) ? 1u : 0u;
}

__global__ void k_bp_large_count(Dev d, uint32_t* cnt) {
  uint32_t k = gid();
  if (k >= d.g.nlarge) return;
  uint4 r = d.g.lrange[k];
  cnt[r.w] = d.lrank[r.z + (r.y - r.x)] - d.lrank[r.z];
}

// Candidates of each small body, sorted by b: lists of up to BP_CACHE are
// sorted in a per-thread LDS slice, longer ones in place.
__global__ void __launch_bounds__(TPB) k_bp_fill(Dev d) {
  __shared__ uint32_t srt[BP_CACHE][TPB];  // column per thread: no bank conflicts between lanes
  uint32_t a = gid();
  const uint32_t total = d.cand_off[d.n];
  if (total > d.cap_cand) {  // recorded; the candidates read as none
    if (a == 0) note_overflow(d, OVF_CAND, total);
    return;
  }
  if (a >= d.n || d.g.is_large[a]) return;
  uint2* out = d.C.ab + d.cand_off[a];
  const uint32_t m = d.cand_off[a + 1] - d.cand_off[a];
  const uint32_t t = threadIdx.x;
  if (m <= BP_CACHE) {
    const uint32_t* cache = d.bp_cache + (size_t)a * BP_CACHE;
    for (uint32_t i = 0; i < m; ++i) {
      uint32_t v = cache[i];
      uint32_t j = i;
      while (j > 0 && srt[j - 1][t] > v) {
        srt[j][t] = srt[j - 1][t];
        --j;
      }
      srt[j][t] = v;
    }
    for (uint32_t i = 0; i < m; ++i) out[i] = make_uint2(a, srt[i][t]);
    return;
  }
  bp_small(d, a, [&](uint32_t i, uint32_t o) { out[i] = make_uint2(a, o); });
  for (uint32_t i = 1; i < m; ++i) {  // insertion sort by b
    uint2 v = out[i];
    uint32_t j = i;
    while (j > 0 && out[j - 1].y > v.y) {
      out[j] = out[j - 1];
      --j;
    }
    out[j] = v;
  }
}

__global__ void k_bp_large_fill(Dev d) {
  uint32_t t = gid();
  if (t >= d.g.lflat || !d.lflag[t] || d.cand_off[d.n] > d.cap_cand) return;
  uint4 r = d.g.lrange[large_row(d, t)];
  uint32_t b = d.g.sitems[r.x + (t - r.z)];
  d.C.ab[d.cand_off[r.w] + d.lrank[t] - d.lrank[r.z]] = make_uint2(r.w, b);
}

// ---------------------------------------------------------------------------
// narrowphase: discover_pair (rows.rs:432-756) for every candidate
// ---------------------------------------------------------------------------
// Index of pair (a, b) in the previous step's table (-1 if absent): a
// binary search for b in a's run of that table.
__device__ __forceinline__ int find_old_pair(const Dev& d, uint32_t a, uint32_t b) {
  uint32_t lo = d.poff_old[a], hi = d.poff_old[a + 1];
  const uint32_t end = hi;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (d.Pold.ab[mid].y < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < end && d.Pold.ab[lo].y == b) ? (int)lo : -1;
}

// First pair of every body's a-run: pairs are grouped by their first body
// in body order, so the run of h starts where h's accepted candidates do.
// Every thread also derives this step's pair and point totals (the
// acceptance scan's); when either exceeds its table the overflow is
// recorded and the step has no pairs (all a-runs empty).
__global__ void k_pair_offsets(Dev d) {
  uint32_t h = gid();
  const uint32_t nc = dev_ncand(d);
  const uint32_t np = d.acc_off[nc], nq = d.pt_off[nc];
  const bool fits = np <= d.cap_pairs && nq <= d.cap_pts;
  if (h <= d.n) d.poff[h] = (fits && nc) ? d.acc_off[d.cand_off[h]] : 0u;
  if (h == 0) {
    d.dc[DC_NPAIRS] = fits ? np : 0u;
    d.dc[DC_NPTS] = fits ? nq : 0u;
    if (np > d.cap_pairs) note_overflow(d, OVF_PAIRS, np);
    if (nq > d.cap_pts) note_overflow(d, OVF_PTS, nq);
  }
}

// Narrowphase gate: filters, the old pair, and for two boxes the SAT gap
// with the creation gate.  Survivors are listed (one atomic per workgroup
// and list) for k_narrow with their SAT result stashed in disc0 (normal,
// gap) and the old pair in disc1.z, so the expensive part runs on full
// wavefronts: overlapping box pairs (whose feature scans are the longest)
// from the back of the candidate range, the rest from the front, so that
// wavefronts hold work of one kind.
__global__ void __launch_bounds__(TPB) k_narrow_gate(Dev d) {
  __shared__ uint32_t s_n, s_base, s_p, s_pbase;  // front and back lists
  const uint32_t ncand = dev_ncand(d);
  if (threadIdx.x == 0) {
    s_n = 0u;
    s_p = 0u;
  }
  __syncthreads();
  const uint32_t c = gid();
  bool live = false, overlap = false;
  if (c < ncand) {
    d.C.acc[c] = 0u;
    d.C.npts[c] = 0u;
    const uint2 ab = d.C.ab[c];
    const uint32_t ia = ab.x, ib = ab.y;
    if (d.ev_emit) {  // trigger pairs overlap at a signed gap <= 0 (rows.rs:496-523, :803-807)
      uint32_t g = 0u;
      if (d.b.trigger[ia] || d.b.trigger[ib]) {
        const shp A = body_shape(d, ia), B = body_shape(d, ib);
        float sgn;
        if (A.kind == SH_BOX && B.kind == SH_BOX) {
          f3 sn;
          sgn = separating_axis(A, A.p, B, B.p, &sn);
        } else {
          float sat;
          sgn = shape_distance<false>(A, B, &sat).dist;
        }
        g = sgn <= 0.0f ? 1u : 0u;
      }
      d.ev_gflag[c] = g;
    }
    if (!d.b.trigger[ia] && !d.b.trigger[ib]) {  // trigger overlaps form no pair
      const int old = find_old_pair(d, ia, ib);
      live = true;
      if (d.b.kind[ia] == SH_BOX && d.b.kind[ib] == SH_BOX) {
        const shp A = body_shape(d, ia), B = body_shape(d, ib);
        f3 sn;
        const float sat = separating_axis(A, A.p, B, B.p, &sn);
        if (old < 0) {
          // Creation gate bound: reach <= GEN_MARGIN + |vb - va| dt (|normal| = 1).
          const f3 va = xyz(d.b.vel[ia]), vb = xyz(d.b.vel[ib]);
          const float reach_hi = K_GEN_MARGIN + len3(sub(vb, va)) * K_DT * 1.0001f + 1e-6f;
          live = !(sat > reach_hi);
        }
        overlap = !(sat > 0.0f);
        d.C.disc0[c] = pk(sn, sat);
      }
      d.C.disc1[c] = make_float4(0.0f, 0.0f, __int_as_float(old), 0.0f);
    }
  }
  const uint32_t slot = live ? atomicAdd(overlap ? &s_p : &s_n, 1u) : 0u;
  __syncthreads();
  if (threadIdx.x == 0 && (s_n || s_p)) {
    // both lists' bases from one atomic on the shared word (no carry: each
    // count stays below the candidate count)
    const unsigned long long b = atomicAdd(reinterpret_cast<unsigned long long*>(d.counters + CTR_NARROW),
                                           ((unsigned long long)s_p << 32) | s_n);
    s_base = (uint32_t)b;
    s_pbase = (uint32_t)(b >> 32);
  }
  __syncthreads();
  if (live) d.nlist[overlap ? ncand - 1u - (s_pbase + slot) : s_base + slot] = c;
}

// Entry t of the gate's list: the first counters[CTR_NARROW] entries of
// nlist and the last counters[CTR_NARROW_PEN]; false between them.
__device__ __forceinline__ bool narrow_entry(const Dev& d, uint32_t t, uint32_t* c) {
  const uint32_t ncand = dev_ncand(d);
  if (t >= ncand) return false;
  if (t >= d.counters[CTR_NARROW] && t < ncand - d.counters[CTR_NARROW_PEN]) return false;
  *c = d.nlist[t];
  return true;
}

// Closest features of the listed box-box candidates, G lanes per candidate
// (box_box_from_sat with the gate's SAT result and the pruned group scan),
// stashed in the candidate's first two point slots for k_narrow, which
// reads them before it writes its points.
template <int G>
__global__ void __launch_bounds__(TPB) k_narrow_dist(Dev d) {
  const uint32_t t = gid() / G;
  const int s = (int)(threadIdx.x % G);
  uint32_t c;
  if (!narrow_entry(d, t, &c)) return;  // whole groups
  const uint2 ab = d.C.ab[c];
  if (!(d.b.kind[ab.x] == SH_BOX && d.b.kind[ab.y] == SH_BOX)) return;
  const shp A = body_shape(d, ab.x), B = body_shape(d, ab.y);
  const float4 g = d.C.disc0[c];
  const dist_t r = grp::box_box<G>(A, A.p, B, B.p, g.w, xyz(g), s);
  if (s != 0) return;
  d.C.fresh[(size_t)c * 16] = make_float4(r.dist, r.pa.x, r.pa.y, r.pa.z);
  d.C.fresh[(size_t)c * 16 + 1] = make_float4(r.pb.x, r.pb.y, r.pb.z, 0.0f);
}

// Narrowphase of the gate's survivors (thread t takes list entry t; threads
// off the list return at once).
__global__ void __launch_bounds__(TPB) k_narrow(Dev d) {
  // the manifold points (local offsets on a and on b) of each thread, one
  // LDS column per thread: the point loops index them at run time
  __shared__ float s_pts[K_MAX_POINTS * 6][TPB];
  const uint32_t t = gid();
  uint32_t c;
  if (!narrow_entry(d, t, &c)) return;
  uint2 ab = d.C.ab[c];
  uint32_t ia = ab.x, ib = ab.y;
  shp A = body_shape(d, ia), B = body_shape(d, ib);
  const int old = __float_as_int(d.C.disc1[c].z);
  f3 va = xyz(d.b.vel[ia]), vb = xyz(d.b.vel[ib]);
  dist_t dd;
  float sgn;
  if (A.kind == SH_BOX && B.kind == SH_BOX) {
    const float4 r0 = d.C.fresh[(size_t)c * 16], r1 = d.C.fresh[(size_t)c * 16 + 1];
    dd = dist_t{r0.x, f3{r0.y, r0.z, r0.w}, f3{r1.x, r1.y, r1.z}};
    sgn = d.C.disc0[c].w;
  } else {
    float sat;
    dd = shape_distance<false>(A, B, &sat);  // never box-box here
    sgn = dd.dist;
  }
  // orient_contact_normal (rows.rs:303-333)
  f3 normal;
  if (A.kind == SH_HALFSPACE) {
    normal = neg(qrot(A.rot, A.p));
  } else if (B.kind == SH_HALFSPACE) {
    normal = qrot(B.rot, B.p);
  } else {
    normal = sub(dd.pa, dd.pb);
    if (dd.dist < 0.0f) normal = neg(normal);
    if (len2(normal) < K_AXIS_REST_LEN2) normal = sub(A.pos, B.pos);
    normal = normalize_or(normal, f3{0.0f, 1.0f, 0.0f});
    if (old >= 0) {
      if (dot3(normal, xyz(d.Pold.n(old))) < 0.0f) normal = neg(normal);
    } else if (dot3(normal, sub(A.pos, B.pos)) < 0.0f) {
      normal = neg(normal);
    }
  }
  float approach = fmax_(dot3(sub(vb, va), normal), 0.0f);
  float reach = K_GEN_MARGIN + approach * K_DT;
  if (sgn > reach && old < 0) return;
  bool separated = sgn > reach;
  float4 ma = d.b.mat[ia], mb = d.b.mat[ib];
  // friction_frame's coefficients (rows.rs:32-52, :567-570): along the
  // axis and across it for a body with an axis, the friction both ways else
  const float4 fa = d.b.fdir[ia], fb = d.b.fdir[ib];
  const float a2 = has_axis(fa) ? fa.w : ma.x, b2 = has_axis(fb) ? fb.w : mb.x;
  float mu1 = fmax_(ma.x, mb.x);
  mu1 = fmax_(mu1, 0.0f);
  float mu2 = fmax_(a2, b2);
  mu2 = fmax_(mu2, 0.0f);
  float ra = d.b.shape[ia].w, rb = d.b.shape[ib].w;
  float patch = fmin_(ra, rb) + K_MARGIN;
  f3 pp = mul(add(dd.pa, dd.pb), 0.5f);
  f3 pa_c = sub(A.pos, mul(normal, dot3(sub(A.pos, pp), normal)));
  f3 pb_c = sub(B.pos, mul(normal, dot3(sub(B.pos, pp), normal)));
  float plane_a = A.kind == SH_BOX ? dot3(normal, A.pos) - 1.0f * dot3(abs3(qrot(qconj(A.rot), normal)), A.p)
                                   : dot3(normal, dd.pa);
  float plane_b = B.kind == SH_BOX ? dot3(normal, B.pos) - (-1.0f) * dot3(abs3(qrot(qconj(B.rot), normal)), B.p)
                                   : dot3(normal, dd.pb);
  qf inv_a = qconj(A.rot), inv_b = qconj(B.rot);
  const int tx = (int)threadIdx.x;
  auto fra = [&](int i) { return f3{s_pts[6 * i][tx], s_pts[6 * i + 1][tx], s_pts[6 * i + 2][tx]}; };
  auto frb = [&](int i) { return f3{s_pts[6 * i + 3][tx], s_pts[6 * i + 4][tx], s_pts[6 * i + 5][tx]}; };
  int nf = 0;
  auto dup = [&](f3 x, f3 y) {
    for (int i = 0; i < nf; ++i)
      if (len3(sub(fra(i), x)) < K_POINT_MATCH && len3(sub(frb(i), y)) < K_POINT_MATCH) return true;
    return false;
  };
  auto push = [&](f3 x, f3 y) {
    s_pts[6 * nf][tx] = x.x;
    s_pts[6 * nf + 1][tx] = x.y;
    s_pts[6 * nf + 2][tx] = x.z;
    s_pts[6 * nf + 3][tx] = y.x;
    s_pts[6 * nf + 4][tx] = y.y;
    s_pts[6 * nf + 5][tx] = y.z;
    ++nf;
  };
  if (!separated) {
    for (int side = 0; side < 2; ++side) {
      const shp& body = side == 0 ? A : B;
      if (body.kind != SH_BOX) continue;
      f3 witness = side == 0 ? dd.pa : dd.pb;
      float sign = side == 0 ? 1.0f : -1.0f;
      f3 pc = side == 0 ? pa_c : pb_c;
      f3 sn = smul(sign, normal);
      float slop = K_EXPAND_SLOP + fmax_(-sgn, 0.0f);
      for (int ci = 0; ci < 8 && nf < K_MAX_POINTS; ++ci) {
        float sx = (ci & 4) ? 1.0f : -1.0f, sy = (ci & 2) ? 1.0f : -1.0f, sz = (ci & 1) ? 1.0f : -1.0f;
        f3 corner{sx * body.p.x, sy * body.p.y, sz * body.p.z};
        f3 world = add(body.pos, qrot(body.rot, corner));
        float along = dot3(sub(world, witness), sn);
        if (fabsf(along) > slop) continue;
        f3 rel = sub(world, pc);
        f3 tang = sub(rel, mul(normal, dot3(rel, normal)));
        if (len3(tang) > patch) continue;
        f3 pa, pb;
        if (side == 0) {
          pa = world;
          pb = add(world, mul(normal, plane_b - dot3(normal, world)));
        } else {
          pa = add(world, mul(normal, plane_a - dot3(normal, world)));
          pb = world;
        }
        f3 ral = qrot(inv_a, sub(pa, A.pos));
        f3 rbl = qrot(inv_b, sub(pb, B.pos));
        if (dup(ral, rbl)) continue;
        push(ral, rbl);
      }
    }
    if (nf >= 2 && nf < K_MAX_POINTS) {
      f3 mid = ra < rb ? pa_c : (rb < ra ? pb_c : mul(add(pa_c, pb_c), 0.5f));
      f3 pa = add(mid, mul(normal, plane_a - dot3(normal, mid)));
      f3 pb = add(mid, mul(normal, plane_b - dot3(normal, mid)));
      f3 ral = qrot(inv_a, sub(pa, A.pos));
      f3 rbl = qrot(inv_b, sub(pb, B.pos));
      if (!dup(ral, rbl)) push(ral, rbl);
    }
    if (nf < 3) {
      f3 ral = qrot(inv_a, sub(dd.pa, A.pos));
      f3 rbl = qrot(inv_b, sub(dd.pb, B.pos));
      if (!dup(ral, rbl)) push(ral, rbl);
    }
  }
  bool wake;
  bool sa = d.b.asleep[ia] != 0, sb = d.b.asleep[ib] != 0;
  if (old >= 0) {
    wake = (sa || sb) && sgn < -K_WAKE_PEN;
  } else {
    wake = sgn < -K_WAKE_PEN;
    if (!wake && sa) {
      if (dot3(sub(vb, va), smul(1.0f, normal)) > K_WAKE_IMPACT) wake = true;
    }
    if (!wake && sb) {
      if (dot3(sub(va, vb), smul(-1.0f, normal)) > K_WAKE_IMPACT) wake = true;
    }
  }
  d.C.acc[c] = 1u;
  d.C.npts[c] = (uint32_t)nf;
  d.C.disc0[c] = pk(normal, sgn);
  d.C.disc1[c] = make_float4(mu1, mu2, __int_as_float(old), wake ? 1.0f : 0.0f);
  for (int k = 0; k < nf; ++k) {
    const f3 x = fra(k), y = frb(k);
    d.C.fresh[(size_t)c * 16 + 2 * k] = make_float4(x.x, x.y, x.z, y.x);
    d.C.fresh[(size_t)c * 16 + 2 * k + 1] = make_float4(y.y, y.z, 0.0f, 0.0f);
  }
}

// Merge (rows.rs:795-931) fused with the warm-start decay (mod.rs:1019-1040).
// K_MAX_POINTS lanes per new pair: lane k merges fresh point k; lane j
// also holds old point j, whose anchor every lane tests (in order, so the
// first match wins as in the serial scan); lane 0 writes the pair.
__global__ void __launch_bounds__(TPB) k_merge(Dev d) {
  constexpr int MG = K_MAX_POINTS;
  const uint32_t p = gid() / MG;
  const uint32_t k = gid() % MG;
  if (p >= dev_npairs(d)) return;  // whole groups
  const uint32_t c = d.pair_cand[p];
  uint32_t start = d.pt_off[c];
  uint32_t cnt = d.C.npts[c];
  uint2 ab = d.C.ab[c];
  float4 d0 = d.C.disc0[c], d1 = d.C.disc1[c];
  int old = __float_as_int(d1.z);
  float4 ma = d.b.mat[ab.x], mb = d.b.mat[ab.y];
  if (k == 0) {
    d.P.ab[p] = ab;
    d.P.n(p) = d0;
    d.P.mu(p) = make_float4(d1.x, d1.y, fmax_(ma.z, mb.z), fmax_(ma.w, mb.w));
    d.P.pts[p] = make_uint2(start, cnt);
    if (d1.w != 0.0f) {
      d.b.flag_a[ab.x] = 1;
      d.b.flag_a[ab.y] = 1;
    }
  }
  bool touching = d0.w <= 0.0f;
  bool smooth = d.b.kind[ab.x] == SH_SPHERE || d.b.kind[ab.y] == SH_SPHERE;
  const float AG = K_ALPHA * K_GAMMA;
  uint2 op = old >= 0 ? d.Pold.pts[old] : make_uint2(0u, 0u);
  const f3 oldra = k < op.y ? xyz(d.Qold.ra(op.x + k)) : f3{0.0f, 0.0f, 0.0f};
  f3 ral{0.0f, 0.0f, 0.0f}, rbl{0.0f, 0.0f, 0.0f};
  if (k < cnt) {
    float4 f0 = d.C.fresh[(size_t)c * 16 + 2 * k];
    float4 f1 = d.C.fresh[(size_t)c * 16 + 2 * k + 1];
    ral = f3{f0.x, f0.y, f0.z};
    rbl = f3{f0.w, f1.x, f1.y};
  }
  int m = -1;
  for (uint32_t j = 0; j < op.y; ++j) {
    const f3 o = grp::from_lane<MG>(oldra, (int)j);
    if (m < 0 && len3(sub(o, ral)) < K_LAMBDA_MATCH) m = (int)(op.x + j);
  }
  if (k >= cnt) return;
  uint32_t q = start + k;
  if (m >= 0) {
    float4 ora = d.Qold.ra(m), orb = d.Qold.rb(m);
    float4 l = d.Qold.lam(m), pe = d.Qold.pen(m), ro = d.Qold.roll[m];
    bool stuck = pe.w != 0.0f;
    bool keep = stuck && touching && !smooth;
    f3 ra = keep ? xyz(ora) : ral;
    f3 rb = keep ? xyz(orb) : rbl;
    d.Q.ra(q) = pk(ra, 0.0f);
    d.Q.rb(q) = pk(rb, 0.0f);
    d.Q.lam(q) = make_float4(l.x * AG, l.y * AG, l.z * AG, 0.0f);
    d.Q.pen(q) = make_float4(clampf_(pe.x * K_GAMMA, K_PEN_MIN, K_PEN_MAX), clampf_(pe.y * K_GAMMA, K_PEN_MIN, K_PEN_MAX),
                             clampf_(pe.z * K_GAMMA, K_PEN_MIN, K_PEN_MAX), stuck ? 1.0f : 0.0f);
    d.Q.roll[q] = make_float4(ro.x * AG, ro.y * AG, ro.z * AG, 0.0f);
    d.Q.src[q] = m;
  } else {
    float z = 0.0f * AG;
    float p1 = clampf_(K_PEN_INIT * K_GAMMA, K_PEN_MIN, K_PEN_MAX);
    d.Q.ra(q) = pk(ral, 0.0f);
    d.Q.rb(q) = pk(rbl, 0.0f);
    d.Q.lam(q) = make_float4(z, z, z, 0.0f);
    d.Q.pen(q) = make_float4(p1, p1, p1, 1.0f);
    d.Q.roll[q] = make_float4(z, z, z, 0.0f);
    d.Q.src[q] = -1;
  }
  d.Q.pair[q] = p;
}

// ---------------------------------------------------------------------------
// joint-component propagation (sleep.rs:155-175) over static labels
// ---------------------------------------------------------------------------
__global__ void k_prop_mark(Dev d, uint8_t* flags, uint8_t value) {
  uint32_t h = gid();
  if (h >= d.n) return;
  if (d.b.type[h] == BT_DYNAMIC && flags[h] == value) d.b.comp_any[d.b.comp[h]] = 1;
}
__global__ void k_prop_apply(Dev d, uint8_t* flags, uint8_t value) {
  uint32_t h = gid();
  if (h >= d.n) return;
  if (d.b.type[h] == BT_DYNAMIC && d.b.comp_any[d.b.comp[h]]) flags[h] = value;
}
__global__ void k_prop_clear(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  d.b.comp_any[h] = 0;
}

// Gear (joints_ext.rs:38-161): a gear j keeps its ratio and constant in
// dref.xy and the indices of the two Revolute / Prismatic joints it couples
// in dref.zw (bit patterns); sacc[2j] holds their step-start raw and
// continuous coordinates, sacc[2j+1].x = 1 once those are set.
struct GearSide {
  uint32_t a, b;
  bool ang;
  f3 axis, ra, rb;
  float coord;
};
struct PoseBodies {  // live poses from the body arrays
  const Dev* d;
  __device__ void operator()(uint32_t h, f3& p, qf& q) const {
    p = xyz(d->b.pos[h]);
    q = asq(d->b.rot[h]);
  }
};
struct PoseRecords {  // live poses from the solver records
  const float4* sb;
  __device__ void operator()(uint32_t h, f3& p, qf& q) const {
    p = xyz(sb[4 * (size_t)h]);
    q = asq(sb[4 * (size_t)h + 1]);
  }
};
__device__ __forceinline__ uint32_t gear_ref(const Dev& d, uint32_t j, int k) {
  const float4 g = d.J.dref[j];
  return __float_as_uint(k == 0 ? g.z : g.w);
}
// joint_members (sleep.rs:9-16): a gear's are its two joints' bodies
__device__ __forceinline__ void joint_members_dev(const Dev& d, uint32_t j, uint32_t m[4]) {
  if (d.J.kind[j] == JK_GEAR) {
    const uint2 x = d.J.ab[gear_ref(d, j, 0)], y = d.J.ab[gear_ref(d, j, 1)];
    m[0] = x.x; m[1] = x.y; m[2] = y.x; m[3] = y.y;
  } else {
    const uint2 ab = d.J.ab[j];
    m[0] = m[2] = ab.x;
    m[1] = m[3] = ab.y;
  }
}
// joint_coordinate (joints_ext.rs:38-51) of joint r, with its gear-side
// gradient data
template <class Pose>
__device__ __forceinline__ GearSide gear_side_raw(const Dev& d, uint32_t r, const Pose& pose) {
  GearSide s;
  const uint2 ab = d.J.ab[r];
  s.a = ab.x;
  s.b = ab.y;
  s.ang = d.J.kind[r] == JK_REVOLUTE;
  f3 pa, pb;
  qf qa, qb;
  pose(ab.x, pa, qa);
  pose(ab.y, pb, qb);
  const f3 ax = xyz(d.J.ax[2 * r]);
  const float ref = d.J.dref[r].x;
  s.axis = normalize_or(qrot(qa, ax), f3{0.0f, 0.0f, 1.0f});
  if (s.ang) {
    s.ra = s.rb = f3{0.0f, 0.0f, 0.0f};
    s.coord = hinge_twist(qa, qb, ax) - ref;
  } else {
    s.ra = qrot(qa, xyz(d.J.la[r]));
    s.rb = qrot(qb, xyz(d.J.lb[r]));
    s.coord = dot3(sub(add(pb, s.rb), add(pa, s.ra)), s.axis) - ref;
  }
  return s;
}
// gear_sides (joints_ext.rs:120-161): a hinge side unwrapped against the
// step-start memory
template <class Pose>
__device__ __forceinline__ GearSide gear_side(const Dev& d, uint32_t j, int k, const Pose& pose) {
  GearSide s = gear_side_raw(d, gear_ref(d, j, k), pose);
  if (s.ang && d.J.sacc[2 * j + 1].x != 0.0f) {
    const float4 m = d.J.sacc[2 * j];
    s.coord = (k == 0 ? m.z : m.w) + wrap_pi(s.coord - (k == 0 ? m.x : m.y));
  }
  return s;
}
template <class Pose>
__device__ __forceinline__ float gear_error(const Dev& d, uint32_t j, const Pose& pose) {
  const GearSide sa = gear_side(d, j, 0, pose), sb = gear_side(d, j, 1, pose);
  const float4 g = d.J.dref[j];
  return sa.coord + g.x * sb.coord - g.y;
}
// update_gear_mem (joints_ext.rs:56-115), once per step after the wake
__global__ void k_gear_mem(Dev d) {
  const uint32_t j = gid();
  if (j >= d.njoints || d.J.kind[j] != JK_GEAR) return;
  const PoseBodies pose{&d};
  const GearSide s0 = gear_side_raw(d, gear_ref(d, j, 0), pose), s1 = gear_side_raw(d, gear_ref(d, j, 1), pose);
  float4 m = d.J.sacc[2 * j];
  const bool set = d.J.sacc[2 * j + 1].x != 0.0f;
  m.z = (set && s0.ang) ? m.z + wrap_pi(s0.coord - m.x) : s0.coord;
  m.w = (set && s1.ang) ? m.w + wrap_pi(s1.coord - m.y) : s1.coord;
  m.x = s0.coord;
  m.y = s1.coord;
  d.J.sacc[2 * j] = m;
  d.J.sacc[2 * j + 1] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ bool teleported(const Dev& d, uint32_t h) {  // a non-dynamic body moved by the host
  if (d.b.type[h] == BT_DYNAMIC) return false;
  const float4 p = d.b.pos[h], pp = d.b.prev_pos[h];
  const float4 q = d.b.rot[h], pq = d.b.prev_rot[h];
  return !(p.x == pp.x && p.y == pp.y && p.z == pp.z) || !(q.x == pq.x && q.y == pq.y && q.z == pq.z && q.w == pq.w);
}

// A joint driven (a non-dynamic member teleported) or off its rest pose
// (joint_rest_error, sleep.rs:30-66).
__device__ __forceinline__ bool joint_unsettled(const Dev& d, uint32_t j) {
  if (d.J.kind[j] == JK_GEAR) {
    uint32_t m[4];
    joint_members_dev(d, j, m);
    for (int i = 0; i < 4; ++i)
      if (teleported(d, m[i])) return true;
    return fabsf(gear_error(d, j, PoseBodies{&d})) > K_JOINT_REST;
  }
  uint2 ab = d.J.ab[j];
  uint32_t m[2] = {ab.x, ab.y};
  for (int i = 0; i < 2; ++i) {
    uint32_t h = m[i];
    if (d.b.type[h] != BT_DYNAMIC) {
      float4 p = d.b.pos[h], pp = d.b.prev_pos[h];
      float4 q = d.b.rot[h], pq = d.b.prev_rot[h];
      if (!(p.x == pp.x && p.y == pp.y && p.z == pp.z) || !(q.x == pq.x && q.y == pq.y && q.z == pq.z && q.w == pq.w))
        return true;
    }
  }
  const int kind = d.J.kind[j];
  if (kind == JK_SPRING && ((uint32_t)d.J.dref[j].w & SPRING_KEEPS_AWAKE)) return true;
  if ((kind == JK_REVOLUTE || kind == JK_PRISMATIC || kind == JK_WHEEL) && ((uint32_t)d.J.dref[j].w & HINGE_MOTOR)) {
    const float4 mo = d.J.qref[j];  // a turning motor with a budget keeps its bodies awake
    if (mo.x != 0.0f && mo.y > 0.0f) return true;
  }
  const qf qa = asq(d.b.rot[ab.x]), qb = asq(d.b.rot[ab.y]);
  f3 pa = add(xyz(d.b.pos[ab.x]), qrot(qa, xyz(d.J.la[j])));
  f3 pb = add(xyz(d.b.pos[ab.y]), qrot(qb, xyz(d.J.lb[j])));
  const f3 delta = sub(pb, pa);
  float err;
  if (kind == JK_MOTOR) {  // MotorDrive::keeps_awake (joint.rs:474-477)
    const float4 lt = d.J.ax[2 * j], at = d.J.ax[2 * j + 1];
    if ((lt.w > 0.0f && (lt.x != 0.0f || lt.y != 0.0f || lt.z != 0.0f)) ||
        (at.w > 0.0f && (at.x != 0.0f || at.y != 0.0f || at.z != 0.0f)))
      return true;
  }
  if (kind == JK_BALL) {
    err = len3(delta);
  } else if (is_tether(kind)) {  // a slack rope is at rest
    const float e = len3(delta) - d.J.dref[j].x;
    err = kind == JK_ROPE ? fmax_(e, 0.0f) : fabsf(e);
  } else if (kind == JK_SIXDOF) {  // sixdof_rest_error (sleep.rs:86-103)
    const f3 lin = sub(qrot(qconj(qa), delta), xyz(d.J.dref[j]));
    const f3 ang = quat_diff_vec(qmul(qconj(qa), qb), asq(d.J.qref[j]));
    const float4 w0 = d.J.six[4 * j], w1 = d.J.six[4 * j + 1], w2 = d.J.six[4 * j + 2], w3 = d.J.six[4 * j + 3];
    err = 0.0f;
    for (int i = 0; i < 3; ++i) {
      for (int s2 = 0; s2 < 2; ++s2) {
        const int k = 3 * s2 + i;
        const float v = s2 ? comp(ang, i) : comp(lin, i);
        const float lo = s2 ? comp(xyz(w2), i) : comp(xyz(w0), i), hi = s2 ? comp(xyz(w3), i) : comp(xyz(w1), i);
        const int cfg = six_cfg(w0.w, k);
        float r = 0.0f;
        if (cfg == AX_LOCKED) r = fabsf(v);
        if (cfg == AX_LIMITED) r = fmax_(fmax_(lo - v, v - hi), 0.0f);
        err = fmax_(err, r);
      }
    }
  } else if (kind == JK_WHEEL) {  // joint_rest_error (sleep.rs:61-76)
    const f3 axis = qrot(qa, xyz(d.J.sus[j]));
    const float4 pr = d.J.dref[j];
    err = len3(sub(delta, mul(axis, dot3(delta, axis))));
    err = fmax_(err, len3(sub(qrot(qa, xyz(d.J.ax[2 * j])), qrot(qb, xyz(d.J.ax[2 * j + 1])))));
    if (!(pr.y > 0.0f)) err = fmax_(err, fabsf(dot3(delta, axis) - pr.x));
  } else if (kind == JK_FIXED || kind == JK_MOTOR) {  // joint_rest_error (sleep.rs:30-82)
    err = fmax_(len3(sub(delta, qrot(qa, xyz(d.J.dref[j])))), len3(quat_diff_vec(qmul(qconj(qa), qb), asq(d.J.qref[j]))));
  } else {
    const f3 axis = qrot(qa, xyz(d.J.ax[2 * j]));
    const float e_axis = len3(sub(axis, qrot(qb, xyz(d.J.ax[2 * j + 1]))));
    if (kind == JK_REVOLUTE) {
      err = fmax_(len3(delta), e_axis);
    } else {
      err = len3(sub(delta, mul(axis, dot3(delta, axis))));
      const float4 pr = d.J.dref[j];
      if ((uint32_t)pr.w & HINGE_LIMIT) {  // outside the travel window
        const float value = dot3(delta, axis) - pr.x;
        err = fmax_(err, fmax_(pr.y - value, value - pr.z));
      }
      err = fmax_(err, e_axis);
    }
  }
  return err > K_JOINT_REST;
}

// wake_joint_motion (sleep.rs:107-137): per-body part.
__device__ __forceinline__ void awake_init_body(const Dev& d, uint32_t h) {
  uint8_t aw = 0;
  if (d.b.type[h] == BT_DYNAMIC) {
    float4 v = d.b.vel[h], w = d.b.angvel[h], t = d.b.torque[h];
    float4 p = d.b.pos[h], pp = d.b.prev_pos[h], q = d.b.rot[h], pq = d.b.prev_rot[h];
    bool nz = v.x != 0.0f || v.y != 0.0f || v.z != 0.0f || w.x != 0.0f || w.y != 0.0f || w.z != 0.0f ||
              t.x != 0.0f || t.y != 0.0f || t.z != 0.0f;
    bool moved = p.x != pp.x || p.y != pp.y || p.z != pp.z || q.x != pq.x || q.y != pq.y || q.z != pq.z ||
                 q.w != pq.w;
    aw = (!d.b.asleep[h] || nz || moved) ? 1 : 0;
  }
  d.b.flag_b[h] = aw;
}
__global__ void k_awake_init(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  awake_init_body(d, h);
}
__global__ void k_awake_joints(Dev d) {
  uint32_t j = gid();
  if (j >= d.njoints) return;
  if (joint_unsettled(d, j)) {
    uint32_t m[4];
    joint_members_dev(d, j, m);
    for (int i = 0; i < 4; ++i)
      if (d.b.type[m[i]] == BT_DYNAMIC) d.b.flag_b[m[i]] = 1;
  }
}
__device__ __forceinline__ void awake_apply_body(const Dev& d, uint32_t h) {
  if (d.b.flag_b[h] && d.b.asleep[h]) wake_body_dev(d, h);
}
__global__ void k_awake_apply(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  awake_apply_body(d, h);
}

// External wake (mod.rs:984-992).
__device__ __forceinline__ void external_wake_body(const Dev& d, uint32_t h) {
  if (!d.b.asleep[h]) return;
  if (len3(xyz(d.b.vel[h])) > K_SLEEP_LIN || len3(xyz(d.b.angvel[h])) > K_SLEEP_ANG) wake_body_dev(d, h);
}
__global__ void k_external_wake(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  external_wake_body(d, h);
}

// ---------------------------------------------------------------------------
// inertial targets + warm-started positions (mod.rs:993-1049)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void inertial_body(const Dev& d, uint32_t h) {
  float4 p = d.b.pos[h];
  float4 q = d.b.rot[h];
  float4 p0, q0;
  if (d.b.type[h] == BT_KINEMATIC) {
    p0 = d.b.prev_pos[h];
    q0 = d.b.prev_rot[h];
  } else {
    p0 = p;
    q0 = q;
  }
  float srad = d.b.kind[h] == SH_SPHERE ? d.b.shape[h].x : -1.0f;
  d.b.pos0[h] = make_float4(p0.x, p0.y, p0.z, srad);
  d.b.rot0[h] = q0;
  bool solvable = d.b.type[h] == BT_DYNAMIC && p.w > 0.0f;
  if (!solvable) {
    d.b.inertial[h] = p;
    d.b.inertial_rot[h] = q;
  } else {
    f3 g{d.gravity[0], d.gravity[1], d.gravity[2]};
    f3 y = add(add(xyz(p), mul(xyz(d.b.vel[h]), K_DT)), mul(g, K_DT * K_DT));
    qf yq = quat_integrate(asq(q), mul(xyz(d.b.angvel[h]), K_DT));
    d.b.inertial[h] = pk(y, 0.0f);
    d.b.inertial_rot[h] = pkq(yq);
    if (!d.b.asleep[h]) {
      p = pk(y, p.w);
      q = pkq(yq);
      d.b.pos[h] = p;
      d.b.rot[h] = q;
    }
  }
  f3 dr = quat_diff_vec(asq(q), asq(q0));
  d.b.drot[h] = pk(dr, 0.0f);
}
__global__ void k_inertial(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  inertial_body(d, h);
}

// Scenes without joints: the wake passes after the narrowphase (votes of
// new contacts, then with sleepers about, joint-free motion and external
// velocities) and the inertial targets, all per body, in one launch.
__global__ void k_wake_inertial(Dev d, uint32_t may_sleep) {
  uint32_t h = gid();
  if (h >= d.n) return;
  apply_wake_body(d, h);
  if (may_sleep) {
    awake_init_body(d, h);
    awake_apply_body(d, h);
    external_wake_body(d, h);
  }
  inertial_body(d, h);
}

// Scenes without joints: the step's opening wake pass (with sleepers about)
// and k_begin's per-body work in one launch.
__global__ void __launch_bounds__(TPB) k_wake_begin(Dev d, uint32_t may_sleep) {
  uint32_t h = gid();
  uint8_t fast = 0;
  if (h < d.n) {
    if (may_sleep) {
      awake_init_body(d, h);
      awake_apply_body(d, h);
    }
    fast = begin_body(d, h);
  }
  block_add(&d.counters[CTR_FAST], fast);
}

// Joint warm-start decay (mod.rs:1034-1041) and the step-start terms of
// the joint rows: the anchor separation (Fixed: plus rot0_a dref,
// assembly.rs:286-290), the hinge rows' tangent basis of the step-start
// axis with t . (axa0 - axb0) (:620-634), the Fixed step-start rotation
// error (:790).
__device__ __forceinline__ float4 decay_lam(float4 l) {
  const float AG = K_ALPHA * K_GAMMA;
  return make_float4(l.x * AG, l.y * AG, l.z * AG, l.w);  // w (a limit accumulator) does not decay
}
__device__ __forceinline__ float4 decay_pen(float4 pe) {
  return make_float4(clampf_(pe.x * K_GAMMA, K_PEN_MIN, K_PEN_MAX), clampf_(pe.y * K_GAMMA, K_PEN_MIN, K_PEN_MAX),
                     clampf_(pe.z * K_GAMMA, K_PEN_MIN, K_PEN_MAX), 0.0f);
}
__global__ void k_joint_prep(Dev d) {
  uint32_t j = gid();
  if (j >= d.njoints) return;
  d.J.lam[j] = decay_lam(d.J.lam[j]);
  d.J.pen[j] = decay_pen(d.J.pen[j]);
  d.J.lama[j] = decay_lam(d.J.lama[j]);
  d.J.pena[j] = decay_pen(d.J.pena[j]);
  const int kind = d.J.kind[j];
  uint2 ab = d.J.ab[j];
  const qf ra0 = asq(d.b.rot0[ab.x]), rb0 = asq(d.b.rot0[ab.y]);
  f3 pa0 = add(xyz(d.b.pos0[ab.x]), qrot(ra0, xyz(d.J.la[j])));
  f3 pb0 = add(xyz(d.b.pos0[ab.y]), qrot(rb0, xyz(d.J.lb[j])));
  f3 c0v = sub(pa0, pb0);
  if (kind == JK_FIXED) c0v = add(c0v, qrot(ra0, xyz(d.J.dref[j])));
  d.J.c0v[j] = pk(c0v, 0.0f);
  if ((kind == JK_REVOLUTE || kind == JK_PRISMATIC) && ((uint32_t)d.J.dref[j].w & HINGE_LIMIT)) {
    const float ref = d.J.dref[j].x;
    float v0;
    if (kind == JK_REVOLUTE) {
      v0 = wrap_pi(hinge_twist(ra0, rb0, xyz(d.J.ax[2 * j])) - ref);
    } else {
      const f3 wa0 = normalize_or(qrot(ra0, xyz(d.J.ax[2 * j])), f3{0.0f, 0.0f, 1.0f});
      v0 = dot3(sub(pb0, pa0), wa0) - ref;
    }
    d.J.c0v[j] = pk(c0v, v0);
  }
  if (kind == JK_REVOLUTE || kind == JK_PRISMATIC || kind == JK_WHEEL) {  // (a wheel: its axle)
    const f3 axa0 = qrot(ra0, xyz(d.J.ax[2 * j])), axb0 = qrot(rb0, xyz(d.J.ax[2 * j + 1]));
    f3 t1, t2;
    tangent_basis(axa0, &t1, &t2);
    const f3 dax = sub(axa0, axb0);
    d.J.a0[2 * j] = pk(t1, dot3(t1, dax));
    d.J.a0[2 * j + 1] = pk(t2, dot3(t2, dax));
  } else if (kind == JK_FIXED) {
    d.J.a0[2 * j] = pk(quat_diff_vec(qmul(qconj(ra0), rb0), asq(d.J.qref[j])), 0.0f);
  }
}

// Per-pair step constants: friction frame, separated-damper stiffness
// (assembly.rs:129-174), step-start violation c0 per point (joints_ext.rs:21-32).
// PREP_LANES lanes per pair, lane s taking points s, s + PREP_LANES, ...
// friction_frame's first direction (rows.rs:32-52) from the live rotations
// qa, qb of the pair's bodies: body a's axis, else b's, projected on the
// contact plane and normalised (UnitVec3::normalize_checked,
// invariants.rs:55-60: none when not finite or shorter than
// DEGENERATE_LEN2), else tangent_basis(n).
__device__ __forceinline__ f3 friction_t1(const Dev& d, uint32_t a, uint32_t b, qf qa, qf qb, f3 n) {
  const float4 fd[2] = {d.b.fdir[a], d.b.fdir[b]};
  const qf q[2] = {qa, qb};
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (!has_axis(fd[k])) continue;
    const f3 world = qrot(q[k], xyz(fd[k]));
    const f3 proj = sub(world, mul(n, dot3(world, n)));
    const float l2 = dot3(proj, proj);
    if (!(isfinite(proj.x) && isfinite(proj.y) && isfinite(proj.z)) || l2 < 1e-12f) continue;
    return mul(proj, 1.0f / sqrtf(dot3(proj, proj)));
  }
  f3 t1, t2;
  tangent_basis(n, &t1, &t2);
  return t1;
}

__global__ void __launch_bounds__(TPB) k_pair_prep(Dev d) {
  const uint32_t p = gid() / PREP_LANES;
  const uint32_t s = threadIdx.x % PREP_LANES;
  if (p >= dev_npairs(d)) return;
  uint2 ab = d.P.ab[p];
  float4 n4 = d.P.n(p);
  f3 n = xyz(n4);
  float gap = n4.w;
  f3 t1, t2u;
  tangent_basis(n, &t1, &t2u);
  f3 t2 = cross3(t1, n);
  uint2 pts = d.P.pts[p];
  float damp = 0.0f;
  if (!(gap <= 0.0f)) {
    float approach = fmax_(dot3(sub(xyz(d.b.vel[ab.y]), xyz(d.b.vel[ab.x])), n), 0.0f);
    if (!(gap - approach * K_DT > 0.0f)) {
      float npts = (float)(pts.y > 1u ? pts.y : 1u);
      float ka = d.b.type[ab.x] == BT_DYNAMIC ? d.b.pos[ab.x].w : 0.0f;
      float kb = d.b.type[ab.y] == BT_DYNAMIC ? d.b.pos[ab.y].w : 0.0f;
      float m_red = ka + kb > K_NEAR_ZERO ? 1.0f / (ka + kb) : 0.0f;
      if (m_red > 0.0f) damp = fmin_(2.0f * m_red / (K_DT * K_DT) / npts, K_PEN_MAX);
    }
  }
  if (s == 0) {
    // t2.w marks a pair whose friction frame follows its bodies' axes: the
    // solves and the dual then take it from the live rotations
    const bool aniso = has_axis(d.b.fdir[ab.x]) || has_axis(d.b.fdir[ab.y]);
    d.P.t1(p) = pk(t1, damp);
    d.P.t2(p) = pk(t2, aniso ? 1.0f : 0.0f);
  }
  if (s >= pts.y) return;
  bool sphere = d.b.kind[ab.x] == SH_SPHERE || d.b.kind[ab.y] == SH_SPHERE;
  float4 pa0 = d.b.pos0[ab.x], pb0 = d.b.pos0[ab.y];
  qf qa0 = asq(d.b.rot0[ab.x]), qb0 = asq(d.b.rot0[ab.y]);
  for (uint32_t k = s; k < pts.y; k += PREP_LANES) {
    uint32_t q = pts.x + k;
    float4 ra = d.Q.ra(q);
    float c0 = 0.0f;
    if (gap <= 0.0f) {
      if (sphere) {
        c0 = gap + K_MARGIN;
      } else {
        f3 pa = add(xyz(pa0), qrot(qa0, xyz(ra)));
        f3 pb = add(xyz(pb0), qrot(qb0, xyz(d.Q.rb(q))));
        c0 = dot3(n, sub(pa, pb)) + K_MARGIN;
      }
    }
    ra.w = c0;
    d.Q.ra(q) = ra;
  }
}

// ---------------------------------------------------------------------------
// per-body incidence (point-sides in pair order) and colouring
// ---------------------------------------------------------------------------
// Incidence is built for swept bodies only (the primal solves nothing
// else); the colouring graph links swept bodies sharing a pair with
// points or a joint (the graph of colour_bodies).

// Per pair: the b side's counts (scattered bodies, few atomics each); the
// a side of a whole a-run is summed by the run's first pair.
__global__ void __launch_bounds__(TPB) k_inc_count(Dev d) {
  uint32_t p = gid();
  uint32_t side_pairs = 0;
  if (p < dev_npairs(d)) {
    uint2 ab = d.P.ab[p];
    uint32_t c = d.P.pts[p].y;
    bool sa = d.sw[ab.x], sb = d.sw[ab.y];
    if (sb) {
      atomicAdd(&d.ptcnt[ab.y], c);
      atomicAdd(&d.bcnt[ab.y], 1u);
      if (c && sa && d.ckey[ab.x] > d.ckey[ab.y]) atomicAdd(&d.nbr_cnt[ab.y], 1u);  // a is above b
    }
    if (sa && p == d.poff[ab.x]) {
      uint32_t pts = 0, nb = 0;
      const unsigned long long ka = d.ckey[ab.x];
      for (uint32_t q = p; q < d.poff[ab.x + 1]; ++q) {
        uint32_t cq = d.P.pts[q].y;
        const uint32_t bq = d.P.ab[q].y;
        pts += cq;
        nb += (cq && d.sw[bq] && d.ckey[bq] > ka) ? 1u : 0u;  // b_q above a
      }
      atomicAdd(&d.ptcnt[ab.x], pts);
      if (nb) atomicAdd(&d.nbr_cnt[ab.x], nb);
    }
    side_pairs = c ? (uint32_t)sa + (uint32_t)sb : 0u;
  }
  block_add(&d.counters[CTR_PAIRSIDES], side_pairs);
}
__global__ void __launch_bounds__(TPB) k_nbr_joint_count(Dev d) {
  uint32_t j = gid();
  uint32_t sides = 0;
  if (j < d.njoints) {
    // a gear's rows read all four of its members: it links each two
    uint32_t m[4];
    joint_members_dev(d, j, m);
    const bool gear = d.J.kind[j] == JK_GEAR;
    for (int x = 0; x < 4; ++x) {
      bool first = gear || x < 2;
      for (int y = 0; y < x; ++y) first = first && m[y] != m[x];
      sides += (first && d.sw[m[x]]) ? 1u : 0u;
      for (int y = x + 1; y < 4; ++y)
        if ((gear || (x == 0 && y == 1)) && m[x] != m[y] && d.sw[m[x]] && d.sw[m[y]])
          atomicAdd(&d.nbr_cnt[d.ckey[m[y]] > d.ckey[m[x]] ? m[x] : m[y]], 1u);
    }
  }
  block_add(&d.counters[CTR_JSIDES], sides);
}
// Colouring neighbours of a swept pair or joint (a, b): the body of lower
// priority lists the other (each body lists its higher-priority neighbours).
__device__ __forceinline__ void nbr_link(const Dev& d, uint2 ab) {
  const bool b_up = d.ckey[ab.y] > d.ckey[ab.x];
  const uint32_t lo = b_up ? ab.x : ab.y, hi = b_up ? ab.y : ab.x;
  d.nbr[d.nbr_off[lo] + atomicAdd(&d.nbr_cur[lo], 1u)] = hi;
}
__global__ void k_inc_fill(Dev d) {
  uint32_t p = gid();
  if (p >= dev_npairs(d)) return;
  uint2 ab = d.P.ab[p];
  bool sa = d.sw[ab.x], sb = d.sw[ab.y];
  if (sb) d.blist[d.boff[ab.y] + atomicAdd(&d.bcur[ab.y], 1u)] = p;
  if (sa && sb && d.P.pts[p].y) nbr_link(d, ab);
}
__global__ void k_nbr_joint_fill(Dev d) {
  uint32_t j = gid();
  if (j >= d.njoints) return;
  uint32_t m[4];
  joint_members_dev(d, j, m);
  const bool gear = d.J.kind[j] == JK_GEAR;
  for (int x = 0; x < 4; ++x)
    for (int y = x + 1; y < 4; ++y)
      if ((gear || (x == 0 && y == 1)) && m[x] != m[y] && d.sw[m[x]] && d.sw[m[y]]) nbr_link(d, make_uint2(m[x], m[y]));
}
// Body h's point-sides in ascending pair index: pairs where h is b (a < h),
// then the contiguous run where h is a.  INC_LANES lanes per body: lane s
// places the b-side pairs s, s + INC_LANES, ... of h's (unordered) b-list
// after the points of the b-side pairs of smaller index, and the a-run
// pairs s, s + INC_LANES, ..., whose points follow each other in the point
// array (they are numbered in pair order).
__global__ void __launch_bounds__(TPB) k_inc_build(Dev d) {
  const uint32_t h = gid() / INC_LANES;
  const uint32_t s = threadIdx.x % INC_LANES;
  if (h >= d.n || !d.sw[h]) return;  // whole groups
  const uint32_t b0 = d.boff[h], nb = d.boff[h + 1] - b0;
  const uint32_t w0 = d.inc_off[h];
  uint32_t bpts = 0;
  for (uint32_t i = s; i < nb; i += INC_LANES) {
    const uint32_t p = d.blist[b0 + i];
    uint32_t pre = 0;
    for (uint32_t j = 0; j < nb; ++j) {
      const uint32_t q = d.blist[b0 + j];
      if (q < p) pre += d.P.pts[q].y;
    }
    const uint2 pts = d.P.pts[p];
    const uint2 ab = d.P.ab[p];
    for (uint32_t k = 0; k < pts.y; ++k) d.inc[w0 + pre + k] = make_uint4((pts.x + k) | 0x80000000u, p, ab.x, ab.y);
    bpts += pts.y;
  }
  for (int m = INC_LANES / 2; m > 0; m >>= 1) bpts += __shfl_xor(bpts, m, INC_LANES);
  const uint32_t a0 = d.acc_off[d.cand_off[h]], a1 = d.acc_off[d.cand_off[h + 1]];
  if (a0 >= a1) return;
  const uint32_t x0 = d.P.pts[a0].x;
  for (uint32_t p = a0 + s; p < a1; p += INC_LANES) {
    const uint2 pts = d.P.pts[p];
    const uint2 ab = d.P.ab[p];
    const uint32_t w = w0 + bpts + (pts.x - x0);
    for (uint32_t k = 0; k < pts.y; ++k) d.inc[w + k] = make_uint4(pts.x + k, p, ab.x, ab.y);
  }
}

// Jones-Plassmann colouring, pull form.  A swept body is coloured once all
// its higher-priority neighbours are, with the smallest colour none of them
// holds -- exactly the sequential greedy colouring in priority order (larger
// hash first, then smaller scene-local index; the key packs both, so a
// larger key is a higher priority; bodies of different scenes are never
// neighbours, so keys of neighbours differ).  Each body lists only its
// higher-priority neighbours; passes over the pending bodies repeat until
// none is left (no counters are decremented, so a pass costs no atomics per
// neighbour).
__device__ __forceinline__ bool swept(const Dev& d, uint32_t h) {
  return d.b.type[h] == BT_DYNAMIC && d.b.pos[h].w > 0.0f && !d.b.asleep[h];
}

// Also clears the per-step incidence and colouring counters.
__global__ void k_colour_init(Dev d) {
  uint32_t h = gid();
  if (h < CSTAT_HEAD + 65536) d.cstat[h] = 0u;
  if (h < 65536) d.colour_cur[h] = 0u;
  if (h < 2 * CCLS) d.ccls[h] = 0u;
  if (h > d.n) return;
  d.ptcnt[h] = 0u;
  d.bcnt[h] = 0u;
  d.bcur[h] = 0u;
  d.nbr_cnt[h] = 0u;
  d.nbr_cur[h] = 0u;
  if (h == d.n) return;
  d.sw[h] = swept(d, h) ? 1 : 0;
  d.ckey[h] = ((unsigned long long)d.b.prio[h] << 32) | (0xffffffffu - d.b.local[h]);
  d.b.colour[h] = -1;
}

// Size class of a body for the primal's colour lists: class 0 holds the
// bodies with the most items (point-sides and joint-sides), class 7 the
// fewest.  Within a colour the list runs from class 0 to class 7, so the
// bodies sharing a wavefront have similar work and the longest solves start
// first.
__device__ __forceinline__ uint32_t item_class(uint32_t items) {
  return items > 96u ? 0u : items > 64u ? 1u : items > 48u ? 2u : items > 32u ? 3u
       : items > 24u ? 4u : items > 16u ? 5u : items > 8u ? 6u : 7u;
}
__device__ __forceinline__ uint32_t body_items(const Dev& d, uint32_t h) {
  return d.inc_off[h + 1] - d.inc_off[h] + d.b.ji_off[h + 1] - d.b.ji_off[h];
}

// Appends v to list for every active lane whose pred holds (one atomic per
// wavefront).
__device__ __forceinline__ void wave_push(uint32_t* list, uint32_t* n, bool pred, uint32_t v) {
  unsigned long long m = __ballot(pred);
  if (!m) return;
  const uint32_t lane = __lane_id();
  const int leader = __ffsll(m) - 1;
  uint32_t base = 0;
  if (lane == (uint32_t)leader) base = atomicAdd(n, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (pred) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = v;
}

// Colour of body h, by the CL lanes of its group: if all of h's
// higher-priority neighbours have colours, h takes the smallest colour none
// of them uses and this returns false; otherwise h stays uncoloured and
// this returns true.  A neighbour coloured concurrently may be seen or not;
// either way a body is coloured only from the final colours of all its
// higher-priority neighbours, so the colouring is the sequential greedy one
// in priority order.  All lanes of the group call this together.
#define CL 8
__device__ __forceinline__ bool colour_body(const Dev& d, uint32_t h, bool act) {
  constexpr int CU = 4;  // colour reads in flight per lane
  const uint32_t s = threadIdx.x % CL;
  const uint32_t k0 = act ? d.nbr_off[h] : 0u, k1 = k0 + (act ? d.nbr_cur[h] : 0u);
  int newc = -1;
  bool waiting = false;
  for (int cb = 0; act && newc < 0 && !waiting; cb += 64) {
    unsigned long long used = 0;
    int wait = 0;
    for (uint32_t k = k0 + s; k < k1; k += CU * CL) {
      int co[CU];
#pragma unroll
      for (int u = 0; u < CU; ++u) co[u] = k + u * CL < k1 ? d.b.colour[d.nbr[k + u * CL]] : 0x7fffffff;
#pragma unroll
      for (int u = 0; u < CU; ++u) {
        wait |= co[u] < 0 ? 1 : 0;
        if (co[u] >= cb && co[u] < cb + 64) used |= 1ull << (co[u] - cb);
      }
    }
    for (int m = CL / 2; m > 0; m >>= 1) {
      used |= __shfl_xor(used, m, CL);
      wait |= __shfl_xor(wait, m, CL);
    }
    waiting = wait != 0;
    if (!waiting && ~used != 0ull) newc = cb + __builtin_ctzll(~used);
  }
  if (act && s == 0 && !waiting) d.b.colour[h] = newc;
  return act && waiting;
}

// One pass over all bodies, CL lanes each: swept bodies still uncoloured
// try to take their colour.  *deferred receives the number left
// uncoloured (one atomic per workgroup).
__global__ void __launch_bounds__(COLOUR_PASS_TPB) k_colour_pass(Dev d, uint32_t* deferred) {
  __shared__ uint32_t s_def;
  if (threadIdx.x == 0) s_def = 0u;
  __syncthreads();
  const uint32_t h = gid() / CL;
  const bool act = h < d.n && d.sw[h] && d.b.colour[h] < 0;
  const bool waiting = colour_body(d, h, act);
  if (waiting && threadIdx.x % CL == 0) atomicAdd(&s_def, 1u);
  __syncthreads();
  if (threadIdx.x == 0 && s_def) atomicAdd(deferred, s_def);
}

// One pass over a list of uncoloured bodies (round r of the ring: reads
// front list r & 1 with counter r % 3, lists the bodies still waiting into
// list (r + 1) & 1 / counter (r + 1) % 3, and clears counter (r + 2) % 3 for
// the next round), CL lanes per body over a grid-stride loop: the list's
// length is only known on the device.  *deferred receives the number left.
__global__ void __launch_bounds__(COLOUR_PASS_TPB) k_colour_pass_list(Dev d, uint32_t round, uint32_t* deferred) {
  constexpr uint32_t PER = COLOUR_PASS_TPB / CL;  // bodies per workgroup and sweep of the loop
  __shared__ uint32_t s_n, s_base, s_def;
  const size_t n = d.n;
  const uint32_t* in_list = d.front + (size_t)(round & 1u) * n;
  uint32_t* out_list = d.front + (size_t)((round + 1u) & 1u) * n;
  uint32_t* out_n = d.fcnt + (round + 1u) % 3u;
  const uint32_t nin = d.fcnt[round % 3u];
  if (gid() == 0) d.fcnt[(round + 2u) % 3u] = 0u;
  if (threadIdx.x == 0) s_def = 0u;
  const bool lead = threadIdx.x % CL == 0;
  // the loop bound is uniform over the workgroup, so its barriers are too;
  // the bodies left are listed with one atomic per workgroup and sweep
  for (uint32_t b0 = blockIdx.x * PER; b0 < nin; b0 += gridDim.x * PER) {
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    const uint32_t t = b0 + threadIdx.x / CL;
    const bool act = t < nin;
    const uint32_t h = act ? in_list[t] : 0u;
    const bool waiting = colour_body(d, h, act);
    const uint32_t slot = waiting && lead ? atomicAdd(&s_n, 1u) : 0u;
    __syncthreads();
    if (threadIdx.x == 0 && s_n) {
      s_base = atomicAdd(out_n, s_n);
      s_def += s_n;
    }
    __syncthreads();
    if (waiting && lead) out_list[s_base + slot] = h;
  }
  if (threadIdx.x == 0 && s_def) atomicAdd(deferred, s_def);
}

// The swept bodies still uncoloured, listed into `list` / *list_n (one
// atomic per workgroup).
__global__ void __launch_bounds__(TPB) k_colour_collect(Dev d, uint32_t* list, uint32_t* list_n) {
  __shared__ uint32_t s_n, s_base;
  if (threadIdx.x == 0) s_n = 0u;
  __syncthreads();
  const uint32_t h = gid();
  const bool pend = h < d.n && d.sw[h] && d.b.colour[h] < 0;
  const uint32_t slot = pend ? atomicAdd(&s_n, 1u) : 0u;
  __syncthreads();
  if (threadIdx.x == 0 && s_n) s_base = atomicAdd(list_n, s_n);
  __syncthreads();
  if (pend) list[s_base + slot] = h;
}

// The remaining passes, from round0 on, in one workgroup over the listed
// bodies: passes are separated by workgroup barriers instead of launches
// (few bodies are still uncoloured by then).  Writes the number of passes to
// cstat[CSTAT_ROUNDS].  The loop ends when no body is left; every pass
// colours at least the listed body of highest priority, so n + 1 passes
// always suffice.  Lists alternate between two buffers, their counters
// rotate over three (pass r reads r % 3, fills (r + 1) % 3, clears
// (r + 2) % 3).
__global__ void __launch_bounds__(COLOUR_TAIL_TPB) k_colour_tail(Dev d, uint32_t round0) {
  constexpr uint32_t PER = COLOUR_TAIL_TPB / CL;
  const size_t n = d.n;
  for (uint32_t round = round0; round <= round0 + (uint32_t)n; ++round) {
    const uint32_t* in_list = d.front + (size_t)(round & 1u) * n;
    uint32_t* out_list = d.front + (size_t)((round + 1u) & 1u) * n;
    uint32_t* out_n = d.fcnt + (round + 1u) % 3u;
    // counters are updated by device atomics: read them past this CU's cache
    const uint32_t nin = __hip_atomic_load(d.fcnt + round % 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (threadIdx.x == 0) {
      __hip_atomic_store(d.fcnt + (round + 2u) % 3u, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (nin == 0u) d.cstat[CSTAT_ROUNDS] = round;
    }
    if (nin == 0u) break;
    for (uint32_t base = 0; base < nin; base += PER) {
      const uint32_t i = base + threadIdx.x / CL;
      const bool act = i < nin;
      const uint32_t h = act ? in_list[i] : 0u;
      const bool waiting = colour_body(d, h, act);
      wave_push(out_list, out_n, waiting && threadIdx.x % CL == 0, h);
    }
    __threadfence();
    __syncthreads();
  }
}

// Bodies per colour and per (colour, size class), and the number of colours,
// from per-block LDS histograms (colours from CCLS_COLOURS on, rare, take
// one atomic per body).
// Exclusive offsets of the colour classes (colour_off[0..ncol], ncol the
// colour count in cstat) from their sizes: one wavefront, 64 colours a round.
// More colours than the `slots` primal launches the host queued is
// recorded as an overflow.
__global__ void __launch_bounds__(64) k_colour_offsets(Dev d, uint32_t slots) {
  constexpr int WV = 64;
  const uint32_t ncol = d.cstat[CSTAT_MAX];
  const int lane = (int)threadIdx.x;
  if (lane == 0 && ncol > slots) note_overflow(d, OVF_COLOURS, ncol);
  uint32_t base = 0;
  for (uint32_t c0 = 0; c0 < ncol; c0 += WV) {
    const uint32_t k = c0 + (uint32_t)lane;
    const uint32_t v = k < ncol ? d.colour_cnt[k] : 0u;
    uint32_t x = v;
    for (int o = 1; o < WV; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, WV);
      if (lane >= o) x += y;
    }
    if (k < ncol) d.colour_off[k] = base + x - v;
    base += __shfl(x, WV - 1, WV);
  }
  if (lane == 0) d.colour_off[ncol] = base;
}

__global__ void __launch_bounds__(TPB) k_colour_classes(Dev d) {
  __shared__ uint32_t cnt[CCLS];
  __shared__ uint32_t s_max;
  uint32_t h = gid();
  if (h == 0) d.counters[CTR_PTSIDES] = d.inc_off[d.n];  // point-sides of swept bodies
  for (uint32_t i = threadIdx.x; i < CCLS; i += TPB) cnt[i] = 0u;
  if (threadIdx.x == 0) s_max = 0u;
  __syncthreads();
  const int c = h < d.n ? d.b.colour[h] : -1;
  if (c >= 0 && c < CCLS_COLOURS) atomicAdd(&cnt[(uint32_t)c * 8u + item_class(body_items(d, h))], 1u);
  if (c >= CCLS_COLOURS) atomicAdd(&d.colour_cnt[c], 1u);
  if (c >= 0) atomicMax(&s_max, (uint32_t)c + 1u);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < CCLS; i += TPB)
    if (cnt[i]) atomicAdd(&d.ccls[i], cnt[i]);
  if (threadIdx.x < CCLS_COLOURS) {
    uint32_t t = 0;
    for (int k = 0; k < 8; ++k) t += cnt[threadIdx.x * 8 + k];
    if (t) atomicAdd(&d.colour_cnt[threadIdx.x], t);
  }
  if (threadIdx.x == 0 && s_max) atomicMax(d.cmax, s_max);
}

// Colour lists: each colour below CCLS_COLOURS is laid out by size class
// (class 0 first); slots within a block come from LDS counters, with one
// global atomic per (colour, class) per block.  Colours beyond (rare) take
// one atomic per body and are not ordered.
__global__ void __launch_bounds__(TPB) k_colour_fill(Dev d) {
  __shared__ uint32_t cnt[CCLS], base[CCLS];
  uint32_t h = gid();
  for (uint32_t i = threadIdx.x; i < CCLS; i += TPB) cnt[i] = 0u;
  __syncthreads();
  const int c = h < d.n ? d.b.colour[h] : -1;
  uint32_t items = 0, slot = 0, local = 0;
  if (c >= 0) items = body_items(d, h);
  if (c >= 0 && c < CCLS_COLOURS) {
    slot = (uint32_t)c * 8u + item_class(items);
    local = atomicAdd(&cnt[slot], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < CCLS; i += TPB)
    if (cnt[i]) base[i] = atomicAdd(&d.ccls[CCLS + i], cnt[i]);
  __syncthreads();
  if (c < 0) return;
  // the primal's per-body record: body (bit 31: has joint sides), first
  // incidence item, item count, first joint side
  const uint32_t i0 = d.inc_off[h], j0 = d.b.ji_off[h];
  const uint32_t nj = d.b.ji_off[h + 1] - j0;
  const uint4 rec = make_uint4(h | (nj ? 0x80000000u : 0u), i0, d.inc_off[h + 1] - i0, j0);
  if (c < CCLS_COLOURS) {
    uint32_t off = d.colour_off[c];
    for (uint32_t k = (uint32_t)c * 8u; k < slot; ++k) off += d.ccls[k];
    d.colour_list[off + base[slot] + local] = rec;
  } else {
    d.colour_list[d.colour_off[c] + atomicAdd(&d.colour_cur[c], 1u)] = rec;
  }
}

// Solver records for the iteration loop (pos, rot, pos0, drot per body);
// the loop reads and writes only these, the arrays are refreshed after it.
__global__ void k_sb_pack(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  float4* rec = d.sb + 4 * (size_t)h;
  rec[0] = d.b.pos[h];
  rec[1] = d.b.rot[h];
  rec[2] = d.b.pos0[h];
  rec[3] = d.b.drot[h];
}
__device__ __forceinline__ void sb_unpack_body(const Dev& d, uint32_t h) {
  const float4* rec = d.sb + 4 * (size_t)h;
  d.b.pos[h] = rec[0];
  d.b.rot[h] = rec[1];
  d.b.drot[h] = rec[3];
}
__global__ void k_sb_unpack(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  sb_unpack_body(d, h);
}

// ---------------------------------------------------------------------------
// primal: solve_body (assembly.rs:73-903) for the bodies of one colour
// ---------------------------------------------------------------------------

// Keeps a loaded value in registers at this point: the compiler may neither
// sink its load past the point nor re-load it later.
__device__ __forceinline__ void pin4(float4& v) { asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w)); }

// A constraint row as the accumulation sees it: a full stamp (stamp_row:
// axis nn, angular lever t, penalty, force), an angular couple
// (rolling/torsion: axis g, ROLL_PEN, force) or a diagonal add to the
// angular block (geometric stiffness).
//
// x holds the row's gradient (linear part, angular part).  The kind is
// carried by the sign of pen (penalties are >= 1, damper stiffnesses > 0):
//   pen > 0  stamp_row: x = (nn, t = r x nn); lower-triangle entry (i, j)
//            receives pen * (x[i] * x[j]), rhs k receives f * x[k];
//   pen < 0  angular couple: x = (0, g); only angular entries, stiffness -pen;
//   pen == 0 geometric stiffness: x[3..5] added to the angular diagonal.
struct Row {
  float x[6];
  float pen;
  float f;
};
static_assert(sizeof(Row) == 32, "rows are read as 8 floats");

__device__ __forceinline__ void make_stamp(Row& r, f3 axis, float pen, float f, f3 lever, float sign) {
  f3 nn = smul(sign, axis);
  f3 t = cross3(lever, nn);
  r.x[0] = nn.x; r.x[1] = nn.y; r.x[2] = nn.z;
  r.x[3] = t.x; r.x[4] = t.y; r.x[5] = t.z;
  r.pen = pen;
  r.f = f;
}

// Rows of one contact point-side of body h.  Row slots in the reference's
// row order: 0 normal (or the damper), 1-2 friction, 3-5 rolling/torsion
// (only when R == 6); returns the mask of slots written.
// `own` is the solver record (pos, rot, pos0, drot) of the body being
// solved, loaded once per body.  Each live row is handed to `emit(kind,
// row)` as soon as it is formed, in row order (kind 0: stamp, 1: angular
// couple), so at most one row is held in registers.
// Partner record of a point item: the current solver record by default.
struct SbPartner {
  const float4* sb;
  __device__ __forceinline__ const float4* operator()(uint32_t h) const { return sb + 4 * (size_t)h; }
};
// The three 64-byte records a point item reads: its pair, its point and
// the partner body's solver record.
struct ItemRec {
  float4 n4, t14, t24, mu;    // pair: normal + gap, t1 + damper, t2, friction
  float4 ra4, rb4, lam, pen;  // point
  float4 o0, o1, o2, o3;      // partner: pos, rot, pos0, drot
};
template <class Partner>
__device__ __forceinline__ const float4* item_partner(const Partner& partner, uint4 it) {
  return partner((it.x >> 31) == 0u ? it.w : it.z);
}

// The rows of a point item from its records (the arithmetic of point_rows).
template <int R, class Emit>
__device__ __forceinline__ void item_rows(const Dev& d, uint4 it, const ItemRec& r, const float4 (&own)[4],
                                          Emit&& emit) {
  uint32_t q = it.x & 0x7fffffffu;
  bool is_a = (it.x >> 31) == 0u;
  float sign = is_a ? 1.0f : -1.0f;
  const float4 n4 = r.n4, t14 = r.t14, t24 = r.t24, mu = r.mu;
  const float4 ra4 = r.ra4, rb4 = r.rb4, lam = r.lam, pen = r.pen;
  const float4 o0 = r.o0, o1 = r.o1, o2 = r.o2, o3 = r.o3;
  const float4 pa4 = is_a ? own[0] : o0, sa1 = is_a ? own[1] : o1, a0 = is_a ? own[2] : o2,
               sa3 = is_a ? own[3] : o3;
  const float4 pb4 = is_a ? o0 : own[0], sb1 = is_a ? o1 : own[1], b0 = is_a ? o2 : own[2],
               sb3 = is_a ? o3 : own[3];
  f3 n = xyz(n4);
  float gap = n4.w;
  bool touching = gap <= 0.0f;
  float c0 = touching ? ra4.w : 0.0f;
  qf qa = asq(sa1), qb = asq(sb1);
  f3 dra = xyz(sa3), drb = xyz(sb3);
  f3 dpa = sub(xyz(pa4), xyz(a0)), dpb = sub(xyz(pb4), xyz(b0));
  // contact_levers (rows.rs:963-973)
  f3 ra_w = a0.w >= 0.0f ? mul(neg(n), a0.w) : qrot(qa, xyz(ra4));
  f3 rb_w = b0.w >= 0.0f ? mul(n, b0.w) : qrot(qb, xyz(rb4));
  // row_c_levers (rows.rs:1006-1025)
  auto row_c = [&](f3 u, float cc, f3 la, f3 lb) {
    f3 d_a = add(dpa, mul(u, dot3(cross3(la, u), dra)));
    f3 d_b = add(dpb, mul(u, dot3(cross3(lb, u), drb)));
    return cc * (1.0f - K_ALPHA) + dot3(u, sub(d_a, d_b));
  };
  float cn = row_c(n, c0, ra_w, rb_w);
  f3 r_side = is_a ? ra_w : rb_w;
  Row row;
  if (!touching) {
    float damp = t14.w;
    if (damp > 0.0f && fabsf(cn) >= K_C_EPS) {
      float f_d = fmin_(damp * cn, 0.0f);
      make_stamp(row, n, damp, f_d, r_side, sign);
      emit(0, row);
    }
    return;
  }
  f3 t1 = xyz(t14), t2 = xyz(t24);
  if (t24.w != 0.0f) {  // the frame from the live orientations (assembly.rs:176-181)
    t1 = friction_t1(d, it.z, it.w, qa, qb, n);
    t2 = cross3(t1, n);
  }
  // friction_levers (rows.rs:987-996)
  float sep = dot3(sub(add(xyz(pa4), ra_w), add(xyz(pb4), rb_w)), n);
  f3 ra_f = sub(ra_w, mul(n, sep * 0.5f));
  f3 rb_f = add(rb_w, mul(n, sep * 0.5f));
  float ct1 = row_c(t1, 0.0f, ra_f, rb_f);
  float ct2 = row_c(t2, 0.0f, ra_f, rb_f);
  // contact_force (rows.rs:1032-1060)
  float fn_c = fmin_(pen.x * cn + lam.x, 0.0f);
  float r0 = pen.y * ct1 + lam.y, r1 = pen.z * ct2 + lam.z;
  float b1 = fabsf(fn_c) * mu.x, b2 = fabsf(fn_c) * mu.y;
  float f1, f2;
  if (b1 <= 0.0f || b2 <= 0.0f) {
    f1 = b1 <= 0.0f ? 0.0f : clampf_(r0, -b1, b1);
    f2 = b2 <= 0.0f ? 0.0f : clampf_(r1, -b2, b2);
  } else {
    float s = (r0 / b1) * (r0 / b1) + (r1 / b2) * (r1 / b2);
    float k = s > 1.0f ? 1.0f / sqrtf(s) : 1.0f;
    f1 = r0 * k;
    f2 = r1 * k;
  }
  f3 r_side_f = is_a ? ra_f : rb_f;
  if (fabsf(cn) >= K_C_EPS || fn_c != 0.0f) {
    make_stamp(row, n, pen.x, fn_c, r_side, sign);
    emit(0, row);
  }
  if (fabsf(ct1) >= K_C_EPS || f1 != 0.0f) {
    make_stamp(row, t1, pen.y, f1, r_side_f, sign);
    emit(0, row);
  }
  if (fabsf(ct2) >= K_C_EPS || f2 != 0.0f) {
    make_stamp(row, t2, pen.z, f2, r_side_f, sign);
    emit(0, row);
  }
  if constexpr (R >= 6) {
    if (!(mu.z > 0.0f || mu.w > 0.0f)) return;
    f3 wrel = sub(dra, drb);
    float4 rl = d.Q.roll[q];
    f3 ax[3] = {t1, t2, n};
    float mur[3] = {mu.z, mu.z, mu.w};
    float rlv[3] = {rl.x, rl.y, rl.z};
#pragma unroll
    for (int li = 0; li < 3; ++li) {
      float c = dot3(wrel, ax[li]);
      float cap = mur[li] * fabsf(fn_c);
      float fr = clampf_(K_ROLL_PEN * c + rlv[li], -cap, cap);
      f3 g = is_a ? ax[li] : neg(ax[li]);
      row.x[0] = 0.0f; row.x[1] = 0.0f; row.x[2] = 0.0f;
      row.x[3] = g.x; row.x[4] = g.y; row.x[5] = g.z;
      row.pen = -K_ROLL_PEN;
      row.f = fr;
      if (!(mur[li] <= 0.0f)) emit(1, row);
    }
  }
}

template <int R, class Emit, class Partner>
__device__ __forceinline__ void point_rows(const Dev& d, uint4 it, const float4 (&own)[4], Emit&& emit,
                                           const Partner& partner) {
  // The three records are loaded together and pinned before any branch, so
  // the item costs one memory round trip instead of one per branch that
  // first uses a field.
  const float4* prec = d.P.rec + 4 * (size_t)it.y;
  const float4* qrec = d.Q.rec + 4 * (size_t)(it.x & 0x7fffffffu);
  const float4* so = item_partner(partner, it);
  ItemRec r;
  r.n4 = prec[0]; r.t14 = prec[1]; r.t24 = prec[2]; r.mu = prec[3];
  r.ra4 = qrec[0]; r.rb4 = qrec[1]; r.lam = qrec[2]; r.pen = qrec[3];
  r.o0 = so[0]; r.o1 = so[1]; r.o2 = so[2]; r.o3 = so[3];
  pin4(r.n4); pin4(r.t14); pin4(r.t24); pin4(r.mu);
  pin4(r.ra4); pin4(r.rb4); pin4(r.lam); pin4(r.pen);
  pin4(r.o0); pin4(r.o1); pin4(r.o2); pin4(r.o3);
  item_rows<R>(d, it, r, own, emit);
}

// A SixDof joint's axis k of A's step-start or live frame (linear: the
// drift from the assembly delta along it; angular: the twist about it
// less the assembly twist), and its window (assembly.rs:452-519, :682-740).
__device__ __forceinline__ float six_lin_value(f3 pa, f3 pb, qf qa, f3 dref, f3 dir) {
  return dot3(sub(sub(pb, pa), qrot(qa, dref)), dir);
}
__device__ __forceinline__ float six_ang_value(qf qa, qf qb, qf qref, f3 e) {
  return hinge_twist(qa, qb, e) - quat_twist(qref, e);
}
__device__ __forceinline__ f3 unit_axis(int i) {
  return f3{i == 0 ? 1.0f : 0.0f, i == 1 ? 1.0f : 0.0f, i == 2 ? 1.0f : 0.0f};
}

// Rows of one joint side, in the reference's order (assembly.rs:253-818):
// the linear rows (Ball / Revolute / Fixed: three world-axis rows at the
// anchors, Fixed offset by the assembly delta; Prismatic: two rows across
// the live slide axis), the geometric-stiffness diagonal, then the angular
// rows (Revolute / Prismatic: two rows aligning the axes; Fixed: three rows
// on the relative-rotation error about A's axes).  Linear rows are stamps
// (kind 0), the diagonal kind 2, angular rows couples (kind 1: x = (0, g),
// its pen holding -K as the rolling rows do).
template <class Emit>
__device__ __forceinline__ void joint_rows(const Dev& d, uint32_t e, Emit&& emit, uint32_t h) {
  uint32_t j = e >> 1;
  bool is_a = (e & 1u) == 0u;
  float sign = is_a ? 1.0f : -1.0f;
  const int kind = d.J.kind[j];
  if (kind == JK_GEAR) {
    // Gear (assembly.rs:426-449, :650-679): C = ca + ratio cb - const over
    // the coupled joints' bodies; body h's slide sides are stamped, then
    // its hinge sides coupled
    const PoseRecords pose{d.sb};
    const GearSide sd[2] = {gear_side(d, j, 0, pose), gear_side(d, j, 1, pose)};
    const float4 g = d.J.dref[j];
    const float c = sd[0].coord + g.x * sd[1].coord - g.y;
    const float pn = d.J.pen[j].x;
    const float f = pn * c + d.J.lam[j].x;
    if (!(fabsf(c) >= K_C_EPS || f != 0.0f)) return;
    const float coef[2] = {1.0f, g.x};
    Row row;
    for (int k = 0; k < 2; ++k)
      if (!sd[k].ang && (h == sd[k].a || h == sd[k].b)) {
        make_stamp(row, sd[k].axis, pn * coef[k] * coef[k], f * coef[k], h == sd[k].a ? sd[k].ra : sd[k].rb,
                   h == sd[k].a ? -1.0f : 1.0f);
        emit(0, row);
      }
    for (int k = 0; k < 2; ++k)
      if (sd[k].ang && (h == sd[k].a || h == sd[k].b)) {
        const f3 gv = h == sd[k].a ? neg(sd[k].axis) : sd[k].axis;
        row.x[0] = 0.0f; row.x[1] = 0.0f; row.x[2] = 0.0f;
        row.x[3] = gv.x; row.x[4] = gv.y; row.x[5] = gv.z;
        row.pen = -(pn * coef[k] * coef[k]);
        row.f = f * coef[k];
        emit(1, row);
      }
    return;
  }
  uint2 ab = d.J.ab[j];
  const float4* sa = d.sb + 4 * (size_t)ab.x;
  const float4* sbb = d.sb + 4 * (size_t)ab.y;
  qf qa = asq(sa[1]), qb = asq(sbb[1]);
  f3 la = xyz(d.J.la[j]), lb = xyz(d.J.lb[j]);
  f3 pa = add(xyz(sa[0]), qrot(qa, la));
  f3 pb = add(xyz(sbb[0]), qrot(qb, lb));
  f3 live = sub(pa, pb);
  f3 c0v = xyz(d.J.c0v[j]);
  float4 lam = d.J.lam[j], pen = d.J.pen[j];
  float lamv[3] = {lam.x, lam.y, lam.z}, penv[3] = {pen.x, pen.y, pen.z};
  f3 r_a = qrot(qa, xyz(d.J.la0[j]));
  f3 r_b = qrot(qb, lb);
  float fl[3] = {0.0f, 0.0f, 0.0f};
  Row row;
  if (kind == JK_MOTOR) {
    // free drive: no position rows (its impulse ran at the step start)
  } else if (is_tether(kind)) {
    // Distance / Rope / Spring: one row along the live anchor delta
    // (assembly.rs:328-425)
    const float4 pr = d.J.dref[j];
    const float len = len3(live);
    const f3 dir = len > K_NEAR_ZERO ? divs(live, len) : f3{0.0f, 1.0f, 0.0f};
    const f3 r_side = is_a ? qrot(qa, la) : r_b;
    if (kind != JK_SPRING) {
      if (kind == JK_DISTANCE || len > pr.x) {  // a rope stamps only while stretched
        const float c = (len - pr.x) - K_ALPHA * (len3(c0v) - pr.x);
        const float f = penv[0] * c + lamv[0];
        if (fabsf(c) >= K_C_EPS || f != 0.0f) {
          make_stamp(row, dir, penv[0], f, r_side, sign);
          emit(0, row);
          fl[0] = dir.x * f;
          fl[1] = dir.y * f;
          fl[2] = dir.z * f;
        }
      }
    } else {
      // position-level spring about the rest length: fixed penalty, no dual
      const float sp = len - pr.x;
      const float s0 = len3(c0v) - pr.x;
      const float4 pa4 = d.b.pos[ab.x], pb4 = d.b.pos[ab.y];
      const float ka = d.b.type[ab.x] == BT_DYNAMIC ? pa4.w : 0.0f;
      const float kb = d.b.type[ab.y] == BT_DYNAMIC ? pb4.w : 0.0f;
      const float m = ka + kb > K_NEAR_ZERO ? 1.0f / (ka + kb) : 0.0f;
      if (m <= 0.0f) return;
      const bool force = ((uint32_t)pr.w & SPRING_FORCE_MODEL) != 0u;
      const float stiff = force ? pr.y : pr.y * m;
      const float dampc = force ? pr.z : pr.z * m;
      if (stiff <= 0.0f && dampc <= 0.0f) return;
      if (stiff > 0.0f) {
        make_stamp(row, dir, stiff, stiff * sp + dampc * (sp - s0) / K_DT, r_side, sign);
      } else {  // a pure damper
        make_stamp(row, dir, dampc / K_DT, dampc * (sp - s0) / K_DT, r_side, sign);
      }
      emit(0, row);
    }
  } else if (kind == JK_SIXDOF) {
    // SixDof (assembly.rs:452-519): per axis of A's live frame a locked row
    // on the drift from the assembly delta, or a one-sided row inside its
    // window warmed by the accumulator; B minus A, so the gradient flips
    const float4 w0 = d.J.six[4 * j], w1 = d.J.six[4 * j + 1];
    const float4 s0 = d.J.sacc[2 * j];
    const f3 dr = xyz(d.J.dref[j]);
    const f3 r_side = is_a ? qrot(qa, la) : r_b;
    const float lsign = is_a ? -1.0f : 1.0f;
    const float lo3[3] = {w0.x, w0.y, w0.z}, hi3[3] = {w1.x, w1.y, w1.z}, sa3[3] = {s0.x, s0.y, s0.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int cfg = six_cfg(w0.w, i);
      if (cfg == AX_FREE) continue;
      const f3 e = unit_axis(i);
      const f3 dir = normalize_or(qrot(qa, e), e);
      const float v = six_lin_value(pa, pb, qa, dr, dir);
      float c, f;
      if (cfg == AX_LOCKED) {
        c = v;
        f = penv[i] * c + lamv[i];
      } else {
        const int st = warm_limit_state(v, lo3[i], hi3[i], K_LIMIT_SLOP_LIN, sa3[i]);
        if (st < 0) continue;
        const bool lower = st == 1;
        const qf ra0 = asq(d.b.rot0[ab.x]), rb0 = asq(d.b.rot0[ab.y]);
        const f3 pa0 = add(xyz(d.b.pos0[ab.x]), qrot(ra0, la)), pb0 = add(xyz(d.b.pos0[ab.y]), qrot(rb0, lb));
        const float initial = six_lin_value(pa0, pb0, ra0, dr, normalize_or(qrot(ra0, e), e));
        c = regularized_limit(v, initial, lower ? lo3[i] : hi3[i], lower);
        const float f_raw = penv[i] * c + sa3[i];
        f = lower ? fmin_(f_raw, 0.0f) : fmax_(f_raw, 0.0f);
      }
      if (!(fabsf(c) >= K_C_EPS || f != 0.0f)) continue;
      make_stamp(row, dir, penv[i], f, r_side, lsign);
      emit(0, row);
    }
  } else if (kind == JK_WHEEL) {
    // Wheel (assembly.rs:521-580): two rows across the live suspension
    // axis, then the suspension spring along it (penalty m w^2, damping
    // 2 m zeta w, no dual) or, without a frequency, a rigid row on slot 2
    const float4 pr = d.J.dref[j];
    const f3 wa = normalize_or(qrot(qa, xyz(d.J.sus[j])), f3{0.0f, 0.0f, 1.0f});
    f3 u, v;
    tangent_basis(wa, &u, &v);
    const f3 dl = sub(live, mul(c0v, K_ALPHA));
    const f3 r_side = is_a ? qrot(qa, la) : r_b;
#pragma unroll
    for (int li = 0; li < 2; ++li) {
      const f3 ax = li == 0 ? u : v;
      float c = dot3(dl, ax);
      float f = penv[li] * c + lamv[li];
      if (!(fabsf(c) >= K_C_EPS || f != 0.0f)) continue;
      make_stamp(row, ax, penv[li], f, r_side, sign);
      emit(0, row);
      fl[0] = fl[0] + ax.x * f;
      fl[1] = fl[1] + ax.y * f;
      fl[2] = fl[2] + ax.z * f;
    }
    const float sv = dot3(sub(pb, pa), wa) - pr.x;
    const float lsign = is_a ? -1.0f : 1.0f;  // C = (B - A).wa
    const qf ra0 = asq(d.b.rot0[ab.x]), rb0 = asq(d.b.rot0[ab.y]);
    const f3 wa0 = normalize_or(qrot(ra0, xyz(d.J.sus[j])), f3{0.0f, 0.0f, 1.0f});
    const f3 pa0 = add(xyz(d.b.pos0[ab.x]), qrot(ra0, la)), pb0 = add(xyz(d.b.pos0[ab.y]), qrot(rb0, lb));
    const float s0 = dot3(sub(pb0, pa0), wa0) - pr.x;
    if (pr.y > 0.0f) {
      const float4 pa4 = d.b.pos[ab.x], pb4 = d.b.pos[ab.y];
      const float ka = d.b.type[ab.x] == BT_DYNAMIC ? pa4.w : 0.0f;
      const float kb = d.b.type[ab.y] == BT_DYNAMIC ? pb4.w : 0.0f;
      const float m = ka + kb > K_NEAR_ZERO ? 1.0f / (ka + kb) : 0.0f;
      if (m > 0.0f) {
        const float omega = K_TAU * pr.y;
        const float stiff = m * omega * omega;
        const float dampc = 2.0f * m * pr.z * omega;
        make_stamp(row, wa, stiff, stiff * sv + dampc * (sv - s0) / K_DT, r_side, lsign);
        emit(0, row);
      }
    } else {
      const float c = sv - K_ALPHA * s0;
      const float f = penv[2] * c + lamv[2];
      if (fabsf(c) >= K_C_EPS || f != 0.0f) {
        make_stamp(row, wa, penv[2], f, r_side, lsign);
        emit(0, row);
      }
    }
  } else if (kind != JK_PRISMATIC) {
    if (kind == JK_FIXED) live = add(live, qrot(qa, xyz(d.J.dref[j])));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float c = comp(live, k) - K_ALPHA * comp(c0v, k);
      float f = penv[k] * c + lamv[k];
      if (!(fabsf(c) >= K_C_EPS || f != 0.0f)) continue;
      f3 axis = zero3();
      setc(axis, k, 1.0f);
      make_stamp(row, axis, penv[k], f, is_a ? r_a : r_b, sign);
      emit(0, row);
      fl[k] = f;
    }
  } else {
    const f3 wa = normalize_or(qrot(qa, xyz(d.J.ax[2 * j])), f3{0.0f, 0.0f, 1.0f});
    f3 u, v;
    tangent_basis(wa, &u, &v);
    const f3 dl = sub(live, mul(c0v, K_ALPHA));
    const f3 r_side = is_a ? qrot(qa, la) : r_b;
#pragma unroll
    for (int li = 0; li < 2; ++li) {
      const f3 ax = li == 0 ? u : v;
      float c = dot3(dl, ax);
      float f = penv[li] * c + lamv[li];
      if (!(fabsf(c) >= K_C_EPS || f != 0.0f)) continue;
      make_stamp(row, ax, penv[li], f, r_side, sign);
      emit(0, row);
      fl[0] = fl[0] + ax.x * f;
      fl[1] = fl[1] + ax.y * f;
      fl[2] = fl[2] + ax.z * f;
    }
  }
  f3 r = is_a ? r_a : neg(r_b);
  float hm[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
  for (int k = 0; k < 3; ++k) {
    float g[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    float rk = comp(r, k);
    g[0][0] = -rk;
    g[1][1] = -rk;
    g[2][2] = -rk;
    g[0][k] += r.x;
    g[1][k] += r.y;
    g[2][k] += r.z;
    for (int x = 0; x < 3; ++x)
      for (int y = 0; y < 3; ++y) hm[x][y] += g[x][y] * fl[k];
  }
  row.x[0] = 0.0f; row.x[1] = 0.0f; row.x[2] = 0.0f;
  row.x[3] = len3(f3{hm[0][0], hm[1][0], hm[2][0]});
  row.x[4] = len3(f3{hm[0][1], hm[1][1], hm[2][1]});
  row.x[5] = len3(f3{hm[0][2], hm[1][2], hm[2][2]});
  row.pen = 0.0f;
  row.f = 0.0f;
  emit(2, row);
  if (kind == JK_BALL || is_tether(kind) || kind == JK_MOTOR) return;
  const float4 lama = d.J.lama[j], pena = d.J.pena[j];
  const float lav[3] = {lama.x, lama.y, lama.z}, pav[3] = {pena.x, pena.y, pena.z};
  auto angular = [&](f3 g, float pn, float f) {
    row.x[0] = 0.0f; row.x[1] = 0.0f; row.x[2] = 0.0f;
    row.x[3] = g.x; row.x[4] = g.y; row.x[5] = g.z;
    row.pen = -pn;
    row.f = f;
    emit(1, row);
  };
  if (kind == JK_SIXDOF) {
    // angular rows about A's live axes (assembly.rs:682-740): locked on the
    // relative-rotation error, limited on the twist window
    const float4 w0 = d.J.six[4 * j], w2 = d.J.six[4 * j + 2], w3 = d.J.six[4 * j + 3];
    const float4 s1 = d.J.sacc[2 * j + 1];
    const qf qr = asq(d.J.qref[j]);
    const f3 diff = quat_diff_vec(qmul(qconj(qa), qb), qr);
    const float lo3[3] = {w2.x, w2.y, w2.z}, hi3[3] = {w3.x, w3.y, w3.z}, sa3[3] = {s1.x, s1.y, s1.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int cfg = six_cfg(w0.w, 3 + i);
      if (cfg == AX_FREE) continue;
      const f3 e = unit_axis(i);
      const f3 dir = normalize_or(qrot(qa, e), e);
      float c, f;
      if (cfg == AX_LOCKED) {
        c = dot3(diff, e);
        f = pav[i] * c + lav[i];
      } else {
        const float v = six_ang_value(qa, qb, qr, e);
        const int st = warm_limit_state(v, lo3[i], hi3[i], K_LIMIT_SLOP_ANG, sa3[i]);
        if (st < 0) continue;
        const bool lower = st == 1;
        const float initial = six_ang_value(asq(d.b.rot0[ab.x]), asq(d.b.rot0[ab.y]), qr, e);
        c = regularized_limit(v, initial, lower ? lo3[i] : hi3[i], lower);
        const float f_raw = pav[i] * c + sa3[i];
        f = lower ? fmin_(f_raw, 0.0f) : fmax_(f_raw, 0.0f);
      }
      if (!(fabsf(c) >= K_C_EPS || f != 0.0f)) continue;
      angular(is_a ? neg(dir) : dir, pav[i], f);
    }
    return;
  }
  if (kind == JK_FIXED) {
    const f3 diff = quat_diff_vec(qmul(qconj(qa), qb), asq(d.J.qref[j]));
    const f3 diff0 = xyz(d.J.a0[2 * j]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float c = comp(diff, k) - K_ALPHA * comp(diff0, k);
      float f = pav[k] * c + lav[k];
      if (!(fabsf(c) >= K_C_EPS || f != 0.0f)) continue;
      f3 ek = zero3();
      setc(ek, k, 1.0f);
      const f3 dir = qrot(qa, ek);
      angular(is_a ? neg(dir) : dir, pav[k], f);
    }
    return;
  }
  const f3 axa = qrot(qa, xyz(d.J.ax[2 * j])), axb = qrot(qb, xyz(d.J.ax[2 * j + 1]));
  const f3 dax = sub(axa, axb);
#pragma unroll
  for (int li = 0; li < 2; ++li) {
    const float4 t4 = d.J.a0[2 * j + li];
    const f3 t = xyz(t4);
    float c = dot3(t, dax) - K_ALPHA * t4.w;
    float f = pav[li] * c + lav[li];
    if (!(fabsf(c) >= K_C_EPS || f != 0.0f)) continue;
    angular(is_a ? cross3(axa, t) : neg(cross3(axb, t)), pav[li], f);
  }
  // the travel limit (assembly.rs:809-874): one-sided, warmed by the dual's
  // accumulator (lama.w revolute, lam.w prismatic), never written here
  const float4 pr = d.J.dref[j];
  if (!((uint32_t)pr.w & HINGE_LIMIT)) return;
  const bool rev = kind == JK_REVOLUTE;
  const f3 wa = normalize_or(axa, f3{0.0f, 0.0f, 1.0f});
  const float v = limit_value(d, j, kind, qa, qb, pa, pb, wa);
  const float accf = rev ? lama.w : lam.w;
  const int st = warm_limit_state(v, pr.y, pr.z, rev ? K_LIMIT_SLOP_ANG : K_LIMIT_SLOP_LIN, accf);
  if (st < 0) return;
  const bool lower = st == 1;
  const float c = regularized_limit(v, d.J.c0v[j].w, lower ? pr.y : pr.z, lower);
  const float lpen = rev ? pav[2] : penv[2];
  const float f_raw = lpen * c + accf;
  const float f = lower ? fmin_(f_raw, 0.0f) : fmax_(f_raw, 0.0f);
  if (!(fabsf(c) >= K_C_EPS || f != 0.0f)) return;
  if (rev) {
    angular(is_a ? neg(wa) : wa, lpen, f);
  } else {
    make_stamp(row, wa, lpen, f, is_a ? qrot(qa, la) : r_b, is_a ? -1.0f : 1.0f);  // C = (B - A).wa
    emit(0, row);
  }
}

// solve_6x6 (assembly.rs:25-69) on the 21-entry lower triangle.
__device__ __forceinline__ int tri(int i, int j) { return i * (i + 1) / 2 + j; }
__device__ bool solve_6x6_dev(const float* L21, const float* rhs, float* x) {
  float l[6][6];
  float dg[6];
  for (int i = 0; i < 6; ++i) {
    for (int j = 0; j <= i; ++j) {
      float s = L21[tri(i, j)];
      for (int k = 0; k < j; ++k) s -= l[i][k] * dg[k] * l[j][k];
      if (i == j) {
        if (s <= K_DEGENERATE) return false;
        dg[i] = s;
        l[i][i] = 1.0f;
      } else {
        l[i][j] = s / dg[j];
      }
    }
  }
  float y[6];
  for (int i = 0; i < 6; ++i) {
    float s = rhs[i];
    for (int k = 0; k < i; ++k) s -= l[i][k] * y[k];
    y[i] = s;
  }
  float z[6];
  for (int i = 0; i < 6; ++i) z[i] = y[i] / dg[i];
  for (int i = 5; i >= 0; --i) {
    float s = z[i];
    for (int k = i + 1; k < 6; ++k) s -= l[k][i] * x[k];
    x[i] = s;
  }
  return true;
}

// Lumped inertial Hessian diagonal [m/dt^2 x 3, I/dt^2 x 3] of a GpuAvbdMass
// {inertia[3], inv_mass} (avbd_inertial_hessian_diag, crates/physics/src/
// gpu/avbd.rs:88-107); statics, non-finite inputs and degenerate dt give zeros.
__device__ __forceinline__ void lumped_diag(float4 m, float dt, float dg[6]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) dg[k] = 0.0f;
  if (isfinite(m.w) && m.w > 0.0f && isfinite(dt) && dt > 0.0f) {
    const float dt2 = dt * dt;
    if (isfinite(dt2) && !(dt2 <= K_DEGENERATE)) {
      const float l = 1.0f / (m.w * dt2);
      dg[0] = l;
      dg[1] = l;
      dg[2] = l;
      dg[3] = m.x / dt2;
      dg[4] = m.y / dt2;
      dg[5] = m.z / dt2;
    }
  }
}

// Rung-1 staged solve (GpuAvbdStub::stage_diag_solve, crates/physics/src/gpu/
// avbd.rs:207-226), one thread per body: the lumped diagonal of its
// GpuAvbdMass record and the diagonal solve (avbd_diag_solve_cpu, :117-128):
// x = rhs / diag on axes above the degenerate threshold, 0 elsewhere.  rhs and
// delta are GpuBodyState records {lin[3], 0, ang[3], 0}.
__global__ void k_solve_diag(const float4* __restrict__ mass, const float4* __restrict__ rhs, uint32_t n, float dt,
                             float4* __restrict__ delta) {
  uint32_t i = gid();
  if (i >= n) return;
  float dg[6];
  lumped_diag(mass[i], dt, dg);
  const float4 lin = rhs[2 * (size_t)i], ang = rhs[2 * (size_t)i + 1];
  const float r[6] = {lin.x, lin.y, lin.z, ang.x, ang.y, ang.z};
  float x[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) x[k] = dg[k] > K_DEGENERATE ? r[k] / dg[k] : 0.0f;
  delta[2 * (size_t)i] = make_float4(x[0], x[1], x[2], 0.0f);
  delta[2 * (size_t)i + 1] = make_float4(x[3], x[4], x[5], 0.0f);
}

// Rung-2 staged systems (avbd_stage_contact_solve, crates/physics/src/gpu/
// avbd.rs:415-449), one thread per 96-byte GpuAvbdSystem record: the lumped
// inertial Hessian diagonal, one stamped row (stamp_row's order: every entry
// receives one product on top of a zero or the diagonal) and solve_6x6; a
// breakdown gives a zero delta and ok = 0.  delta holds GpuBodyState records
// {lin[3], 0, ang[3], 0}.
__global__ void k_solve_systems(const float4* __restrict__ sys, uint32_t n, float dt, float4* __restrict__ delta,
                                float* __restrict__ ok) {
  uint32_t i = gid();
  if (i >= n) return;
  const float4* s = sys + 6 * (size_t)i;
  const float4 m = s[0], lin = s[1], ang = s[2], ax = s[3], lv = s[4], sg = s[5];
  float dg[6];
  lumped_diag(m, dt, dg);
  const f3 nn = smul(sg.x, xyz(ax));
  const f3 t = cross3(xyz(lv), nn);
  const float pen = ax.w, f = lv.w;
  const float v[6] = {nn.x, nn.y, nn.z, t.x, t.y, t.z};
  float L21[21], rhs[6] = {lin.x, lin.y, lin.z, ang.x, ang.y, ang.z};
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int c = 0; c <= r; ++c) {
      // lhs[r][c] += pen * (a * b) with the factor order stamp_row uses for
      // each block: nn nn, t t, and for the lower-left block nn[c] t[r - 3]
      const float prod = (r >= 3 && c < 3) ? pen * (v[c] * v[r]) : pen * (v[r] * v[c]);
      L21[tri(r, c)] = (r == c ? dg[r] : 0.0f) + prod;
    }
#pragma unroll
  for (int k = 0; k < 6; ++k) rhs[k] += f * v[k];
  float x[6];
  const bool solved = solve_6x6_dev(L21, rhs, x);
  delta[2 * (size_t)i] = solved ? make_float4(x[0], x[1], x[2], 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  delta[2 * (size_t)i + 1] = solved ? make_float4(x[3], x[4], x[5], 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ok[i] = solved ? 1.0f : 0.0f;
}

// Inputs of the inertial block of body h: its solver record (pos, rot), its
// body-frame inertia and its inertial targets.
struct InertialIn {
  float4 p, q, inertia, target, target_rot;
};
__device__ __forceinline__ InertialIn load_inertial(const Dev& d, uint32_t h) {
  InertialIn in;
  in.p = d.sb[4 * (size_t)h];
  in.q = d.sb[4 * (size_t)h + 1];
  in.inertia = d.b.inertia[h];
  in.target = d.b.inertial[h];
  in.target_rot = d.b.inertial_rot[h];
  return in;
}

// Inertial block and rhs of a body (assembly.rs:74-96): out[0..20] the
// lower triangle (m/dt^2 on the linear diagonal, I_world/dt^2 in the angular
// block), out[21..26] the rhs (m/dt^2 (x - x~), I_world dq / dt^2), out[27] 0.
__device__ __forceinline__ void inertial_block(const InertialIn& in, float out[28]) {
  const float m_dt2 = 1.0f / in.p.w / (K_DT * K_DT);
  const qf q = asq(in.q);
  float iw[3][3];
  world_inertia(xyz(in.inertia), q, iw);
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j)
      out[tri(i, j)] = i < 3 ? (i == j ? m_dt2 : 0.0f) : (j < 3 ? 0.0f : iw[i - 3][j - 3] / (K_DT * K_DT));
  const f3 rl = smul(m_dt2, sub(xyz(in.p), xyz(in.target)));
  const f3 dv = quat_diff_vec(q, asq(in.target_rot));
  const f3 ra = divs(mat3_vec(iw, dv), K_DT * K_DT);
  out[21] = rl.x;
  out[22] = rl.y;
  out[23] = rl.z;
  out[24] = ra.x;
  out[25] = ra.y;
  out[26] = ra.z;
  out[27] = 0.0f;
}

#define WAVE 64

// Primal solve of one colour class (solve_body, assembly.rs:73-903), sixteen
// lanes per body: a wavefront holds four bodies, a workgroup W wavefronts.
// Lane s of a body owns slot s: it evaluates the body's items s, s + 16,
// s + 32, ... (point-sides in pair order, then joint sides; the three
// 64-byte records an item reads are loaded together, one memory round trip
// per item) and adds their rows, in item order, to 27 partial sums in
// registers (the lower triangle of the 6x6 system, then the rhs), starting
// at +0.  A row stamp adds pen * (x_i * x_j) to entry (i, j) and f * x_k to
// rhs k (stamp_row's products); a rolling / torsion couple touches the
// angular block and rhs only; a geometric-stiffness row adds to the angular
// diagonal.  The sixteen slots are then added pairwise (tree16) with DPP
// exchanges (every lane of the body ends with the same bits), the total is
// added to the body's inertial block
// (evaluated once per iteration beside the dual update), and the 6x6 LDL^T
// solve runs on every lane of the body; slot 0 writes the pose.  The CPU
// oracle's colour order performs exactly these operations.  R = row slots
// per point item (4: contact rows; 6: with rolling / torsion).
constexpr int BODY_LANES = 16;

__device__ __forceinline__ void acc_stamp(float (&acc)[27], const Row& r) {
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) acc[tri(i, j)] += r.pen * (r.x[i] * r.x[j]);
#pragma unroll
  for (int k = 0; k < 6; ++k) acc[21 + k] += r.f * r.x[k];
}
// Angular couple: x = (0, g), stiffness -pen.
__device__ __forceinline__ void acc_couple(float (&acc)[27], const Row& r) {
  const float pen = -r.pen;
#pragma unroll
  for (int i = 3; i < 6; ++i)
#pragma unroll
    for (int j = 3; j <= i; ++j) acc[tri(i, j)] += pen * (r.x[i] * r.x[j]);
#pragma unroll
  for (int k = 3; k < 6; ++k) acc[21 + k] += r.f * r.x[k];
}
// Geometric-stiffness diagonal x[3..5].
__device__ __forceinline__ void acc_diag(float (&acc)[27], const Row& r) {
  acc[tri(3, 3)] += r.x[3];
  acc[tri(4, 4)] += r.x[4];
  acc[tri(5, 5)] += r.x[5];
}

// Pairwise tree over the 16 lanes of a DPP row: v + the value of lane
// (this ^ 1), then ^ 2 within each quad, then the mirrored lane of the
// 8-lane half row, then the mirrored lane of the row.  After each step the
// lanes of a group hold identical bits (a + b == b + a), so the mirrors pair
// whole groups: the result is ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 +
// s7)), and the same for s8-s15, added, in every lane.
__device__ __forceinline__ float tree16(float v) {
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));  // quad_perm 1,0,3,2
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));  // quad_perm 2,3,0,1
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));  // row_half_mirror
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false));  // row_mirror
  return v;
}

// The three records of a point item, loaded without waiting for them.
__device__ __forceinline__ ItemRec load_item(const Dev& d, uint4 it) {
  const float4* prec = d.P.rec + 4 * (size_t)it.y;
  const float4* qrec = d.Q.rec + 4 * (size_t)(it.x & 0x7fffffffu);
  const float4* so = item_partner(SbPartner{d.sb}, it);
  ItemRec r;
  r.n4 = prec[0]; r.t14 = prec[1]; r.t24 = prec[2]; r.mu = prec[3];
  r.ra4 = qrec[0]; r.rb4 = qrec[1]; r.lam = qrec[2]; r.pen = qrec[3];
  r.o0 = so[0]; r.o1 = so[1]; r.o2 = so[2]; r.o3 = so[3];
  return r;
}

// Partner of a point item and of a joint side (the body whose solver
// record the item reads besides the body's own).
__device__ __forceinline__ uint32_t item_partner_id(uint4 it) { return (it.x >> 31) == 0u ? it.w : it.z; }
__device__ __forceinline__ uint32_t joint_partner_id(const Dev& d, uint32_t e) {
  const uint2 ab = d.J.ab[e >> 1];
  return (e & 1u) == 0u ? ab.y : ab.x;
}

// One body's solve on sixteen lanes (slot = lane within the body's lane
// group; lanes with valid == false hold no body and store nothing).  `cr` is
// the body's colour-list record: body (bit 31: has joint sides), first
// incidence item, item count, first joint side.  A lane calls wait(p) before
// it reads the record of partner p.
// PF = 1: the records of a slot's next item are requested before its
// current item is evaluated (one item of records in flight behind the
// arithmetic, 48 more registers); PF = 0: only the next descriptor.
// JT = false compiles the joint rows out (a context without joints): their
// code would otherwise set the register budget of every launch.
template <int R, int PF, bool JT = true, class Wait>
__device__ __forceinline__ void primal_body(const Dev& d, uint4 cr, bool valid, int slot, Wait&& wait) {
  uint32_t h = 0, i0 = 0, np = 0, j0 = 0, nj = 0;
  if (valid) {
    h = cr.x & 0x7fffffffu;
    i0 = cr.y;
    np = cr.z;
    j0 = cr.w;
    if (cr.x >> 31) nj = d.b.ji_off[h + 1] - j0;
  }
  float acc[27];
#pragma unroll
  for (int e = 0; e < 27; ++e) acc[e] = 0.0f;
  float4 own[4];
  {
    const float4* rec = d.sb + 4 * (size_t)h;
    own[0] = rec[0];
    own[1] = rec[1];
    own[2] = rec[2];
    own[3] = rec[3];
  }
  // the inertial block (evaluated beside the previous dual update); with
  // PF = 1 it is requested before the items, else after them
  const float4* inb = reinterpret_cast<const float4*>(d.inb + 28 * (size_t)h);
  float4 ib[7];
  if constexpr (PF == 1) {
#pragma unroll
    for (int q = 0; q < 7; ++q) ib[q] = inb[q];
  }
  auto add_row = [&](int kind, const Row& row) {
    if (kind == 0)
      acc_stamp(acc, row);
    else if (kind == 1)
      acc_couple(acc, row);
    else
      acc_diag(acc, row);
  };
  if constexpr (PF == 0) {
    // the next item's descriptor is fetched while the current item is
    // evaluated, so an item waits for one memory round trip, not two
    uint4 next = slot < np ? d.inc[i0 + slot] : make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t k = slot; k < np; k += BODY_LANES) {
      const uint4 it = next;
      if (k + BODY_LANES < np) next = d.inc[i0 + k + BODY_LANES];
      wait(item_partner_id(it));
      point_rows<R>(d, it, own, add_row, SbPartner{d.sb});
    }
  } else if ((uint32_t)slot < np) {
    // The loads run one item ahead and are unconditional (past the slot's
    // last item they re-read the body's last item), so no branch joins on
    // a load in flight and the next item's records arrive while the
    // current item is evaluated.
    const uint32_t last = i0 + np - 1u;
    uint4 cur = d.inc[i0 + slot];
    wait(item_partner_id(cur));
    ItemRec rc = load_item(d, cur);
    uint4 nxt = d.inc[min(i0 + slot + BODY_LANES, last)];
    for (uint32_t k = slot; k < np; k += BODY_LANES) {
      wait(item_partner_id(nxt));
      const ItemRec rn = load_item(d, nxt);
      const uint4 nn = d.inc[min(i0 + k + 2 * BODY_LANES, last)];
      item_rows<R>(d, cur, rc, own, add_row);
      cur = nxt;
      rc = rn;
      nxt = nn;
    }
  }
  // joint sides continue the item numbering after the point-sides
  if constexpr (JT) {
    const uint32_t total = np + nj;
    for (uint32_t k = np + (uint32_t)((slot - (int)(np % BODY_LANES) + BODY_LANES) % BODY_LANES); k < total;
         k += BODY_LANES) {
      const uint32_t e = d.b.ji[j0 + (k - np)];
      if (d.J.kind[e >> 1] == JK_GEAR) {
        uint32_t m[4];
        joint_members_dev(d, e >> 1, m);
        for (int i = 0; i < 4; ++i)
          if (m[i] != h) wait(m[i]);
      } else {
        wait(joint_partner_id(d, e));
      }
      joint_rows(d, e, add_row, h);
    }
  }
  float L21[21], rhs[6];
  if constexpr (PF == 0) {
#pragma unroll
    for (int q = 0; q < 7; ++q) ib[q] = inb[q];
  }
  float blk[28];
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const float4 v = valid ? ib[q] : make_float4(1.0f, 0.0f, 1.0f, 0.0f);
    blk[4 * q] = v.x;
    blk[4 * q + 1] = v.y;
    blk[4 * q + 2] = v.z;
    blk[4 * q + 3] = v.w;
  }
#pragma unroll
  for (int e = 0; e < 21; ++e) L21[e] = blk[e] + tree16(acc[e]);
#pragma unroll
  for (int k = 0; k < 6; ++k) rhs[k] = -(blk[21 + k] + tree16(acc[21 + k]));
  float dx[6];
  const bool solved = solve_6x6_dev(L21, rhs, dx);
  if (!valid || slot != 0) return;
  if (solved) {
    float4* rec = d.sb + 4 * (size_t)h;
    float4 p = rec[0];
    rec[0] = make_float4(p.x + dx[0], p.y + dx[1], p.z + dx[2], p.w);
    qf q = quat_integrate(asq(rec[1]), f3{dx[3], dx[4], dx[5]});
    rec[1] = pkq(q);
    rec[3] = pk(quat_diff_vec(q, asq(d.b.rot0[h])), 0.0f);
  } else {
    atomicAdd(&d.counters[CTR_BREAKDOWN], 1u);
  }
}

struct NoWait {
  __device__ __forceinline__ void operator()(uint32_t) const {}
};

// Primal solve of one colour class: W wavefronts per workgroup, four bodies
// per wavefront.  Lanes past the list take an empty body.
template <int W, int R, int PF = 0, bool JT = true>
__global__ void __launch_bounds__(WAVE* W) k_primal8(Dev d, const uint4* __restrict__ list, uint32_t count) {
  constexpr int NB = (WAVE / BODY_LANES) * W;  // bodies per workgroup
  const int tid = threadIdx.x;
  const uint32_t idx = blockIdx.x * NB + tid / BODY_LANES;
  const bool valid = idx < count;
  const uint4 cr = valid ? list[idx] : make_uint4(0u, 0u, 0u, 0u);
  primal_body<R, PF, JT>(d, cr, valid, tid % BODY_LANES, NoWait{});
}

// The same solve of colour class `k` with the class's range read from the
// device colour offsets (a step whose colour sizes the host has not read):
// workgroups loop over the class's body groups, so any grid covers it.
template <int W, int R, int PF = 0, bool JT = true>
__global__ void __launch_bounds__(WAVE* W) k_primal8_class(Dev d, uint32_t k) {
  constexpr int NB = (WAVE / BODY_LANES) * W;  // bodies per workgroup
  const int tid = threadIdx.x;
  const uint32_t count = d.colour_cnt[k];
  const uint4* __restrict__ list = d.colour_list + d.colour_off[k];
  for (uint32_t g = blockIdx.x; g * NB < count; g += gridDim.x) {
    const uint32_t idx = g * NB + tid / BODY_LANES;
    const bool valid = idx < count;
    const uint4 cr = valid ? list[idx] : make_uint4(0u, 0u, 0u, 0u);
    primal_body<R, PF, JT>(d, cr, valid, tid % BODY_LANES, NoWait{});
  }
}

// The last colour classes of a sweep in one launch.  Each class's bodies
// (its run of the colour list, from the device colour offsets) form
// batches of four, a batch never spanning two classes; batches are numbered
// class by class, and wavefront w of the `nwaves` launched solves batches
// w, w + nwaves, ... in that order, four bodies at once on sixteen lanes
// each, as k_primal8 does.  Before a lane reads the record of a partner
// whose class lies in [kfirst, the body's class) it waits until the
// partner's done word holds `stamp` (the partner was solved in this sweep);
// after a body's solve its slot-0 lane stores `stamp` to the body's done
// word (release; the waits acquire).  A body therefore reads exactly the
// partner records the class-by-class launches give it: lower classes
// already solved, higher ones not yet (they wait for it).  The bodies of a
// batch share a class and so never wait for each other, and every wait is
// for an earlier batch; with all nwaves wavefronts resident (the host sizes
// the grid from the kernel's occupancy) the earliest unfinished batch can
// always proceed.  A wait gives up after TAIL_SPIN_LIMIT polls and raises
// the error word, so every wavefront ends.
constexpr int TAIL_W = 4;  // wavefronts per workgroup
constexpr uint32_t TAIL_SPIN_LIMIT = 1u << 22;
__device__ __forceinline__ void tail_wait(const Dev& d, uint32_t p, int kfirst, int kown, uint32_t stamp) {
  const int cp = d.b.colour[p];
  if (cp < kfirst || cp >= kown) return;
  uint32_t spins = 0;
  while (__hip_atomic_load(d.b.done + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != stamp) {
    if (++spins == TAIL_SPIN_LIMIT) {
      atomicOr(d.tail_err, 1u);
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
template <int R>
__global__ void __launch_bounds__(WAVE* TAIL_W) k_primal_tail(Dev d, uint32_t kfirst, uint32_t ncol,
                                                              uint32_t nbatch, uint32_t nwaves, uint32_t stamp) {
  constexpr uint32_t BPW = WAVE / BODY_LANES;  // bodies per batch (per wavefront)
  const int lane = threadIdx.x % WAVE;
  const uint32_t w = blockIdx.x * TAIL_W + threadIdx.x / WAVE;
  for (uint32_t b = w; b < nbatch; b += nwaves) {
    // the batch's class k: the first whose batches reach past b
    uint32_t k = kfirst, b0 = 0, c0 = 0, cn = 0;
    for (;;) {
      c0 = d.colour_off[k];
      cn = d.colour_off[k + 1] - c0;
      const uint32_t nb = (cn + BPW - 1) / BPW;
      if (b < b0 + nb || k + 1 >= ncol) break;
      b0 += nb;
      ++k;
    }
    const uint32_t j = BPW * (b - b0) + (uint32_t)lane / BODY_LANES;
    const bool valid = j < cn;
    const uint4 cr = valid ? d.colour_list[c0 + j] : make_uint4(0u, 0u, 0u, 0u);
    const int kown = (int)k;
    primal_body<R, 1>(d, cr, valid, lane % BODY_LANES,
                      [&](uint32_t p) { tail_wait(d, p, (int)kfirst, kown, stamp); });
    if (valid && lane % BODY_LANES == 0)
      __hip_atomic_store(d.b.done + (cr.x & 0x7fffffffu), stamp, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Primal solve of a colour class with one body per wavefront, for classes
// too small to hide memory latency with many bodies in flight.  All 64
// lanes first stage the body's item descriptors, then the three records of
// every item (12 float4 each), in LDS, with every load of a phase in flight
// at once: two memory round trips for the whole body instead of one per
// item and slot.  Lanes 0-15 then evaluate the items from LDS exactly as
// k_primal8 does (slot s: items s, s + 16, ..., the same item_rows
// arithmetic, tree16 and solve), so the two kernels give identical bits.
// Items past PRIMAL_LDS_ITEMS are read from memory as in k_primal8.
constexpr int PRIMAL_LDS_ITEMS = 64;
template <int W, int R>
__global__ void __launch_bounds__(WAVE* W) k_primal_lds(Dev d, const uint4* __restrict__ list, uint32_t count) {
  constexpr int NE = PRIMAL_LDS_ITEMS * 12;  // float4 records per body
  __shared__ uint4 s_desc[W][PRIMAL_LDS_ITEMS];
  __shared__ float4 s_rec[W][NE];
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const uint32_t idx = blockIdx.x * W + wv;
  const bool valid = idx < count;
  uint32_t h = 0, i0 = 0, np = 0, j0 = 0, nj = 0;
  if (valid) {
    const uint4 cr = list[idx];
    h = cr.x & 0x7fffffffu;
    i0 = cr.y;
    np = cr.z;
    j0 = cr.w;
    if (cr.x >> 31) nj = d.b.ji_off[h + 1] - j0;
  }
  const uint32_t ns = np < (uint32_t)PRIMAL_LDS_ITEMS ? np : (uint32_t)PRIMAL_LDS_ITEMS;
  if ((uint32_t)lane < ns) s_desc[wv][lane] = d.inc[i0 + lane];
  __syncthreads();
  {
    constexpr int U = NE / WAVE;
    const uint32_t ne = ns * 12u;
    float4 v[U];
    // every lane loads (a lane past the body's records reads pair, point
    // and body 0, which always exist), so all loads are issued together
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t e = (uint32_t)lane + (uint32_t)u * WAVE;
      const bool live = e < ne;
      const uint32_t part = e % 12u;
      const uint4 it = live ? s_desc[wv][e / 12u] : make_uint4(0u, 0u, 0u, 0u);
      const float4* src = part < 4u   ? d.P.rec + 4 * (size_t)it.y + part
                          : part < 8u ? d.Q.rec + 4 * (size_t)(it.x & 0x7fffffffu) + (part - 4u)
                                      : item_partner(SbPartner{d.sb}, it) + (part - 8u);
      v[u] = *src;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t e = (uint32_t)lane + (uint32_t)u * WAVE;
      if (e < ne) s_rec[wv][e] = v[u];
    }
  }
  __syncthreads();
  float acc[27];
#pragma unroll
  for (int e = 0; e < 27; ++e) acc[e] = 0.0f;
  float4 own[4];
  {
    const float4* rec = d.sb + 4 * (size_t)h;
    own[0] = rec[0];
    own[1] = rec[1];
    own[2] = rec[2];
    own[3] = rec[3];
  }
  auto add_row = [&](int kind, const Row& row) {
    if (kind == 0)
      acc_stamp(acc, row);
    else if (kind == 1)
      acc_couple(acc, row);
    else
      acc_diag(acc, row);
  };
  const int slot = lane;  // lanes 16-63 hold no items: their sums stay +0
  if (slot < BODY_LANES) {
    for (uint32_t k = slot; k < np; k += BODY_LANES) {
      if (k < ns) {
        const float4* r4 = &s_rec[wv][12 * k];
        ItemRec r;
        r.n4 = r4[0]; r.t14 = r4[1]; r.t24 = r4[2]; r.mu = r4[3];
        r.ra4 = r4[4]; r.rb4 = r4[5]; r.lam = r4[6]; r.pen = r4[7];
        r.o0 = r4[8]; r.o1 = r4[9]; r.o2 = r4[10]; r.o3 = r4[11];
        item_rows<R>(d, s_desc[wv][k], r, own, add_row);
      } else {
        point_rows<R>(d, d.inc[i0 + k], own, add_row, SbPartner{d.sb});
      }
    }
    const uint32_t total = np + nj;
    for (uint32_t k = np + (uint32_t)((slot - (int)(np % BODY_LANES) + BODY_LANES) % BODY_LANES); k < total;
         k += BODY_LANES)
      joint_rows(d, d.b.ji[j0 + (k - np)], add_row, h);
  }
  float L21[21], rhs[6];
  const float4* inb = reinterpret_cast<const float4*>(d.inb + 28 * (size_t)h);
  float blk[28];
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const float4 v = valid ? inb[q] : make_float4(1.0f, 0.0f, 1.0f, 0.0f);
    blk[4 * q] = v.x;
    blk[4 * q + 1] = v.y;
    blk[4 * q + 2] = v.z;
    blk[4 * q + 3] = v.w;
  }
#pragma unroll
  for (int e = 0; e < 21; ++e) L21[e] = blk[e] + tree16(acc[e]);
#pragma unroll
  for (int k = 0; k < 6; ++k) rhs[k] = -(blk[21 + k] + tree16(acc[21 + k]));
  float dx[6];
  const bool solved = solve_6x6_dev(L21, rhs, dx);
  if (!valid || lane != 0) return;
  if (solved) {
    float4* rec = d.sb + 4 * (size_t)h;
    float4 p = rec[0];
    rec[0] = make_float4(p.x + dx[0], p.y + dx[1], p.z + dx[2], p.w);
    qf q = quat_integrate(asq(rec[1]), f3{dx[3], dx[4], dx[5]});
    rec[1] = pkq(q);
    rec[3] = pk(quat_diff_vec(q, asq(d.b.rot0[h])), 0.0f);
  } else {
    atomicAdd(&d.counters[CTR_BREAKDOWN], 1u);
  }
}

// ---------------------------------------------------------------------------
// dual update (joints_ext.rs:464-598), one thread per point / joint
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dual_point(const Dev& d, uint32_t q) {
  uint32_t p = d.Q.pair[q];
  uint2 ab = d.P.ab[p];
  float4 n4 = d.P.n(p);
  f3 n = xyz(n4);
  bool touching = n4.w <= 0.0f;
  float4 ra4 = d.Q.ra(q), rb4 = d.Q.rb(q);
  float c0 = touching ? ra4.w : 0.0f;
  const float4* sa = d.sb + 4 * (size_t)ab.x;
  const float4* sbb = d.sb + 4 * (size_t)ab.y;
  float4 pa4 = sa[0], pb4 = sbb[0];
  float4 a0 = sa[2], b0 = sbb[2];
  qf qa = asq(sa[1]), qb = asq(sbb[1]);
  f3 dra = xyz(sa[3]), drb = xyz(sbb[3]);
  f3 dpa = sub(xyz(pa4), xyz(a0)), dpb = sub(xyz(pb4), xyz(b0));
  f3 ra_w = a0.w >= 0.0f ? mul(neg(n), a0.w) : qrot(qa, xyz(ra4));
  f3 rb_w = b0.w >= 0.0f ? mul(n, b0.w) : qrot(qb, xyz(rb4));
  auto row_c = [&](f3 u, float cc, f3 la, f3 lb) {
    f3 d_a = add(dpa, mul(u, dot3(cross3(la, u), dra)));
    f3 d_b = add(dpb, mul(u, dot3(cross3(lb, u), drb)));
    return cc * (1.0f - K_ALPHA) + dot3(u, sub(d_a, d_b));
  };
  float cn = row_c(n, c0, ra_w, rb_w);
  const float4 t24 = d.P.t2(p);
  f3 t1 = xyz(d.P.t1(p)), t2 = xyz(t24);
  if (t24.w != 0.0f) {  // joints_ext.rs:470-474, live orientations
    t1 = friction_t1(d, ab.x, ab.y, qa, qb, n);
    t2 = cross3(t1, n);
  }
  float sep = dot3(sub(add(xyz(pa4), ra_w), add(xyz(pb4), rb_w)), n);
  f3 ra_f = sub(ra_w, mul(n, sep * 0.5f));
  f3 rb_f = add(rb_w, mul(n, sep * 0.5f));
  float ct1 = row_c(t1, 0.0f, ra_f, rb_f);
  float ct2 = row_c(t2, 0.0f, ra_f, rb_f);
  float4 pen = d.Q.pen(q), lam = d.Q.lam(q);
  float4 mu = d.P.mu(p);
  float fn_c = fmin_(pen.x * cn + lam.x, 0.0f);
  float r0 = pen.y * ct1 + lam.y, r1 = pen.z * ct2 + lam.z;
  float b1 = fabsf(fn_c) * mu.x, b2 = fabsf(fn_c) * mu.y;
  float f1, f2;
  if (b1 <= 0.0f || b2 <= 0.0f) {
    f1 = b1 <= 0.0f ? 0.0f : clampf_(r0, -b1, b1);
    f2 = b2 <= 0.0f ? 0.0f : clampf_(r1, -b2, b2);
  } else {
    float s = (r0 / b1) * (r0 / b1) + (r1 / b2) * (r1 / b2);
    float k = s > 1.0f ? 1.0f / sqrtf(s) : 1.0f;
    f1 = r0 * k;
    f2 = r1 * k;
  }
  float npn_w = 0.0f;
  {
    float bnd = fabsf(fn_c) * fmax_(mu.x, mu.y);
    float fs = sqrtf(r0 * r0 + r1 * r1);
    bool stuck = fs <= bnd && sqrtf(ct1 * ct1 + ct2 * ct2) < K_STICK;
    npn_w = stuck ? 1.0f : 0.0f;
  }
  float4 nl = lam, npn = pen;
  if (touching && fabsf(cn) >= K_C_EPS) {
    nl.x = fn_c;
    if (fn_c < 0.0f) npn.x = fmin_(pen.x + K_BETA * fabsf(cn), K_PEN_MAX);
  }
  float bb1 = fabsf(fn_c) * mu.x, bb2 = fabsf(fn_c) * mu.y;
  float s_ell = (bb1 > 0.0f && bb2 > 0.0f) ? (r0 / bb1) * (r0 / bb1) + (r1 / bb2) * (r1 / bb2) : __builtin_inff();
  if (touching) {
    nl.y = f1;
    nl.z = f2;
    if (s_ell <= 1.0f) {
      if (fabsf(ct1) >= K_C_EPS) npn.y = fmin_(pen.y + K_BETA * fabsf(ct1), K_PEN_MAX);
      if (fabsf(ct2) >= K_C_EPS) npn.z = fmin_(pen.z + K_BETA * fabsf(ct2), K_PEN_MAX);
    }
  }
  // the duals and penalties (with the stuck flag in pen.w) fill the
  // record's second 32-byte half; the anchors are not rewritten
  npn.w = npn_w;
  d.Q.lam(q) = nl;
  d.Q.pen(q) = npn;
  if (touching && (mu.z > 0.0f || mu.w > 0.0f)) {
    f3 wrel = sub(dra, drb);
    float lam_n = fabsf(nl.x);
    float4 rl = d.Q.roll[q];
    f3 ax[3] = {t1, t2, n};
    float mur[3] = {mu.z, mu.z, mu.w};
    float rlv[3] = {rl.x, rl.y, rl.z};
    for (int li = 0; li < 3; ++li) {
      if (mur[li] <= 0.0f) continue;
      float c = dot3(wrel, ax[li]);
      float cap = mur[li] * lam_n;
      rlv[li] = clampf_(K_ROLL_PEN * c + rlv[li], -cap, cap);
    }
    d.Q.roll[q] = make_float4(rlv[0], rlv[1], rlv[2], 0.0f);
  }
}

// Inertial block (27 values) of every coloured body for the next primal
// sweep.  It depends only on the body's own pose, which changes only when
// the body itself is solved, so it is evaluated once per iteration off the
// sweep's critical path.
__device__ __forceinline__ void store_inertial_block(const Dev& d, uint32_t h) {
  if (d.b.colour[h] < 0) return;
  float v[28];
  inertial_block(load_inertial(d, h), v);
  float4* out = reinterpret_cast<float4*>(d.inb + 28 * (size_t)h);
#pragma unroll
  for (int k = 0; k < 7; ++k) out[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
}

// Dual update of every point (threads below npts) and, when `next` is set,
// the inertial blocks of the next iteration's sweep (threads above).
// A grid-stride loop: the point count is the device's.
__global__ void __launch_bounds__(TPB) k_dual(Dev d, uint32_t next) {
  const uint32_t npts = dev_npts(d);
  const uint32_t total = npts + (next ? d.n : 0u);
  for (uint32_t g = gid(); g < total; g += gridDim.x * TPB) {
    if (g < npts)
      dual_point(d, g);
    else
      store_inertial_block(d, g - npts);
  }
}

__global__ void k_inertial_blocks(Dev d) {
  uint32_t h = gid();
  if (h < d.n) store_inertial_block(d, h);
}

// Joint duals (joints_ext.rs:543-822): the linear rows accumulate
// lam += K c and ramp K by BETA |c|; the angular rows the same with
// BETA_ANG, the hinge rows committing lam = K c + lam.
#define K_BETA_ANG 100.0f
__global__ void k_dual_joint(Dev d) {
  uint32_t j = gid();
  if (j >= d.njoints) return;
  const int kind = d.J.kind[j];
  if (kind == JK_GEAR) {  // the gear pass (joints_ext.rs:897-917)
    const float c = gear_error(d, j, PoseRecords{d.sb});
    if (fabsf(c) >= K_C_EPS) {
      float4 lam = d.J.lam[j], pen = d.J.pen[j];
      lam.x += pen.x * c;
      pen.x = fmin_(pen.x + K_BETA * fabsf(c), K_LIM_PEN_MAX);
      d.J.lam[j] = lam;
      d.J.pen[j] = pen;
    }
    return;
  }
  uint2 ab = d.J.ab[j];
  const float4* sa = d.sb + 4 * (size_t)ab.x;
  const float4* sbb = d.sb + 4 * (size_t)ab.y;
  const qf qa = asq(sa[1]), qb = asq(sbb[1]);
  f3 pa = add(xyz(sa[0]), qrot(qa, xyz(d.J.la[j])));
  f3 pb = add(xyz(sbb[0]), qrot(qb, xyz(d.J.lb[j])));
  f3 live = sub(pa, pb);
  f3 c0v = xyz(d.J.c0v[j]);
  float4 lam = d.J.lam[j], pen = d.J.pen[j];
  float lv[3] = {lam.x, lam.y, lam.z}, pv[3] = {pen.x, pen.y, pen.z};
  if (is_tether(kind) || kind == JK_MOTOR) {
    if (kind == JK_SPRING || kind == JK_MOTOR) return;  // no dual state
    const float ref = d.J.dref[j].x;
    const float len = len3(live);
    if (kind == JK_DISTANCE) {
      const float c = (len - ref) - K_ALPHA * (len3(c0v) - ref);
      if (fabsf(c) >= K_C_EPS) {
        lv[0] += pv[0] * c;
        pv[0] = fmin_(pv[0] + K_BETA * fabsf(c), K_PEN_MAX);
      }
    } else if (len > ref) {  // rope: tension only, capped like a limit row
      const float c = (len - ref) - K_ALPHA * (len3(c0v) - ref);
      if (fabsf(c) >= K_C_EPS) {
        lv[0] = fmax_(lv[0] + pv[0] * c, 0.0f);
        pv[0] = fmin_(pv[0] + K_BETA * fabsf(c), K_LIM_PEN_MAX);
      }
    } else {
      lv[0] = 0.0f;  // slack forgets the force
    }
    d.J.lam[j] = make_float4(lv[0], lv[1], lv[2], 0.0f);
    d.J.pen[j] = make_float4(pv[0], pv[1], pv[2], 0.0f);
    return;
  }
  float4 sacc0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sacc1 = sacc0;
  if (kind == JK_SIXDOF) {
    // per axis: locked accumulates, limited commits into its accumulator
    // (zeroed when clear), penalties ramp capped (joints_ext.rs:646-686)
    sacc0 = d.J.sacc[2 * j];
    sacc1 = d.J.sacc[2 * j + 1];
    const float4 w0 = d.J.six[4 * j], w1 = d.J.six[4 * j + 1];
    const f3 dr = xyz(d.J.dref[j]);
    float sa3[3] = {sacc0.x, sacc0.y, sacc0.z};
    const float lo3[3] = {w0.x, w0.y, w0.z}, hi3[3] = {w1.x, w1.y, w1.z};
    for (int i = 0; i < 3; ++i) {
      const int cfg = six_cfg(w0.w, i);
      if (cfg == AX_FREE) continue;
      const f3 e = unit_axis(i);
      const f3 dir = normalize_or(qrot(qa, e), e);
      const float v = six_lin_value(pa, pb, qa, dr, dir);
      if (cfg == AX_LOCKED) {
        if (fabsf(v) >= K_C_EPS) {
          lv[i] += pv[i] * v;
          pv[i] = fmin_(pv[i] + K_BETA * fabsf(v), K_PEN_MAX);
        }
        continue;
      }
      const int st = warm_limit_state(v, lo3[i], hi3[i], K_LIMIT_SLOP_LIN, sa3[i]);
      if (st < 0) {
        sa3[i] = 0.0f;
        continue;
      }
      const bool lower = st == 1;
      const qf ra0 = asq(d.b.rot0[ab.x]), rb0 = asq(d.b.rot0[ab.y]);
      const f3 pa0 = add(xyz(d.b.pos0[ab.x]), qrot(ra0, xyz(d.J.la[j])));
      const f3 pb0 = add(xyz(d.b.pos0[ab.y]), qrot(rb0, xyz(d.J.lb[j])));
      const float initial = six_lin_value(pa0, pb0, ra0, dr, normalize_or(qrot(ra0, e), e));
      const float c = regularized_limit(v, initial, lower ? lo3[i] : hi3[i], lower);
      if (fabsf(c) >= K_C_EPS) {
        const float f = pv[i] * c + sa3[i];
        sa3[i] = lower ? fmin_(f, 0.0f) : fmax_(f, 0.0f);
        pv[i] = fmin_(pv[i] + K_BETA * fabsf(c), K_LIM_PEN_MAX);
      }
    }
    sacc0 = make_float4(sa3[0], sa3[1], sa3[2], 0.0f);
  } else if (kind == JK_WHEEL) {
    // the perp rows, and the rigid row on slot 2 without a suspension
    // frequency (joints_ext.rs:738-766)
    const f3 wa = normalize_or(qrot(qa, xyz(d.J.sus[j])), f3{0.0f, 0.0f, 1.0f});
    f3 u, v;
    tangent_basis(wa, &u, &v);
    const f3 dl = sub(live, mul(c0v, K_ALPHA));
    for (int li = 0; li < 2; ++li) {
      float c = dot3(dl, li == 0 ? u : v);
      if (fabsf(c) >= K_C_EPS) {
        lv[li] += pv[li] * c;
        pv[li] = fmin_(pv[li] + K_BETA * fabsf(c), K_PEN_MAX);
      }
    }
    const float4 pr = d.J.dref[j];
    if (!(pr.y > 0.0f)) {
      const f3 wa0 = normalize_or(qrot(asq(d.b.rot0[ab.x]), xyz(d.J.sus[j])), f3{0.0f, 0.0f, 1.0f});
      const float c = (-dot3(live, wa) - pr.x) - K_ALPHA * (-dot3(c0v, wa0) - pr.x);
      if (fabsf(c) >= K_C_EPS) {
        lv[2] += pv[2] * c;
        pv[2] = fmin_(pv[2] + K_BETA * fabsf(c), K_LIM_PEN_MAX);
      }
    }
  } else if (kind != JK_PRISMATIC) {
    if (kind == JK_FIXED) live = add(live, qrot(qa, xyz(d.J.dref[j])));
    for (int k = 0; k < 3; ++k) {
      float c = comp(live, k) - K_ALPHA * comp(c0v, k);
      if (fabsf(c) >= K_C_EPS) {
        lv[k] += pv[k] * c;
        pv[k] = fmin_(pv[k] + K_BETA * fabsf(c), K_PEN_MAX);
      }
    }
  } else {
    const f3 wa = normalize_or(qrot(qa, xyz(d.J.ax[2 * j])), f3{0.0f, 0.0f, 1.0f});
    f3 u, v;
    tangent_basis(wa, &u, &v);
    const f3 dl = sub(live, mul(c0v, K_ALPHA));
    for (int li = 0; li < 2; ++li) {
      float c = dot3(dl, li == 0 ? u : v);
      if (fabsf(c) >= K_C_EPS) {
        lv[li] += pv[li] * c;
        pv[li] = fmin_(pv[li] + K_BETA * fabsf(c), K_PEN_MAX);
      }
    }
  }
  if (kind == JK_BALL) {
    d.J.lam[j] = make_float4(lv[0], lv[1], lv[2], lam.w);
    d.J.pen[j] = make_float4(pv[0], pv[1], pv[2], 0.0f);
    return;
  }
  float4 la4 = d.J.lama[j], pa4 = d.J.pena[j];
  float av[3] = {la4.x, la4.y, la4.z}, qv[3] = {pa4.x, pa4.y, pa4.z};
  if (kind == JK_SIXDOF) {
    const float4 w0 = d.J.six[4 * j], w2 = d.J.six[4 * j + 2], w3 = d.J.six[4 * j + 3];
    const qf qr = asq(d.J.qref[j]);
    const f3 diff = quat_diff_vec(qmul(qconj(qa), qb), qr);
    float sa3[3] = {sacc1.x, sacc1.y, sacc1.z};
    const float lo3[3] = {w2.x, w2.y, w2.z}, hi3[3] = {w3.x, w3.y, w3.z};
    for (int i = 0; i < 3; ++i) {
      const int cfg = six_cfg(w0.w, 3 + i);
      if (cfg == AX_FREE) continue;
      const f3 e = unit_axis(i);
      if (cfg == AX_LOCKED) {
        const float c = dot3(diff, e);
        if (fabsf(c) >= K_C_EPS) {
          av[i] += qv[i] * c;
          qv[i] = fmin_(qv[i] + K_BETA_ANG * fabsf(c), K_PEN_MAX);
        }
        continue;
      }
      const float v = six_ang_value(qa, qb, qr, e);
      const int st = warm_limit_state(v, lo3[i], hi3[i], K_LIMIT_SLOP_ANG, sa3[i]);
      if (st < 0) {
        sa3[i] = 0.0f;
        continue;
      }
      const bool lower = st == 1;
      const float initial = six_ang_value(asq(d.b.rot0[ab.x]), asq(d.b.rot0[ab.y]), qr, e);
      const float c = regularized_limit(v, initial, lower ? lo3[i] : hi3[i], lower);
      if (fabsf(c) >= K_C_EPS) {
        const float f = qv[i] * c + sa3[i];
        sa3[i] = lower ? fmin_(f, 0.0f) : fmax_(f, 0.0f);
        qv[i] = fmin_(qv[i] + K_BETA_ANG * fabsf(c), K_LIM_PEN_MAX);
      }
    }
    d.J.sacc[2 * j] = sacc0;
    d.J.sacc[2 * j + 1] = make_float4(sa3[0], sa3[1], sa3[2], 0.0f);
  } else if (kind == JK_FIXED) {
    const f3 diff = quat_diff_vec(qmul(qconj(qa), qb), asq(d.J.qref[j]));
    const f3 diff0 = xyz(d.J.a0[2 * j]);
    for (int k = 0; k < 3; ++k) {
      float c = comp(diff, k) - K_ALPHA * comp(diff0, k);
      if (fabsf(c) < K_C_EPS) continue;
      av[k] += qv[k] * c;
      qv[k] = fmin_(qv[k] + K_BETA_ANG * fabsf(c), K_PEN_MAX);
    }
  } else {
    const f3 dax = sub(qrot(qa, xyz(d.J.ax[2 * j])), qrot(qb, xyz(d.J.ax[2 * j + 1])));
    for (int li = 0; li < 2; ++li) {
      const float4 t4 = d.J.a0[2 * j + li];
      float c = dot3(xyz(t4), dax) - K_ALPHA * t4.w;
      if (fabsf(c) < K_C_EPS) continue;
      av[li] = qv[li] * c + av[li];
      qv[li] = fmin_(qv[li] + K_BETA_ANG * fabsf(c), K_PEN_MAX);
    }
  }
  // the limit's one-sided accumulator (joints_ext.rs:820-890): acc_lim
  // (lama.w) for a hinge, lim_dual (lam.w) for a slider, zeroed when the
  // joint is clear of both bounds; penalties capped at K_LIM_PEN_MAX
  float lim_dual = lam.w, acc_lim = la4.w;
  const float4 pr = d.J.dref[j];
  if ((kind == JK_REVOLUTE || kind == JK_PRISMATIC) && ((uint32_t)pr.w & HINGE_LIMIT)) {
    const bool rev = kind == JK_REVOLUTE;
    const f3 wa = normalize_or(qrot(qa, xyz(d.J.ax[2 * j])), f3{0.0f, 0.0f, 1.0f});
    const float v = limit_value(d, j, kind, qa, qb, pa, pb, wa);
    const int st = warm_limit_state(v, pr.y, pr.z, rev ? K_LIMIT_SLOP_ANG : K_LIMIT_SLOP_LIN, rev ? acc_lim : lim_dual);
    if (st >= 0) {
      const bool lower = st == 1;
      const float c = regularized_limit(v, d.J.c0v[j].w, lower ? pr.y : pr.z, lower);
      if (fabsf(c) >= K_C_EPS) {
        if (rev) {
          const float f = qv[2] * c + acc_lim;
          acc_lim = lower ? fmin_(f, 0.0f) : fmax_(f, 0.0f);
          qv[2] = fmin_(qv[2] + K_BETA_ANG * fabsf(c), K_LIM_PEN_MAX);
        } else {
          const float f_raw = pv[2] * c + lim_dual;
          lim_dual = lower ? fmin_(f_raw, 0.0f) : fmax_(f_raw, 0.0f);
          pv[2] = fmin_(pv[2] + K_BETA * fabsf(c), K_LIM_PEN_MAX);
        }
      }
    } else {
      acc_lim = 0.0f;
      if (!rev) lim_dual = 0.0f;
    }
  }
  d.J.lam[j] = make_float4(lv[0], lv[1], lv[2], lim_dual);
  d.J.pen[j] = make_float4(pv[0], pv[1], pv[2], 0.0f);
  d.J.lama[j] = make_float4(av[0], av[1], av[2], acc_lim);
  d.J.pena[j] = make_float4(qv[0], qv[1], qv[2], 0.0f);
}

// ---------------------------------------------------------------------------
// finish: BDF1 (mod.rs:1062-1078) and sleep (mod.rs:1088-1122)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bdf1_body(const Dev& d, uint32_t h) {
  float4 p = d.b.pos[h];
  d.b.supported[h] = 0;
  if (!(d.b.type[h] == BT_DYNAMIC && p.w > 0.0f) || d.b.asleep[h]) return;
  f3 v = divs(sub(xyz(p), xyz(d.b.pos0[h])), K_DT);
  qf q = asq(d.b.rot[h]);
  f3 spin = quat_diff_vec(q, asq(d.b.rot0[h]));
  f3 w = len3(spin) < K_SPIN_REST ? zero3() : divs(spin, K_DT);
  float4 v4 = d.b.vel[h], w4 = d.b.angvel[h];
  d.b.vel[h] = pk(v, v4.w);
  d.b.angvel[h] = pk(w, w4.w);
  d.b.rot[h] = pkq(qnormalize(q));
}
__global__ void k_bdf1(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  bdf1_body(d, h);
}

__global__ void k_supported(Dev d) {
  uint32_t p = gid();
  if (p < dev_npairs(d)) {
    if (d.P.n(p).w <= 0.0f && d.P.pts[p].y > 0u) {
      uint2 ab = d.P.ab[p];
      d.b.supported[ab.x] = 1;
      d.b.supported[ab.y] = 1;
    }
  }
  if (p < d.njoints) {
    uint2 ab = d.J.ab[p];
    d.b.supported[ab.x] = 1;
    d.b.supported[ab.y] = 1;
  }
}

// joint_sleep_ready (sleep.rs:139-150): flag_b = ready.
__global__ void k_ready_init(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  d.b.flag_b[h] = 1;
}

// Scenes without joints: the solver records back to the bodies, BDF1 and
// the sleep readiness flag, per body in one launch.
__global__ void k_finish_bodies(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  sb_unpack_body(d, h);
  bdf1_body(d, h);
  d.b.flag_b[h] = 1;
}
__global__ void k_ready_joints(Dev d) {
  uint32_t j = gid();
  if (j >= d.njoints) return;
  if (joint_unsettled(d, j)) {
    uint32_t m[4];
    joint_members_dev(d, j, m);
    for (int i = 0; i < 4; ++i) d.b.flag_b[m[i]] = 0;
  }
}

// Sleep timers; flag_a = freeze.
__device__ __forceinline__ void sleep_timer_body(const Dev& d, uint32_t h) {
  uint8_t asl = d.b.asleep[h];
  d.b.flag_a[h] = asl;
  if (!(d.b.type[h] == BT_DYNAMIC && d.b.pos[h].w > 0.0f) || asl) return;
  if (!d.b.supported[h] || !d.b.flag_b[h]) {
    d.b.timer[h] = 0.0f;
    return;
  }
  if (len3(xyz(d.b.vel[h])) < K_SLEEP_LIN && len3(xyz(d.b.angvel[h])) < K_SLEEP_ANG) {
    float t = d.b.timer[h] + K_DT;
    d.b.timer[h] = t;
    d.b.flag_a[h] = t >= K_SLEEP_TIME ? 1 : 0;
  } else {
    d.b.timer[h] = 0.0f;
  }
}
__global__ void k_sleep_timer(Dev d) {
  uint32_t h = gid();
  if (h >= d.n) return;
  sleep_timer_body(d, h);
}

// sleep_body (mod.rs:822-828) where frozen; refresh the driver baseline
// (mod.rs:1139-1142); count sleepers.
__device__ __forceinline__ bool sleep_apply_body(const Dev& d, uint32_t h) {
  if (d.b.flag_a[h] && !d.b.asleep[h] && d.b.type[h] == BT_DYNAMIC) {
    float4 v = d.b.vel[h], w = d.b.angvel[h], in = d.b.inertia[h], p = d.b.pos[h];
    d.b.vel[h] = make_float4(0.0f, 0.0f, 0.0f, v.w);
    d.b.angvel[h] = make_float4(0.0f, 0.0f, 0.0f, w.w);
    d.b.inertia[h] = make_float4(0.0f, 0.0f, 0.0f, in.w);
    d.b.pos[h] = make_float4(p.x, p.y, p.z, 0.0f);
    d.b.asleep[h] = 1;
  }
  d.b.flag_a[h] = 0;
  d.b.prev_pos[h] = d.b.pos[h];
  d.b.prev_rot[h] = d.b.rot[h];
  return d.b.asleep[h] != 0;
}
__global__ void __launch_bounds__(TPB) k_sleep_apply(Dev d) {
  uint32_t h = gid();
  bool asleep = false;
  if (h < d.n) asleep = sleep_apply_body(d, h);
  block_add(&d.counters[CTR_SLEEPING], asleep ? 1u : 0u);
}
// Scenes without joints (no readiness to propagate): timer and sleep in
// one launch.
__global__ void __launch_bounds__(TPB) k_sleep_bodies(Dev d) {
  uint32_t h = gid();
  bool asleep = false;
  if (h < d.n) {
    sleep_timer_body(d, h);
    asleep = sleep_apply_body(d, h);
  }
  block_add(&d.counters[CTR_SLEEPING], asleep ? 1u : 0u);
}

// ---------------------------------------------------------------------------
// contact and trigger events (emit_events, mod.rs:1606-1669)
// ---------------------------------------------------------------------------
// Per pair at the end of the step: its deepest anchor pair under the final
// poses (the witness is the anchor on body a), touching when deeper than
// CONTACT_BEGIN_SLOP (trigger.rs:68); a touching pair closing faster than
// CONTACT_HIT_THRESHOLD (trigger.rs:63), from the velocities before the
// step's clamp, gets a hit record -- reported when the pair did not touch
// at the last report (the host holds that set).
__global__ void __launch_bounds__(TPB) k_ev_touch(Dev d, uint32_t* tflag, uint32_t* hflag, uint4* hrec) {
  const uint32_t p = gid();
  if (p >= dev_npairs(d)) return;
  const uint2 ab = d.P.ab[p];
  const f3 pa = xyz(d.b.pos[ab.x]), pb = xyz(d.b.pos[ab.y]);
  const qf qa = asq(d.b.rot[ab.x]), qb = asq(d.b.rot[ab.y]);
  const f3 n = xyz(d.P.n(p));
  float deepest = 0.0f;
  f3 witness = mul(add(pa, pb), 0.5f);
  const uint2 pts = d.P.pts[p];
  for (uint32_t k = 0; k < pts.y; ++k) {
    const uint32_t q = pts.x + k;
    const f3 wa = add(pa, qrot(qa, xyz(d.Q.ra(q))));
    const f3 wb = add(pb, qrot(qb, xyz(d.Q.rb(q))));
    const float gap = dot3(n, sub(wa, wb));
    if (-gap > deepest) {
      deepest = -gap;
      witness = wa;
    }
  }
  const bool touch = deepest > 0.005f;
  uint32_t hit = 0u;
  if (touch) {
    const float approach = -dot3(sub(xyz(d.pre_vel[ab.x]), xyz(d.pre_vel[ab.y])), n);
    if (approach > 1.0f) {
      const f3 nn = neg(n);
      hrec[3 * (size_t)p] = make_uint4(EVK_HIT, ab.x, ab.y, 0u);
      hrec[3 * (size_t)p + 1] = make_uint4(__float_as_uint(witness.x), __float_as_uint(witness.y),
                                           __float_as_uint(witness.z), __float_as_uint(approach));
      hrec[3 * (size_t)p + 2] = make_uint4(__float_as_uint(nn.x), __float_as_uint(nn.y), __float_as_uint(nn.z), 0u);
      hit = 1u;
    }
  }
  tflag[p] = touch ? 1u : 0u;
  hflag[p] = hit;
}

// Stream compaction of flagged entries (exclusive offsets from a scan):
// body-index pairs, and three-word event records.
__global__ void k_ev_compact_ab(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ off,
                                const uint2* __restrict__ ab, uint2* __restrict__ out, uint32_t n) {
  const uint32_t i = gid();
  if (i < n && flag[i]) out[off[i]] = ab[i];
}
__global__ void k_ev_compact_rec(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ off,
                                 const uint4* __restrict__ rec, uint4* __restrict__ out, uint32_t n) {
  const uint32_t i = gid();
  if (i >= n || !flag[i]) return;
  const size_t o = 3 * (size_t)off[i];
  out[o] = rec[3 * (size_t)i];
  out[o + 1] = rec[3 * (size_t)i + 1];
  out[o + 2] = rec[3 * (size_t)i + 2];
}
