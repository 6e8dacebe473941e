// avbd_adjoint.hpp — reverse-mode derivatives of the AVBD step's pieces.
//
// The step is differentiated through its fixed active set: the pair list,
// each point's body-frame anchors, the pair normals and friction frames, the
// colouring, the warm-start map, sleep states and every branch taken
// (touching, live rows, force clamps, penalty updates) are those of the
// forward step; everything continuous is live.  Each function below takes
// the forward inputs and the adjoint (reverse-mode cotangent) of its
// outputs and ADDS the adjoints of its inputs.  Templated on the scalar
// type: float in the device kernels, double in the host checks that compare
// every function with central finite differences.
//
// Forward formulas restated here follow the step (avbd_math.hpp and the
// primal / dual kernels): qrot is glam's Quat * Vec3, quat_integrate
// rows.rs:110-119, quat_diff_vec rows.rs:128-146, contact_force
// rows.rs:1032-1060, row_c_levers rows.rs:1006-1025, friction_levers
// rows.rs:987-996, stamp_row rows.rs:1064-1092, the inertial system
// assembly.rs:74-96 and the dual update joints_ext.rs:464-573.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define ADJ_HD __host__ __device__ __forceinline__
#else
#define ADJ_HD inline
#endif

namespace adj {

template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
struct Q4 {
  T x, y, z, w;
};

template <class T> ADJ_HD V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> ADJ_HD V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> ADJ_HD V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> ADJ_HD V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> ADJ_HD V3<T>& operator+=(V3<T>& a, V3<T> b) {
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
  return a;
}
template <class T> ADJ_HD V3<T>& operator-=(V3<T>& a, V3<T> b) {
  a.x -= b.x;
  a.y -= b.y;
  a.z -= b.z;
  return a;
}
template <class T> ADJ_HD Q4<T>& operator+=(Q4<T>& a, Q4<T> b) {
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
  a.w += b.w;
  return a;
}
template <class T> ADJ_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> ADJ_HD V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> ADJ_HD T len(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> ADJ_HD V3<T> zero() { return {T(0), T(0), T(0)}; }
template <class T> ADJ_HD Q4<T> qzero() { return {T(0), T(0), T(0), T(0)}; }
template <class T> ADJ_HD V3<T> vec(Q4<T> q) { return {q.x, q.y, q.z}; }

// ---- quaternion pieces --------------------------------------------------

// glam Quat * Vec3: v (w^2 - b.b) + b 2 (v.b) + (b x v) 2 w, b = (x, y, z).
template <class T>
ADJ_HD V3<T> qrot(Q4<T> q, V3<T> v) {
  V3<T> b = vec(q);
  T b2 = dot(b, b);
  return (v * (q.w * q.w - b2) + b * (dot(v, b) * T(2))) + cross(b, v) * (q.w * T(2));
}
template <class T>
ADJ_HD void qrot_vjp(Q4<T> q, V3<T> v, V3<T> rb, Q4<T>* qb, V3<T>* vb) {
  V3<T> b = vec(q);
  T w = q.w;
  T rv = dot(rb, v), rbb = dot(rb, b), vbd = dot(v, b);
  if (vb) *vb += rb * (w * w - dot(b, b)) + b * (T(2) * rbb) + cross(rb, b) * (T(2) * w);
  if (qb) {
    V3<T> g = b * (T(-2) * rv) + rb * (T(2) * vbd) + v * (T(2) * rbb) + cross(v, rb) * (T(2) * w);
    qb->x += g.x;
    qb->y += g.y;
    qb->z += g.z;
    qb->w += T(2) * w * rv + T(2) * dot(rb, cross(b, v));
  }
}

// Hamilton product (rows.rs:91-98).
template <class T>
ADJ_HD Q4<T> qmul(Q4<T> a, Q4<T> b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class T>
ADJ_HD void qmul_vjp(Q4<T> a, Q4<T> b, Q4<T> p, Q4<T>* ab, Q4<T>* bb) {
  if (ab) {
    ab->x += p.x * b.w - p.y * b.z + p.z * b.y - p.w * b.x;
    ab->y += p.x * b.z + p.y * b.w - p.z * b.x - p.w * b.y;
    ab->z += -p.x * b.y + p.y * b.x + p.z * b.w - p.w * b.z;
    ab->w += p.x * b.x + p.y * b.y + p.z * b.z + p.w * b.w;
  }
  if (bb) {
    bb->x += p.x * a.w + p.y * a.z - p.z * a.y - p.w * a.x;
    bb->y += -p.x * a.z + p.y * a.w + p.z * a.x - p.w * a.y;
    bb->z += p.x * a.y - p.y * a.x + p.z * a.w - p.w * a.z;
    bb->w += p.x * a.x + p.y * a.y + p.z * a.z + p.w * a.w;
  }
}

template <class T>
ADJ_HD Q4<T> qnormalize(Q4<T> q) {
  T l = std::sqrt((q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w));
  return {q.x / l, q.y / l, q.z / l, q.w / l};
}
// d(q / |q|): (o - n (n . o)) / |q|.
template <class T>
ADJ_HD void qnormalize_vjp(Q4<T> q, Q4<T> o, Q4<T>* qb) {
  T l = std::sqrt((q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w));
  Q4<T> n{q.x / l, q.y / l, q.z / l, q.w / l};
  T no = n.x * o.x + n.y * o.y + n.z * o.z + n.w * o.w;
  qb->x += (o.x - n.x * no) / l;
  qb->y += (o.y - n.y * no) / l;
  qb->z += (o.z - n.z * no) / l;
  qb->w += (o.w - n.w * no) / l;
}

// quat_integrate (rows.rs:110-119): normalize((v k, cos(th/2)) * q), k =
// sin(th/2) / th, th = |v|; the identity below 1e-9.
template <class T>
ADJ_HD Q4<T> quat_integrate(Q4<T> q, V3<T> v) {
  T th = len(v);
  if (th < T(1e-9)) return q;
  T s = std::sin(T(0.5) * th), c = std::cos(T(0.5) * th);
  T k = s / th;
  return qnormalize(qmul(Q4<T>{v.x * k, v.y * k, v.z * k, c}, q));
}
template <class T>
ADJ_HD void quat_integrate_vjp(Q4<T> q, V3<T> v, Q4<T> o, Q4<T>* qb, V3<T>* vb) {
  T th = len(v);
  if (th < T(1e-9)) {
    if (qb) *qb += o;
    return;
  }
  T s = std::sin(T(0.5) * th), c = std::cos(T(0.5) * th);
  T k = s / th;
  Q4<T> dq{v.x * k, v.y * k, v.z * k, c};
  Q4<T> p = qmul(dq, q);
  Q4<T> pb = qzero<T>();
  qnormalize_vjp(p, o, &pb);
  Q4<T> dqb = qzero<T>();
  qmul_vjp(dq, q, pb, &dqb, qb);
  if (vb) {
    // dq = (v k(th), c(th)): k' = (th cos(th/2) / 2 - sin(th/2)) / th^2,
    // c' = -sin(th/2) / 2, dth/dv = v / th.
    T kp = (T(0.5) * th * c - s) / (th * th);
    T cp = T(-0.5) * s;
    V3<T> dxyz{dqb.x, dqb.y, dqb.z};
    T scal = (dot(dxyz, v) * kp + dqb.w * cp) / th;
    *vb += dxyz * k + v * scal;
  }
}

// quat_diff_vec (rows.rs:128-146): r = a * conj(b), xyz flipped if r.w < 0,
// s = min(|xyz|, 1), xyz * 2 asin(s) / s (0 below 1e-9).
template <class T>
ADJ_HD V3<T> quat_diff_vec(Q4<T> a, Q4<T> b) {
  Q4<T> r = qmul(a, Q4<T>{-b.x, -b.y, -b.z, b.w});
  V3<T> xyz = vec(r);
  if (r.w < T(0)) xyz = -xyz;
  T l = len(xyz);
  T s = l < T(1) ? l : T(1);
  if (s < T(1e-9)) return zero<T>();
  return xyz * (T(2) * std::asin(s) / s);
}
template <class T>
ADJ_HD void quat_diff_vec_vjp(Q4<T> a, Q4<T> b, V3<T> o, Q4<T>* ab, Q4<T>* bb) {
  Q4<T> cb{-b.x, -b.y, -b.z, b.w};
  Q4<T> r = qmul(a, cb);
  T sg = r.w < T(0) ? T(-1) : T(1);
  V3<T> xyz = vec(r) * sg;
  T l = len(xyz);
  T s = l < T(1) ? l : T(1);
  if (s < T(1e-9)) return;
  T g = T(2) * std::asin(s) / s;
  V3<T> xb = o * g;
  if (l < T(1)) {
    // g'(s) = 2 / (s sqrt(1 - s^2)) - 2 asin(s) / s^2; ds/dxyz = xyz / s
    T gp = T(2) / (s * std::sqrt(T(1) - s * s)) - T(2) * std::asin(s) / (s * s);
    xb += xyz * (gp * dot(xyz, o) / s);
  }
  Q4<T> rb{xb.x * sg, xb.y * sg, xb.z * sg, T(0)};
  Q4<T> cbb = qzero<T>();
  qmul_vjp(a, cb, rb, ab, bb ? &cbb : nullptr);
  if (bb) {
    bb->x -= cbb.x;
    bb->y -= cbb.y;
    bb->z -= cbb.z;
    bb->w += cbb.w;
  }
}

// Rotation matrix of a quaternion (glam Mat3::from_quat), row-major.
template <class T>
ADJ_HD void rotmat(Q4<T> q, T R[3][3]) {
  T x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
  T xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
  T yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
  T wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
  R[0][0] = T(1) - (yy + zz);
  R[1][0] = xy + wz;
  R[2][0] = xz - wy;
  R[0][1] = xy - wz;
  R[1][1] = T(1) - (xx + zz);
  R[2][1] = yz + wx;
  R[0][2] = xz + wy;
  R[1][2] = yz - wx;
  R[2][2] = T(1) - (xx + yy);
}
// q-adjoint of a matrix adjoint G (dS = sum G_ij dR_ij).
template <class T>
ADJ_HD void rotmat_vjp(Q4<T> q, const T G[3][3], Q4<T>* qb) {
  const T x = q.x, y = q.y, z = q.z, w = q.w;
  qb->x += T(2) * (y * (G[0][1] + G[1][0]) + z * (G[0][2] + G[2][0]) + w * (G[2][1] - G[1][2])) -
           T(4) * x * (G[1][1] + G[2][2]);
  qb->y += T(2) * (x * (G[0][1] + G[1][0]) + z * (G[1][2] + G[2][1]) + w * (G[0][2] - G[2][0])) -
           T(4) * y * (G[0][0] + G[2][2]);
  qb->z += T(2) * (x * (G[0][2] + G[2][0]) + y * (G[1][2] + G[2][1]) + w * (G[1][0] - G[0][1])) -
           T(4) * z * (G[0][0] + G[1][1]);
  qb->w += T(2) * (x * (G[2][1] - G[1][2]) + y * (G[0][2] - G[2][0]) + z * (G[1][0] - G[0][1]));
}
// S = w^T R diag(D) R^T z: adjoints of q (through R) and of D.
template <class T>
ADJ_HD T inertia_form(Q4<T> q, V3<T> D, V3<T> w, V3<T> z) {
  T R[3][3];
  rotmat(q, R);
  T rw[3], rz[3];
  for (int i = 0; i < 3; ++i) {
    rw[i] = R[0][i] * w.x + R[1][i] * w.y + R[2][i] * w.z;
    rz[i] = R[0][i] * z.x + R[1][i] * z.y + R[2][i] * z.z;
  }
  return D.x * rw[0] * rz[0] + D.y * rw[1] * rz[1] + D.z * rw[2] * rz[2];
}
template <class T>
ADJ_HD void inertia_form_vjp(Q4<T> q, V3<T> D, V3<T> w, V3<T> z, T sb, Q4<T>* qb, V3<T>* Db, V3<T>* wb,
                             V3<T>* zb) {
  T R[3][3];
  rotmat(q, R);
  T rw[3], rz[3];
  const T wv[3] = {w.x, w.y, w.z}, zv[3] = {z.x, z.y, z.z}, Dv[3] = {D.x, D.y, D.z};
  for (int i = 0; i < 3; ++i) {
    rw[i] = R[0][i] * w.x + R[1][i] * w.y + R[2][i] * w.z;
    rz[i] = R[0][i] * z.x + R[1][i] * z.y + R[2][i] * z.z;
  }
  if (Db) *Db += V3<T>{sb * rw[0] * rz[0], sb * rw[1] * rz[1], sb * rw[2] * rz[2]};
  // dS/dR_ij = D_j (w_i rz_j + z_i rw_j)
  T G[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) G[i][j] = sb * Dv[j] * (wv[i] * rz[j] + zv[i] * rw[j]);
  if (qb) rotmat_vjp(q, G, qb);
  // S = sum_j D_j rw_j rz_j: dS/dw = R (D .* rz), dS/dz = R (D .* rw)
  if (wb || zb) {
    T a[3], c[3];
    for (int j = 0; j < 3; ++j) {
      a[j] = sb * Dv[j] * rz[j];
      c[j] = sb * Dv[j] * rw[j];
    }
    if (wb)
      *wb += V3<T>{R[0][0] * a[0] + R[0][1] * a[1] + R[0][2] * a[2], R[1][0] * a[0] + R[1][1] * a[1] + R[1][2] * a[2],
                   R[2][0] * a[0] + R[2][1] * a[1] + R[2][2] * a[2]};
    if (zb)
      *zb += V3<T>{R[0][0] * c[0] + R[0][1] * c[1] + R[0][2] * c[2], R[1][0] * c[0] + R[1][1] * c[1] + R[1][2] * c[2],
                   R[2][0] * c[0] + R[2][1] * c[1] + R[2][2] * c[2]};
  }
}

// ---- contact rows -----------------------------------------------------

// contact_force (rows.rs:1032-1060) for the branch the forward took.
template <class T>
struct Force {
  T fn, f1, f2;
};
template <class T>
ADJ_HD Force<T> contact_force(T cn, T ct1, T ct2, const T lam[3], const T pen[3], T mu1, T mu2) {
  Force<T> o;
  T a = pen[0] * cn + lam[0];
  o.fn = a < T(0) ? a : T(0);
  T r0 = pen[1] * ct1 + lam[1], r1 = pen[2] * ct2 + lam[2];
  T b1 = std::fabs(o.fn) * mu1, b2 = std::fabs(o.fn) * mu2;
  if (b1 <= T(0) || b2 <= T(0)) {
    o.f1 = b1 <= T(0) ? T(0) : (r0 < -b1 ? -b1 : (r0 > b1 ? b1 : r0));
    o.f2 = b2 <= T(0) ? T(0) : (r1 < -b2 ? -b2 : (r1 > b2 ? b2 : r1));
  } else {
    T s = (r0 / b1) * (r0 / b1) + (r1 / b2) * (r1 / b2);
    T k = s > T(1) ? T(1) / std::sqrt(s) : T(1);
    o.f1 = r0 * k;
    o.f2 = r1 * k;
  }
  return o;
}
// Adds the adjoints of (cn, ct1, ct2, lam, pen, mu1, mu2) for force
// adjoints (fnb, f1b, f2b).
template <class T>
ADJ_HD void contact_force_vjp(T cn, T ct1, T ct2, const T lam[3], const T pen[3], T mu1, T mu2, T fnb, T f1b, T f2b,
                              T* cnb, T* ct1b, T* ct2b, T lamb[3], T penb[3], T* mu1b, T* mu2b) {
  T a = pen[0] * cn + lam[0];
  T fn = a < T(0) ? a : T(0);
  T r0 = pen[1] * ct1 + lam[1], r1 = pen[2] * ct2 + lam[2];
  T afn = std::fabs(fn), sfn = fn < T(0) ? T(-1) : T(1);
  T b1 = afn * mu1, b2 = afn * mu2;
  T r0b = T(0), r1b = T(0), b1b = T(0), b2b = T(0);
  if (b1 <= T(0) || b2 <= T(0)) {
    if (!(b1 <= T(0))) {
      if (r0 < -b1)
        b1b -= f1b;
      else if (r0 > b1)
        b1b += f1b;
      else
        r0b += f1b;
    }
    if (!(b2 <= T(0))) {
      if (r1 < -b2)
        b2b -= f2b;
      else if (r1 > b2)
        b2b += f2b;
      else
        r1b += f2b;
    }
  } else {
    T u = r0 / b1, v = r1 / b2;
    T s = u * u + v * v;
    if (s > T(1)) {
      T k = T(1) / std::sqrt(s);
      r0b += f1b * k;
      r1b += f2b * k;
      T kb = f1b * r0 + f2b * r1;
      T sb = T(-0.5) * k * k * k * kb;
      T ub = T(2) * u * sb, vb = T(2) * v * sb;
      r0b += ub / b1;
      b1b -= ub * r0 / (b1 * b1);
      r1b += vb / b2;
      b2b -= vb * r1 / (b2 * b2);
    } else {
      r0b += f1b;
      r1b += f2b;
    }
  }
  // b_i = |fn| mu_i
  T fnbt = fnb + (b1b * mu1 + b2b * mu2) * sfn;
  *mu1b += b1b * afn;
  *mu2b += b2b * afn;
  *ct1b += r0b * pen[1];
  penb[1] += r0b * ct1;
  lamb[1] += r0b;
  *ct2b += r1b * pen[2];
  penb[2] += r1b * ct2;
  lamb[2] += r1b;
  if (a < T(0)) {
    *cnb += fnbt * pen[0];
    penb[0] += fnbt * cn;
    lamb[0] += fnbt;
  }
}

// Step-frame data of one pair and one of its points, and the poses the
// rows read: body A, body B (current x, q; step-start x0, q0).
template <class T>
struct Contact {
  V3<T> n, t1, t2;   // normal (B -> A) and friction frame
  T mu1, mu2;        // friction per frame axis
  T damp;            // separated-pair damper stiffness (0: none)
  bool touching;     // pair gap <= 0
  V3<T> ra, rb;      // body-frame anchors
  T rad_a, rad_b;    // sphere radius (< 0: not a sphere)
  V3<T> xa, xb, xa0, xb0;
  Q4<T> qa, qb, qa0, qb0;
  T lam[3], pen[3];
};
// Adjoints of a Contact's live inputs.
template <class T>
struct ContactBar {
  V3<T> xa, xb, xa0, xb0;
  Q4<T> qa, qb, qa0, qb0;
  T lam[3], pen[3];
  T mu1, mu2, damp;
};
template <class T>
ADJ_HD void clear(ContactBar<T>& b) {
  b.xa = b.xb = b.xa0 = b.xb0 = zero<T>();
  b.qa = b.qb = b.qa0 = b.qb0 = qzero<T>();
  for (int k = 0; k < 3; ++k) b.lam[k] = b.pen[k] = T(0);
  b.mu1 = b.mu2 = b.damp = T(0);
}

// Row values of a contact point at its current poses: levers, C values and
// forces (rows.rs:947-1060).
template <class T>
struct RowVals {
  V3<T> ra_w, rb_w, ra_f, rb_f, dra, drb;
  T sep, c0, cn, ct1, ct2;
  Force<T> f;
};
template <class T>
ADJ_HD T row_c(V3<T> u, T cc, V3<T> la, V3<T> lb, V3<T> dpa, V3<T> dpb, V3<T> dra, V3<T> drb) {
  V3<T> d_a = dpa + u * dot(cross(la, u), dra);
  V3<T> d_b = dpb + u * dot(cross(lb, u), drb);
  return cc * (T(1) - T(0.99)) + dot(u, d_a - d_b);
}
template <class T>
ADJ_HD void row_c_vjp(V3<T> u, V3<T> la, V3<T> lb, V3<T> dra, V3<T> drb, T cb, T* ccb, V3<T>* lab, V3<T>* lbb,
                      V3<T>* dpab, V3<T>* dpbb, V3<T>* drab, V3<T>* drbb) {
  if (ccb) *ccb += cb * (T(1) - T(0.99));
  V3<T> dab = u * cb;  // d(u . d_a)
  *dpab += dab;
  *dpbb -= dab;
  T ta = dot(dab, u), tb = -dot(dab, u);
  // (la x u) . dra = la . (u x dra)
  *lab += cross(u, dra) * ta;
  *drab += cross(la, u) * ta;
  *lbb += cross(u, drb) * tb;
  *drbb += cross(lb, u) * tb;
}
template <class T>
ADJ_HD RowVals<T> row_vals(const Contact<T>& c) {
  RowVals<T> r;
  r.ra_w = c.rad_a >= T(0) ? (-c.n) * c.rad_a : qrot(c.qa, c.ra);
  r.rb_w = c.rad_b >= T(0) ? c.n * c.rad_b : qrot(c.qb, c.rb);
  r.dra = quat_diff_vec(c.qa, c.qa0);
  r.drb = quat_diff_vec(c.qb, c.qb0);
  V3<T> dpa = c.xa - c.xa0, dpb = c.xb - c.xb0;
  r.c0 = T(0);
  if (c.touching && c.rad_a < T(0) && c.rad_b < T(0))
    r.c0 = dot(c.n, (c.xa0 + qrot(c.qa0, c.ra)) - (c.xb0 + qrot(c.qb0, c.rb))) + T(0.01);
  r.cn = row_c(c.n, r.c0, r.ra_w, r.rb_w, dpa, dpb, r.dra, r.drb);
  r.sep = dot((c.xa + r.ra_w) - (c.xb + r.rb_w), c.n);
  r.ra_f = r.ra_w - c.n * (r.sep * T(0.5));
  r.rb_f = r.rb_w + c.n * (r.sep * T(0.5));
  r.ct1 = row_c(c.t1, T(0), r.ra_f, r.rb_f, dpa, dpb, r.dra, r.drb);
  r.ct2 = row_c(c.t2, T(0), r.ra_f, r.rb_f, dpa, dpb, r.dra, r.drb);
  r.f = contact_force(r.cn, r.ct1, r.ct2, c.lam, c.pen, c.mu1, c.mu2);
  return r;
}
// Adjoints of a Contact from adjoints of its C values, forces and levers.
template <class T>
struct RowValsBar {
  V3<T> ra_w, rb_w, ra_f, rb_f;
  T cn, ct1, ct2, fn, f1, f2;
};
template <class T>
ADJ_HD void row_vals_vjp(const Contact<T>& c, const RowVals<T>& r, RowValsBar<T> g, ContactBar<T>& o) {
  V3<T> dpa = c.xa - c.xa0, dpb = c.xb - c.xb0;
  // forces
  contact_force_vjp(r.cn, r.ct1, r.ct2, c.lam, c.pen, c.mu1, c.mu2, g.fn, g.f1, g.f2, &g.cn, &g.ct1, &g.ct2, o.lam,
                    o.pen, &o.mu1, &o.mu2);
  V3<T> dpab = zero<T>(), dpbb = zero<T>(), drab = zero<T>(), drbb = zero<T>();
  // friction rows
  row_c_vjp(c.t1, r.ra_f, r.rb_f, r.dra, r.drb, g.ct1, (T*)nullptr, &g.ra_f, &g.rb_f, &dpab, &dpbb, &drab, &drbb);
  row_c_vjp(c.t2, r.ra_f, r.rb_f, r.dra, r.drb, g.ct2, (T*)nullptr, &g.ra_f, &g.rb_f, &dpab, &dpbb, &drab, &drbb);
  // friction levers: ra_f = ra_w - n sep / 2, rb_f = rb_w + n sep / 2
  g.ra_w += g.ra_f;
  g.rb_w += g.rb_f;
  T sepb = T(-0.5) * dot(g.ra_f, c.n) + T(0.5) * dot(g.rb_f, c.n);
  o.xa += c.n * sepb;
  g.ra_w += c.n * sepb;
  o.xb -= c.n * sepb;
  g.rb_w -= c.n * sepb;
  // normal row
  T c0b = T(0);
  row_c_vjp(c.n, r.ra_w, r.rb_w, r.dra, r.drb, g.cn, &c0b, &g.ra_w, &g.rb_w, &dpab, &dpbb, &drab, &drbb);
  if (c.touching && c.rad_a < T(0) && c.rad_b < T(0)) {
    o.xa0 += c.n * c0b;
    o.xb0 -= c.n * c0b;
    qrot_vjp(c.qa0, c.ra, c.n * c0b, &o.qa0, (V3<T>*)nullptr);
    qrot_vjp(c.qb0, c.rb, -(c.n * c0b), &o.qb0, (V3<T>*)nullptr);
  }
  o.xa += dpab;
  o.xa0 -= dpab;
  o.xb += dpbb;
  o.xb0 -= dpbb;
  quat_diff_vec_vjp(c.qa, c.qa0, drab, &o.qa, &o.qa0);
  quat_diff_vec_vjp(c.qb, c.qb0, drbb, &o.qb, &o.qb0);
  if (c.rad_a < T(0)) qrot_vjp(c.qa, c.ra, g.ra_w, &o.qa, (V3<T>*)nullptr);
  if (c.rad_b < T(0)) qrot_vjp(c.qb, c.rb, g.rb_w, &o.qb, (V3<T>*)nullptr);
}

// The row terms of one point-side in S = w^T g + w^T H dx (a body solve's
// sensitivity to its inputs, with w = H^-1 s_bar and dx the solve's step):
// each live row r adds f_r (w . J_r) + pen_r (w . J_r)(J_r . dx), J_r =
// (sign u, lever x sign u).  Adds -dS/d(inputs) to `o` (the adjoint of the
// inputs of dx = -H^-1 g).  `is_a`: the solved body is the pair's A side.
template <class T>
ADJ_HD T row_term(V3<T> u, V3<T> lever, T sign, const T w[6], const T dx[6], T* ab, T* bb) {
  V3<T> nn = u * sign;
  V3<T> t = cross(lever, nn);
  *ab = w[0] * nn.x + w[1] * nn.y + w[2] * nn.z + w[3] * t.x + w[4] * t.y + w[5] * t.z;
  *bb = dx[0] * nn.x + dx[1] * nn.y + dx[2] * nn.z + dx[3] * t.x + dx[4] * t.y + dx[5] * t.z;
  return T(0);
}
// lever adjoint of a row term with adjoints (a_bar, b_bar) of a = w . J and
// b = dx . J: t = lever x nn, so lever_bar = nn x t_bar.
template <class T>
ADJ_HD V3<T> row_term_lever_vjp(V3<T> u, T sign, const T w[6], const T dx[6], T abar, T bbar) {
  V3<T> nn = u * sign;
  V3<T> tb{abar * w[3] + bbar * dx[3], abar * w[4] + bbar * dx[4], abar * w[5] + bbar * dx[5]};
  return cross(nn, tb);
}
template <class T>
ADJ_HD void point_side_vjp(const Contact<T>& c, bool is_a, const T w[6], const T dx[6], ContactBar<T>& o) {
  const T sign = is_a ? T(1) : T(-1);
  RowVals<T> r = row_vals(c);
  RowValsBar<T> g{zero<T>(), zero<T>(), zero<T>(), zero<T>(), T(0), T(0), T(0), T(0), T(0), T(0)};
  V3<T> r_side = is_a ? r.ra_w : r.rb_w;
  if (!c.touching) {
    // damper: f = min(damp cn, 0), stiffness damp
    if (!(c.damp > T(0) && std::fabs(r.cn) >= T(1e-7))) return;
    T a, b;
    row_term(c.n, r_side, sign, w, dx, &a, &b);
    T fd = c.damp * r.cn;
    bool act = fd < T(0);
    T f = act ? fd : T(0);
    // S = f a + damp a b; S_bar = -1
    T abar = -(f + c.damp * b), bbar = -(c.damp * a);
    T fbar = -a;
    o.damp += -(a * b) + (act ? fbar * r.cn : T(0));
    g.cn += act ? fbar * c.damp : T(0);
    V3<T> lb = row_term_lever_vjp(c.n, sign, w, dx, abar, bbar);
    if (is_a) g.ra_w += lb; else g.rb_w += lb;
    // the damper's normal row reads the normal C only (no friction rows)
    T c0b = T(0);
    V3<T> dpab = zero<T>(), dpbb = zero<T>(), drab = zero<T>(), drbb = zero<T>();
    row_c_vjp(c.n, r.ra_w, r.rb_w, r.dra, r.drb, g.cn, &c0b, &g.ra_w, &g.rb_w, &dpab, &dpbb, &drab, &drbb);
    o.xa += dpab;
    o.xa0 -= dpab;
    o.xb += dpbb;
    o.xb0 -= dpbb;
    quat_diff_vec_vjp(c.qa, c.qa0, drab, &o.qa, &o.qa0);
    quat_diff_vec_vjp(c.qb, c.qb0, drbb, &o.qb, &o.qb0);
    if (c.rad_a < T(0)) qrot_vjp(c.qa, c.ra, g.ra_w, &o.qa, (V3<T>*)nullptr);
    if (c.rad_b < T(0)) qrot_vjp(c.qb, c.rb, g.rb_w, &o.qb, (V3<T>*)nullptr);
    return;
  }
  V3<T> r_side_f = is_a ? r.ra_f : r.rb_f;
  const V3<T> ax[3] = {c.n, c.t1, c.t2};
  const T cv[3] = {r.cn, r.ct1, r.ct2};
  const T fv[3] = {r.f.fn, r.f.f1, r.f.f2};
  T fb[3] = {T(0), T(0), T(0)};
  for (int k = 0; k < 3; ++k) {
    if (!(std::fabs(cv[k]) >= T(1e-7) || fv[k] != T(0))) continue;
    const V3<T> lev = k == 0 ? r_side : r_side_f;
    T a, b;
    row_term(ax[k], lev, sign, w, dx, &a, &b);
    const T pen = c.pen[k];
    T abar = -(fv[k] + pen * b), bbar = -(pen * a);
    fb[k] = -a;
    o.pen[k] += -(a * b);
    V3<T> lb = row_term_lever_vjp(ax[k], sign, w, dx, abar, bbar);
    if (k == 0) {
      if (is_a) g.ra_w += lb; else g.rb_w += lb;
    } else {
      if (is_a) g.ra_f += lb; else g.rb_f += lb;
    }
  }
  g.fn = fb[0];
  g.f1 = fb[1];
  g.f2 = fb[2];
  row_vals_vjp(c, r, g, o);
}

// Dual update of a touching point (joints_ext.rs:464-545), forward: the new
// lam and pen from the forces at the current poses.
template <class T>
ADJ_HD void dual_forward(const Contact<T>& c, T lam_o[3], T pen_o[3]) {
  RowVals<T> r = row_vals(c);
  for (int k = 0; k < 3; ++k) {
    lam_o[k] = c.lam[k];
    pen_o[k] = c.pen[k];
  }
  if (!c.touching) return;
  const T CE = T(1e-7), BETA = T(10000), PMAX = T(1e10);
  if (std::fabs(r.cn) >= CE) {
    lam_o[0] = r.f.fn;
    if (r.f.fn < T(0)) {
      T p = c.pen[0] + BETA * std::fabs(r.cn);
      pen_o[0] = p < PMAX ? p : PMAX;
    }
  }
  T r0 = c.pen[1] * r.ct1 + c.lam[1], r1 = c.pen[2] * r.ct2 + c.lam[2];
  T bb1 = std::fabs(r.f.fn) * c.mu1, bb2 = std::fabs(r.f.fn) * c.mu2;
  bool ell = (bb1 > T(0) && bb2 > T(0)) && ((r0 / bb1) * (r0 / bb1) + (r1 / bb2) * (r1 / bb2) <= T(1));
  lam_o[1] = r.f.f1;
  lam_o[2] = r.f.f2;
  if (ell) {
    const T ct[2] = {r.ct1, r.ct2};
    for (int k = 1; k < 3; ++k)
      if (std::fabs(ct[k - 1]) >= CE) {
        T p = c.pen[k] + BETA * std::fabs(ct[k - 1]);
        pen_o[k] = p < PMAX ? p : PMAX;
      }
  }
}
// Reverse of dual_forward: adjoints of the new (lam, pen) to the inputs.
template <class T>
ADJ_HD void dual_vjp(const Contact<T>& c, const T lamb_o[3], const T penb_o[3], ContactBar<T>& o) {
  if (!c.touching) {
    for (int k = 0; k < 3; ++k) {
      o.lam[k] += lamb_o[k];
      o.pen[k] += penb_o[k];
    }
    return;
  }
  RowVals<T> r = row_vals(c);
  RowValsBar<T> g{zero<T>(), zero<T>(), zero<T>(), zero<T>(), T(0), T(0), T(0), T(0), T(0), T(0)};
  const T CE = T(1e-7), BETA = T(10000), PMAX = T(1e10);
  // normal
  if (std::fabs(r.cn) >= CE) {
    g.fn += lamb_o[0];
    if (r.f.fn < T(0)) {
      T p = c.pen[0] + BETA * std::fabs(r.cn);
      if (p < PMAX) {
        o.pen[0] += penb_o[0];
        g.cn += penb_o[0] * BETA * (r.cn < T(0) ? T(-1) : T(1));
      }
    } else {
      o.pen[0] += penb_o[0];
    }
  } else {
    o.lam[0] += lamb_o[0];
    o.pen[0] += penb_o[0];
  }
  // friction
  T r0 = c.pen[1] * r.ct1 + c.lam[1], r1 = c.pen[2] * r.ct2 + c.lam[2];
  T bb1 = std::fabs(r.f.fn) * c.mu1, bb2 = std::fabs(r.f.fn) * c.mu2;
  bool ell = (bb1 > T(0) && bb2 > T(0)) && ((r0 / bb1) * (r0 / bb1) + (r1 / bb2) * (r1 / bb2) <= T(1));
  g.f1 += lamb_o[1];
  g.f2 += lamb_o[2];
  const T ct[2] = {r.ct1, r.ct2};
  for (int k = 1; k < 3; ++k) {
    bool upd = ell && std::fabs(ct[k - 1]) >= CE;
    if (upd) {
      T p = c.pen[k] + BETA * std::fabs(ct[k - 1]);
      if (p < PMAX) {
        o.pen[k] += penb_o[k];
        T d = penb_o[k] * BETA * (ct[k - 1] < T(0) ? T(-1) : T(1));
        if (k == 1) g.ct1 += d; else g.ct2 += d;
      }
    } else {
      o.pen[k] += penb_o[k];
    }
  }
  row_vals_vjp(c, r, g, o);
}

// ---- inertial system -----------------------------------------------------

// The inertial terms of S = w^T g + w^T H dx (assembly.rs:74-96):
// H_in = diag(m/dt^2, I_w/dt^2), g_in = (m/dt^2 (x - y), I_w dq/dt^2) with
// I_w = R diag(D) R^T, dq = quat_diff_vec(q, q_y), D = m * dunit (the body
// inertia per unit mass).  Adds -dS/d(x, q, y, q_y, m) to the outputs.
template <class T>
ADJ_HD void inertial_vjp(V3<T> x, Q4<T> q, V3<T> y, Q4<T> qy, T m, V3<T> dunit, const T w[6], const T dx[6],
                         T dt, V3<T>* xb, Q4<T>* qb, V3<T>* yb, Q4<T>* qyb, T* mb) {
  const T dt2 = dt * dt;
  V3<T> wl{w[0], w[1], w[2]}, wa{w[3], w[4], w[5]};
  V3<T> dl{dx[0], dx[1], dx[2]}, da{dx[3], dx[4], dx[5]};
  // linear: S_l = m/dt^2 wl.(x - y + dl)
  *xb -= wl * (m / dt2);
  *yb += wl * (m / dt2);
  *mb -= dot(wl, (x - y) + dl) / dt2;
  // angular: S_a = w_a^T R D R^T (dq + da) / dt^2
  V3<T> dq = quat_diff_vec(q, qy);
  V3<T> z = (dq + da) * (T(1) / dt2);
  V3<T> D = dunit * m;
  Q4<T> qbb = qzero<T>();
  V3<T> Db = zero<T>(), zb = zero<T>();
  inertia_form_vjp(q, D, wa, z, T(-1), &qbb, &Db, (V3<T>*)nullptr, &zb);
  *qb += qbb;
  *mb += dot(Db, dunit);
  quat_diff_vec_vjp(q, qy, zb * (T(1) / dt2), qb, qyb);
}

// The 6x6 solve's transpose: w = H^-1 s for the same LDL^T factors
// (solve_6x6, assembly.rs:25-69; H symmetric).  L21: lower triangle.
template <class T>
ADJ_HD bool solve6(const T L21[21], const T rhs[6], T x[6]) {
  T l[6][6];
  T dg[6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) {
      T s = L21[i * (i + 1) / 2 + j];
      for (int k = 0; k < j; ++k) s -= l[i][k] * dg[k] * l[j][k];
      if (i == j) {
        if (s <= T(1e-12)) return false;
        dg[i] = s;
        l[i][i] = T(1);
      } else {
        l[i][j] = s / dg[j];
      }
    }
  T y[6];
  for (int i = 0; i < 6; ++i) {
    T s = rhs[i];
    for (int k = 0; k < i; ++k) s -= l[i][k] * y[k];
    y[i] = s;
  }
  for (int i = 0; i < 6; ++i) y[i] = y[i] / dg[i];
  for (int i = 5; i >= 0; --i) {
    T s = y[i];
    for (int k = i + 1; k < 6; ++k) s -= l[k][i] * x[k];
    x[i] = s;
  }
  return true;
}

}  // namespace adj
