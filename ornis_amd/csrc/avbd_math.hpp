// avbd_math.hpp — single-precision vector / quaternion arithmetic of the
// AVBD step, for gfx950 device code (and host code of the same library).
//
// Every function reproduces the operation order of the glam 0.34.3 call it
// stands for (the reference's math library, Cargo.lock:1357-1358): Vec3 is
// glam's scalar struct, Quat follows glam's SSE2 path.  The library is
// compiled with -ffp-contract=off, so no multiply-add is fused and each
// result is the correctly rounded IEEE single-precision value of the
// written expression.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/oavbd_fmath.h"

#define HD __host__ __device__ __forceinline__

struct f3 {
  float x, y, z;
};
struct qf {
  float x, y, z, w;
};

HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
HD f3 zero3() { return f3{0.0f, 0.0f, 0.0f}; }
HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
HD f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
HD f3 mul(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
HD f3 smul(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
HD f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
HD f3 cmul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
HD void setc(f3& a, int i, float v) {
  if (i == 0)
    a.x = v;
  else if (i == 1)
    a.y = v;
  else
    a.z = v;
}
HD bool eq3(f3 a, f3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
HD bool eqq(qf a, qf b) { return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w; }
// glam Vec3::dot: (x*x' + y*y') + z*z'.
HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HD f3 cross3(f3 a, f3 b) { return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
HD float len2(f3 a) { return dot3(a, a); }
HD float len3(f3 a) { return sqrtf(dot3(a, a)); }
HD bool finite3(f3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
// glam normalize_or: reciprocal multiply, fallback unless the reciprocal is
// a finite positive number.
HD f3 normalize_or(f3 a, f3 fb) {
  float rcp = 1.0f / len3(a);
  if (isfinite(rcp) && rcp > 0.0f) return mul(a, rcp);
  return fb;
}
HD f3 normalize3(f3 a) { return mul(a, 1.0f / len3(a)); }
HD f3 abs3(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
// f32::min / f32::max with the x86-64 libm result for every input,
// including (+0, -0): the second operand unless the first is strictly
// smaller (larger), NaN yields the other operand.
HD float fmin_(float a, float b) { return (a < b || b != b) ? a : b; }
HD float fmax_(float a, float b) { return (a > b || b != b) ? a : b; }
HD f3 max3(f3 a, f3 b) { return f3{fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
HD f3 min3(f3 a, f3 b) { return f3{fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
// Rust f32::clamp (comparison form).
HD float clampf_(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
HD float signumf(float x) { return isnan(x) ? x : copysignf(1.0f, x); }

HD qf qid() { return qf{0.0f, 0.0f, 0.0f, 1.0f}; }

// wrap_pi (crates/physics/src/avbd/rows.rs:149-151): (x + pi) rem_euclid
// tau, minus pi.  The remainder is exact: below two periods it is one
// exact subtraction (Sterbenz), which equals fmod there bit for bit.
#define K_PI 3.14159265358979323846f
#define K_TAU 6.28318530717958647692f
HD float wrap_pi(float x) {
  const float v = x + K_PI;
  float r;
  if (fabsf(v) < 2.0f * K_TAU)
    r = fabsf(v) < K_TAU ? v : (v > 0.0f ? v - K_TAU : v + K_TAU);
  else
    r = fmodf(v, K_TAU);
  if (r < 0.0f) r = r + K_TAU;
  return r - K_PI;
}
HD qf qconj(qf q) { return qf{-q.x, -q.y, -q.z, q.w}; }
// glam Quat * Vec3 (SSE2 mul_vec3a).
HD f3 qrot(qf q, f3 v) {
  f3 b{q.x, q.y, q.z};
  float b2 = dot3(b, b);
  f3 t0 = mul(v, q.w * q.w - b2);
  f3 t1 = mul(b, dot3(v, b) * 2.0f);
  f3 t2 = mul(cross3(b, v), q.w * 2.0f);
  return add(add(t0, t1), t2);
}
// glam Quat::normalize (SSE2): divide by sqrt((x^2 + z^2) + (y^2 + w^2)).
HD qf qnormalize(qf q) {
  float d = (q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w);
  float l = sqrtf(d);
  return qf{q.x / l, q.y / l, q.z / l, q.w / l};
}
// Hamilton product, rows.rs:91-98.
HD qf qmul(qf a, qf b) {
  return qf{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}

// quat_twist / hinge_twist (crates/physics/src/sequential_impulse/joints.rs:
// 38-48): the twist of a relative rotation about an axis, in [-pi, pi].
HD float quat_twist(qf q, f3 axis) { return wrap_pi(2.0f * oavbd_atan2f(dot3(f3{q.x, q.y, q.z}, axis), q.w)); }
HD float hinge_twist(qf qa, qf qb, f3 axis) { return quat_twist(qmul(qconj(qa), qb), axis); }
// quat_integrate, rows.rs:110-119.
HD qf quat_integrate(qf q, f3 v) {
  float theta = len3(v);
  if (theta < 1e-9f) return q;
  float s, c;
  oavbd_sincosf(0.5f * theta, &s, &c);
  float k = s / theta;
  return qnormalize(qmul(qf{v.x * k, v.y * k, v.z * k, c}, q));
}
// quat_diff_vec, rows.rs:128-146.
HD f3 quat_diff_vec(qf a, qf b) {
  qf r = qmul(a, qf{-b.x, -b.y, -b.z, b.w});
  f3 xyz{r.x, r.y, r.z};
  if (r.w < 0.0f) xyz = neg(xyz);
  float s = fmin_(len3(xyz), 1.0f);
  if (s < 1e-9f) return zero3();
  return mul(xyz, 2.0f * oavbd_asinf(s) / s);
}
// glam Mat3::from_quat, as three columns.
struct m3c {
  f3 c0, c1, c2;
};
HD m3c mat3_from_quat(qf q) {
  float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
  float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
  float yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
  float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
  m3c m;
  m.c0 = f3{1.0f - (yy + zz), xy + wz, xz - wy};
  m.c1 = f3{xy - wz, 1.0f - (xx + zz), yz + wx};
  m.c2 = f3{xz + wy, yz - wx, 1.0f - (xx + yy)};
  return m;
}
// Column i (0, 1 or 2), selected with bit masks: with a lane-dependent i a
// select between the columns is otherwise lowered to a table in private
// (scratch) memory.
HD float bitsel3(float a, float b, float c, int i) {
  const int m0 = -(int)(i == 0), m1 = -(int)(i == 1), m2 = -(int)(i == 2);
  return __builtin_bit_cast(float, (__builtin_bit_cast(int, a) & m0) | (__builtin_bit_cast(int, b) & m1) |
                                       (__builtin_bit_cast(int, c) & m2));
}
HD f3 mcol(const m3c& m, int i) {
  return f3{bitsel3(m.c0.x, m.c1.x, m.c2.x, i), bitsel3(m.c0.y, m.c1.y, m.c2.y, i),
            bitsel3(m.c0.z, m.c1.z, m.c2.z, i)};
}

// tangent_basis (crates/physics/src/math.rs:152-160).
HD void tangent_basis(f3 n, f3* t1, f3* t2) {
  f3 axis = fabsf(n.x) < 0.9f ? f3{1.0f, 0.0f, 0.0f} : f3{0.0f, 1.0f, 0.0f};
  *t1 = normalize_or(cross3(n, axis), f3{0.0f, 0.0f, 1.0f});
  *t2 = cross3(*t1, n);
}

// world_inertia (rows.rs:194-207), symmetric 3x3 row-major.
// orthogonalize_axle (crates/physics/src/math.rs:179-186): the axle less
// its suspension component, normalised; the suspension's first tangent
// when that leaves (almost) nothing.
HD f3 orthogonalize_axle(f3 suspension, f3 axle) {
  const f3 a = sub(axle, mul(suspension, dot3(axle, suspension)));
  if (len2(a) < 1e-6f) {
    f3 t1, t2;
    tangent_basis(suspension, &t1, &t2);
    return t1;
  }
  return normalize3(a);
}

HD void world_inertia(f3 inertia, qf rot, float m[3][3]) {
  m3c r = mat3_from_quat(rot);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      float ca0 = comp(r.c0, a), ca1 = comp(r.c1, a), ca2 = comp(r.c2, a);
      float cb0 = comp(r.c0, b), cb1 = comp(r.c1, b), cb2 = comp(r.c2, b);
      m[a][b] = ca0 * inertia.x * cb0 + ca1 * inertia.y * cb1 + ca2 * inertia.z * cb2;
    }
}

HD f3 mat3_vec(const float m[3][3], f3 v) {
  return f3{m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
            m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
}

// inverse_symmetric (rows.rs:211-238).
HD void inverse_symmetric(const float m[3][3], f3 diag, float o[3][3]) {
  float a = m[0][0], b = m[0][1], c = m[0][2];
  float d = m[1][1], e = m[1][2], f = m[2][2];
  float det = a * (d * f - e * e) - b * (b * f - c * e) + c * (b * e - c * d);
  if (fabsf(det) < 1e-12f) {
    float ix = diag.x > 0.0f ? 1.0f / diag.x : 0.0f;
    float iy = diag.y > 0.0f ? 1.0f / diag.y : 0.0f;
    float iz = diag.z > 0.0f ? 1.0f / diag.z : 0.0f;
    o[0][0] = ix; o[0][1] = 0.0f; o[0][2] = 0.0f;
    o[1][0] = 0.0f; o[1][1] = iy; o[1][2] = 0.0f;
    o[2][0] = 0.0f; o[2][1] = 0.0f; o[2][2] = iz;
    return;
  }
  float id = 1.0f / det;
  o[0][0] = (d * f - e * e) * id;
  o[0][1] = (c * e - b * f) * id;
  o[0][2] = (b * e - c * d) * id;
  o[1][0] = (c * e - b * f) * id;
  o[1][1] = (a * f - c * c) * id;
  o[1][2] = (b * c - a * e) * id;
  o[2][0] = (b * e - c * d) * id;
  o[2][1] = (b * c - a * e) * id;
  o[2][2] = (a * d - b * b) * id;
}

// Load / store helpers for float4-packed arrays.
HD f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
HD qf asq(float4 v) { return qf{v.x, v.y, v.z, v.w}; }
HD float4 pk(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
HD float4 pkq(qf q) { return make_float4(q.x, q.y, q.z, q.w); }
