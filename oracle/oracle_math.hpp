// oracle_math.hpp — vector/quaternion arithmetic of the CPU oracle.
//
// TEST INFRASTRUCTURE: this directory is the CPU checker of the AVBD step.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it; the product library never includes or links anything here.
//
// Restates, operation for operation, the glam 0.34.3 arithmetic the
// reference AVBD engine runs on (Cargo.lock:1357-1358; call sites e.g.
// crates/physics/src/avbd/rows.rs:118, :196): Vec3 is glam's scalar struct,
// Quat follows glam's SSE2 code path (the x86-64 CI target that checks the
// reference's golden snapshot).  Every sum is written left to right in the
// order glam evaluates it; build with -ffp-contract=off.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

#include "../include/oavbd_fmath.h"

namespace orc {

template <class T>
struct V3 {
  T x, y, z;
  T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

template <class T>
struct Q {
  T x, y, z, w;
};

template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> cmul(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline bool operator==(V3<T> a, V3<T> b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
template <class T> inline bool operator!=(V3<T> a, V3<T> b) { return !(a == b); }

// glam Vec3::dot: (x*x' + y*y') + z*z'.
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// glam Vec3::cross.
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> inline T length_sq(V3<T> a) { return dot(a, a); }
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// glam Vec3::normalize_or: multiply by the reciprocal length, fall back when
// that reciprocal is not a finite positive number.
template <class T> inline V3<T> normalize_or(V3<T> a, V3<T> fb) {
  T rcp = T(1) / length(a);
  if (std::isfinite(rcp) && rcp > T(0)) return a * rcp;
  return fb;
}
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / length(a)); }
template <class T> inline V3<T> vabs(V3<T> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
// f32::max / f32::min as x86-64 libm evaluates them (fmaxf/fminf): the
// first operand when strictly larger (smaller) or when the second is NaN,
// else the second, so (+0, -0) yields -0 and (-0, +0) yields +0.
template <class T> inline T fmax_(T a, T b) { return (a > b || b != b) ? a : b; }
template <class T> inline T fmin_(T a, T b) { return (a < b || b != b) ? a : b; }
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) { return {fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
// glam Vec3::clamp = max then min.
template <class T> inline V3<T> vclamp(V3<T> a, V3<T> lo, V3<T> hi) { return vmin(vmax(a, lo), hi); }
// Rust f32::clamp (comparisons, NaN passes through).
template <class T> inline T clampf(T x, T lo, T hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
template <class T> inline T min_element(V3<T> a) { return fmin_(a.x, fmin_(a.y, a.z)); }
// Rust f32::signum.
template <class T> inline T signum(T x) { return std::isnan(x) ? x : std::copysign(T(1), x); }
template <class T> inline bool finite3(V3<T> a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

template <class T> inline Q<T> qident() { return Q<T>{T(0), T(0), T(0), T(1)}; }
template <class T> inline Q<T> conj(Q<T> q) { return Q<T>{-q.x, -q.y, -q.z, q.w}; }
template <class T> inline bool operator==(Q<T> a, Q<T> b) { return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w; }
template <class T> inline bool operator!=(Q<T> a, Q<T> b) { return !(a == b); }

// glam Quat * Vec3 (SSE2 mul_vec3a):
//   v*(w*w - b.b) + b*((v.b)*2) + (b x v)*(w*2),   b = (x, y, z).
template <class T> inline V3<T> rotate(Q<T> q, V3<T> v) {
  V3<T> b{q.x, q.y, q.z};
  T b2 = dot(b, b);
  V3<T> t0 = v * (q.w * q.w - b2);
  V3<T> t1 = b * (dot(v, b) * T(2));
  V3<T> t2 = cross(b, v) * (q.w * T(2));
  return (t0 + t1) + t2;
}

// glam Quat::normalize (SSE2 Vec4::normalize): divide by sqrt of the
// lane-paired dot product (x*x + z*z) + (y*y + w*w).
template <class T> inline Q<T> qnormalize(Q<T> q) {
  T d = (q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w);
  T len = std::sqrt(d);
  return Q<T>{q.x / len, q.y / len, q.z / len, q.w / len};
}

// glam Mat3::from_quat columns.
template <class T> struct M3c {
  V3<T> c[3];
};
template <class T> inline M3c<T> mat3_from_quat(Q<T> q) {
  T x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
  T xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
  T yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
  T wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
  M3c<T> m;
  m.c[0] = V3<T>{T(1) - (yy + zz), xy + wz, xz - wy};
  m.c[1] = V3<T>{xy - wz, T(1) - (xx + zz), yz + wx};
  m.c[2] = V3<T>{xz + wy, yz - wx, T(1) - (xx + yy)};
  return m;
}

// Transcendentals: libm (what the reference links on Linux) or the
// portable float kernels the device shares (include/oavbd_fmath.h).
enum MathMode { MATH_LIBM = 0, MATH_PORTABLE = 1 };

inline void sincos_t(float x, float* s, float* c, int mode) {
  if (mode == MATH_PORTABLE) {
    oavbd_sincosf(x, s, c);
  } else {
    *s = ::sinf(x);
    *c = ::cosf(x);
  }
}
inline void sincos_t(double x, double* s, double* c, int) {
  *s = std::sin(x);
  *c = std::cos(x);
}
inline float asin_t(float x, int mode) { return mode == MATH_PORTABLE ? oavbd_asinf(x) : ::asinf(x); }
inline double asin_t(double x, int) { return std::asin(x); }
inline float atan2_t(float y, float x, int mode) { return mode == MATH_PORTABLE ? oavbd_atan2f(y, x) : ::atan2f(y, x); }
inline double atan2_t(double y, double x, int) { return std::atan2(y, x); }
// Rust f32::rem_euclid: the remainder of fmod, moved into [0, |m|).
template <class T> inline T rem_euclid(T x, T m) {
  T r = std::fmod(x, m);
  return r < T(0) ? r + std::fabs(m) : r;
}

}  // namespace orc
