/*
 * oavbd_fmath.h — portable single-precision sin/cos/asin/atan2 for the AVBD step.
 *
 * The AVBD rotation helpers need sin/cos (exponential-map integration,
 * crates/physics/src/avbd/rows.rs:110-119) and asin (angle-axis recovery,
 * crates/physics/src/avbd/rows.rs:128-146).  The reference calls the
 * platform libm through Rust's f32::sin_cos / f32::asin.  A GPU step and
 * its CPU checker only produce identical trajectories when both evaluate
 * these functions with the same float operations, so this header defines
 * them once, from first principles, using nothing but IEEE-754 single
 * precision +, -, *, / and sqrt (all correctly rounded on x86-64 SSE and
 * on gfx950 under hipcc's default correctly-rounded divide/sqrt):
 *
 *   - range reduction: Cody & Waite (1980) three-part split of pi/2;
 *     the first two parts carry 12 significant bits, so k*C1 and k*C2 are
 *     exact for |k| < 4096 (|x| below ~6400 rad);
 *   - sin/cos on [-pi/4, pi/4]: the Maclaurin series truncated after the
 *     x^11 (sin) and x^10 (cos) terms, whose truncation error there is
 *     below 3e-9 relative; cos uses the w + ((1 - w) - r^2/2) split to keep
 *     the leading terms exact;
 *   - asin: Maclaurin series of asin (coefficients (2k)!/(4^k (k!)^2 (2k+1)))
 *     through x^23 on |x| <= 1/2, and asin(x) = pi/2 - 2 asin(sqrt((1-x)/2))
 *     above, so the series argument never exceeds 1/2 (truncation < 3e-9).
 *
 *   - atan2: the quadrant and octant are reduced away exactly (|y| <= |x|
 *     after a swap), then t = |y|/|x| in [0, 1] is moved next to zero with
 *     the addition formula atan(t) = atan(c) + atan((t - c) / (1 + t c)):
 *     c = 0 below 7/16, c = 1/2 up to 11/16 (u = (2t - 1) / (2 + t), whose
 *     numerator is exact there), c = 1 above (u = (t - 1) / (t + 1), exact
 *     numerator), so |u| <= 7/16; the Maclaurin series of atan through
 *     u^23 (truncation below 1e-9 relative); atan(1/2), pi/4, pi/2 and pi
 *     are carried as float high + low parts.
 *
 * Measured accuracy against double-precision libm is pinned by the test
 * suite (tests/test_fmath.py).  Compile every consumer with
 * -ffp-contract=off: a fused multiply-add changes the bits.
 */
#ifndef OAVBD_FMATH_H
#define OAVBD_FMATH_H

#ifdef __HIPCC__
#define OAVBD_HD __host__ __device__
#else
#define OAVBD_HD
#endif

#include <math.h>

#ifdef __cplusplus
#define OAVBD_INLINE static inline
#else
#define OAVBD_INLINE static inline
#endif

/* pi/2 = C1 + C2 + C3 (Cody-Waite).  C1 = 1.5703125, C2 = 0x39fda000. */
#define OAVBD_PIO2_C1 1.5703125f
#define OAVBD_PIO2_C2 4.837512969970703125e-4f
#define OAVBD_PIO2_C3 7.549790126404332e-08f
#define OAVBD_TWO_OVER_PI 0.636619772367581343f
/* pi/2 = HI + LO with HI the nearest float. */
#define OAVBD_PIO2_HI 1.57079637050628662109f
#define OAVBD_PIO2_LO (-4.37113882867379300296e-08f)

/* sin(r) and cos(r) for |r| <= pi/4 (Maclaurin). */
OAVBD_INLINE OAVBD_HD float oavbd_sin_kernel(float r) {
  const float s1 = (float)(-1.0 / 6.0);
  const float s2 = (float)(1.0 / 120.0);
  const float s3 = (float)(-1.0 / 5040.0);
  const float s4 = (float)(1.0 / 362880.0);
  const float s5 = (float)(-1.0 / 39916800.0);
  float z = r * r;
  float p = s5;
  p = p * z + s4;
  p = p * z + s3;
  p = p * z + s2;
  p = p * z + s1;
  return r + (r * z) * p;
}

OAVBD_INLINE OAVBD_HD float oavbd_cos_kernel(float r) {
  const float c2 = (float)(1.0 / 24.0);
  const float c3 = (float)(-1.0 / 720.0);
  const float c4 = (float)(1.0 / 40320.0);
  const float c5 = (float)(-1.0 / 3628800.0);
  float z = r * r;
  float q = c5;
  q = q * z + c4;
  q = q * z + c3;
  q = q * z + c2;
  float hz = 0.5f * z;
  float w = 1.0f - hz;
  return w + (((1.0f - w) - hz) + (z * z) * q);
}

/* sin and cos of x (radians). */
OAVBD_INLINE OAVBD_HD void oavbd_sincosf(float x, float* s, float* c) {
  float kf = floorf(x * OAVBD_TWO_OVER_PI + 0.5f);
  float r = ((x - kf * OAVBD_PIO2_C1) - kf * OAVBD_PIO2_C2) - kf * OAVBD_PIO2_C3;
  float sr = oavbd_sin_kernel(r);
  float cr = oavbd_cos_kernel(r);
  /* quadrant = k mod 4, taken from the float value (|k| < 2^24). */
  float kq = kf - 4.0f * floorf(kf * 0.25f);
  int q = (int)kq;
  if (q == 0) {
    *s = sr;
    *c = cr;
  } else if (q == 1) {
    *s = cr;
    *c = -sr;
  } else if (q == 2) {
    *s = -sr;
    *c = -cr;
  } else {
    *s = -cr;
    *c = sr;
  }
}

/* asin(x) for |x| <= 1/2 without the sign handling. */
OAVBD_INLINE OAVBD_HD float oavbd_asin_series(float a) {
  const float k1 = (float)(1.0 / 6.0);
  const float k2 = (float)(3.0 / 40.0);
  const float k3 = (float)(5.0 / 112.0);
  const float k4 = (float)(35.0 / 1152.0);
  const float k5 = (float)(63.0 / 2816.0);
  const float k6 = (float)(231.0 / 13312.0);
  const float k7 = (float)(143.0 / 10240.0);
  const float k8 = (float)(6435.0 / 557056.0);
  const float k9 = (float)(12155.0 / 1245184.0);
  const float k10 = (float)(46189.0 / 5505024.0);
  const float k11 = (float)(88179.0 / 12058624.0);
  float z = a * a;
  float p = k11;
  p = p * z + k10;
  p = p * z + k9;
  p = p * z + k8;
  p = p * z + k7;
  p = p * z + k6;
  p = p * z + k5;
  p = p * z + k4;
  p = p * z + k3;
  p = p * z + k2;
  p = p * z + k1;
  return a + (a * z) * p;
}

/* asin(x) for x in [-1, 1]. */
OAVBD_INLINE OAVBD_HD float oavbd_asinf(float x) {
  float a = fabsf(x);
  float y;
  if (a <= 0.5f) {
    y = oavbd_asin_series(a);
  } else {
    float z = (1.0f - a) * 0.5f;
    float t = sqrtf(z);
    float u = oavbd_asin_series(t);
    y = (OAVBD_PIO2_HI - 2.0f * u) + OAVBD_PIO2_LO;
  }
  return x < 0.0f ? -y : y;
}

/* pi/4 and pi as float high + low parts. */
#define OAVBD_PIO4_HI 0.785398185253143310546875f
#define OAVBD_PIO4_LO (-2.1855694143368964e-08f)
#define OAVBD_PI_HI 3.1415927410125732421875f
#define OAVBD_PI_LO (-8.742277657347586e-08f)

/* atan(u) for |u| <= 7/16 (Maclaurin through u^23). */
OAVBD_INLINE OAVBD_HD float oavbd_atan_series(float u) {
  const float a1 = (float)(-1.0 / 3.0);
  const float a2 = (float)(1.0 / 5.0);
  const float a3 = (float)(-1.0 / 7.0);
  const float a4 = (float)(1.0 / 9.0);
  const float a5 = (float)(-1.0 / 11.0);
  const float a6 = (float)(1.0 / 13.0);
  const float a7 = (float)(-1.0 / 15.0);
  const float a8 = (float)(1.0 / 17.0);
  const float a9 = (float)(-1.0 / 19.0);
  const float a10 = (float)(1.0 / 21.0);
  const float a11 = (float)(-1.0 / 23.0);
  float z = u * u;
  float p = a11;
  p = p * z + a10;
  p = p * z + a9;
  p = p * z + a8;
  p = p * z + a7;
  p = p * z + a6;
  p = p * z + a5;
  p = p * z + a4;
  p = p * z + a3;
  p = p * z + a2;
  p = p * z + a1;
  return u + (u * z) * p;
}

/* atan(1/2) as float high + low parts. */
#define OAVBD_ATANHALF_HI 0.46364760398864746f
#define OAVBD_ATANHALF_LO 5.01215868808913e-09f

/* atan(t) for t in [0, 1]. */
OAVBD_INLINE OAVBD_HD float oavbd_atan01(float t) {
  if (t < 0.4375f) return oavbd_atan_series(t);
  if (t < 0.6875f) {
    float u = (2.0f * t - 1.0f) / (2.0f + t);
    return OAVBD_ATANHALF_HI + (oavbd_atan_series(u) + OAVBD_ATANHALF_LO);
  }
  float u = (t - 1.0f) / (t + 1.0f);
  return OAVBD_PIO4_HI + (oavbd_atan_series(u) + OAVBD_PIO4_LO);
}

/* atan2(y, x) with the IEEE signed-zero and infinity conventions of C's
 * atan2f; NaN in, NaN out. */
OAVBD_INLINE OAVBD_HD float oavbd_atan2f(float y, float x) {
  if (x != x || y != y) return x + y;
  float ax = fabsf(x), ay = fabsf(y);
  float a;
  if (ax == 0.0f && ay == 0.0f) {
    a = 0.0f;
  } else if (isinf(ax) && isinf(ay)) {
    a = OAVBD_PIO4_HI;
  } else if (ay <= ax) {
    a = oavbd_atan01(ay / ax);
  } else {
    a = (OAVBD_PIO2_HI - oavbd_atan01(ax / ay)) + OAVBD_PIO2_LO;
  }
  /* x < 0 reflects about pi/2; of the zeros only atan2(+-0, -0) does */
  if (x < 0.0f || (a == 0.0f && signbit(x))) a = (OAVBD_PI_HI - a) + OAVBD_PI_LO;
  return copysignf(a, y);
}

#endif /* OAVBD_FMATH_H */
