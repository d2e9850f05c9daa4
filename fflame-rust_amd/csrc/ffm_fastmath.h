// ffm_fastmath.h -- single-precision sin/cos/atan2 for the walker's hot path.
//
// sin/cos: Cody-Waite reduction by pi/2 with the constant split into three
// floats (pi/2 rounded to float, then the rounded remainders; 72 bits in
// all), applied with fused multiply-adds: for |x| <= 32768 the first
// product-difference is exact (its result lies on x's grid and is below 1),
// so the reduced argument carries two roundings and the split's 2^-76
// error. Then the Cephes minimax polynomials on [-pi/4, pi/4]:
//   sin r = r + r^3 (-1.6666654611e-1 + r^2 (8.3321608736e-3 - r^2 1.9515295891e-4))
//   cos r = 1 - r^2/2 + r^4 (4.166664568298827e-2 + r^2 (-1.388731625493765e-3
//                                                  + r^2 2.443315711809948e-5))
// and the quadrant k mod 4 selects sign and function. Larger arguments
// (rare on the chaos game's attractors) are reduced in f64 in a separate
// non-inlined function, then take the same polynomials.
//
// atan2: a = min(|x|,|y|) / max(|x|,|y|) in [0, 1] with the quotient's
// residual a_lo = (min - a max) / max, then one odd polynomial on the whole
// of [0, 1]: atan a = a + a z P(z), z = a^2, P of degree 7 fitted here for
// the least maximum relative error of atan (2.8e-1 ulp before rounding),
// plus a_lo / (1 + z) (taken as a_lo (1 - z/2): the residual is at most
// half an ulp of a, so the factor's 12% error is invisible); octant
// fix-ups with two-part pi/2 and pi.
//
// Measured on the host build (the same text as C, IEEE divisions) against
// double-precision libm: sin and cos within 1.55 ulp for |x| <= 32768,
// atan2 within 1.53 ulp (tests/test_fastmath.py).
//
// The same text compiles as C (for the host accuracy check) and as HIP.
#pragma once

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define FFM_FM_INLINE __device__ __forceinline__
#define FFM_FM_NOINLINE __device__ __noinline__
#else
#include <math.h>
#define FFM_FM_INLINE static inline
#define FFM_FM_NOINLINE static
#endif

// Reciprocal and quotient from the hardware's 1-ulp v_rcp_f32 refined by
// Newton-Raphson with fused multiply-adds: r' = r + r (1 - b r), then the
// quotient's residual correction q' = q + r' (a - b q). Four FMAs after the
// rcp instead of the division expansion's scaling/fixup sequence; for
// operands away from the exponent extremes (|b| within 2^+-126, where the
// walk lives) the quotient is within 0.5 ulp plus the residual's rounding.
// Square roots: ffm_sqrt is the hardware's 1-ulp v_sqrt_f32 (WGSL's own
// sqrt bound is 2.5 ulp), where a root scales one output; ffm_sqrt_cr is
// correctly rounded (the compiler's expansion: v_sqrt_f32 and a residual
// test of its neighbours), as the IEEE oracle's, where a root's last bit
// is amplified: disc's pi r is rounded at magnitude up to 4, so a 1-ulp root
// there moves the angle by a whole ulp of 4; horseshoe divides both of its
// outputs by r, and a step summing it measured 4 ulp with the fast root.
// The host build (accuracy checks) uses the IEEE operations.
FFM_FM_INLINE float ffm_rcp(float b) {
#ifdef __HIPCC__
  const float r = __builtin_amdgcn_rcpf(b);
  return fmaf(r, fmaf(-b, r, 1.0f), r);
#else
  return 1.0f / b;
#endif
}

FFM_FM_INLINE float ffm_fdiv(float a, float b) {
#ifdef __HIPCC__
  const float r = ffm_rcp(b);
  const float q = a * r;
  return fmaf(r, fmaf(-b, q, a), q);
#else
  return a / b;
#endif
}

FFM_FM_INLINE float ffm_sqrt(float x) {
#ifdef __HIPCC__
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}

FFM_FM_INLINE float ffm_sqrt_cr(float x) { return sqrtf(x); }

// Sign and function by the quadrant k mod 4 from the Cephes polynomials of
// the reduced argument r in [-pi/4, pi/4].
FFM_FM_INLINE void ffm_sincos_poly(float r, int q, float* s, float* c) {
  const float r2 = r * r;
  const float sp = fmaf(fmaf(fmaf(-1.9515295891e-4f, r2, 8.3321608736e-3f), r2,
                             -1.6666654611e-1f),
                        r2 * r, r);
  const float cp = fmaf(fmaf(fmaf(2.443315711809948e-5f, r2, -1.388731625493765e-3f), r2,
                             4.166664568298827e-2f),
                        r2 * r2, fmaf(-0.5f, r2, 1.0f));
  const float ss = (q & 1) ? cp : sp;
  const float cc = (q & 1) ? sp : cp;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

// Past |x| = 32768 the reduction runs in f64 (two-part pi/2, fused
// multiply-adds) before the same polynomials: no library call, so device
// and host give the same bits; infinities and NaN give NaN. Kept out of
// line: such arguments are rare on the chaos game's attractors.
FFM_FM_NOINLINE void ffm_sincos_wide(float x, float* s, float* c) {
  if (!(fabsf(x) <= 3.4028235e38f)) {
    *s = *c = x - x;
    return;
  }
  const double xd = (double)x, k = rint(xd * 0.63661977236758134);
  double rd = fma(-k, 1.5707963267948966, xd);
  rd = fma(-k, 6.123233995736766e-17, rd);
  ffm_sincos_poly((float)rd, (int)fmod(k, 4.0) & 3, s, c);
}

FFM_FM_INLINE void ffm_sincos_reduced(float x, float* s, float* c) {
  const float k = rintf(x * 0.636619772367581343f);
  float r = fmaf(-k, 1.57079637050628662109375f, x);
  r = fmaf(-k, -4.37113882867379300e-8f, r);
  r = fmaf(-k, -1.71512451000588190e-15f, r);
  ffm_sincos_poly(r, (int)k & 3, s, c);
}

FFM_FM_INLINE void ffm_sincosf(float x, float* s, float* c) {
  if (fabsf(x) <= 32768.0f) {
    ffm_sincos_reduced(x, s, c);
  } else {
    ffm_sincos_wide(x, s, c);
  }
}

FFM_FM_INLINE float ffm_atan2f(float y, float x) {
  // pi/2 and pi as float + remainder (two-part constants)
  const float PIO2_HI = 1.57079637050628662109375f, PIO2_LO = -4.37113883e-8f;
  const float PI_HI = 3.1415927410125732421875f, PI_LO = -8.74227766e-8f;
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  float a = 0.0f, a_lo = 0.0f;
  if (mx > 0.0f) {
#ifdef __HIPCC__
    const float rc = ffm_rcp(mx);
    const float q = mn * rc;
    a = fmaf(rc, fmaf(-mx, q, mn), q);
    a_lo = fmaf(-mx, a, mn) * rc;
#else
    a = mn / mx;
    a_lo = fmaf(-mx, a, mn) / mx;
#endif
  }
  const float z = a * a;
  float p = fmaf(2.920681611e-03f, z, -1.636788808e-02f);
  p = fmaf(p, z, 4.321179539e-02f);
  p = fmaf(p, z, -7.552208751e-02f);
  p = fmaf(p, z, 1.066600233e-01f);
  p = fmaf(p, z, -1.421105415e-01f);
  p = fmaf(p, z, 1.999377310e-01f);
  p = fmaf(p, z, -3.333315253e-01f);
  float t = a + fmaf(a * z, p, a_lo * fmaf(-0.5f, z, 1.0f));
  if (ay > ax) t = (PIO2_HI - t) + PIO2_LO;
#ifdef __HIPCC__
  const bool xneg = (__float_as_uint(x) & 0x80000000u) != 0u;
  const bool yneg = (__float_as_uint(y) & 0x80000000u) != 0u;
#else
  const int xneg = signbit(x) != 0, yneg = signbit(y) != 0;
#endif
  if (xneg) t = (PI_HI - t) + PI_LO;
  return yneg ? -t : t;
}

FFM_FM_INLINE float ffm_sinf(float x) {
  float s, c;
  ffm_sincosf(x, &s, &c);
  return s;
}
