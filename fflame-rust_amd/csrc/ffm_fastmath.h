// ffm_fastmath.h -- single-precision sin/cos for the walker's hot path.
//
// Method: Cody-Waite reduction by pi/2 with a three-part constant
// (1.5703125 + 4.837512969970703125e-4 + 7.54978995489188216e-8, as in
// Moshier's Cephes sinf/cosf, applied with fused multiply-adds), then the
// Cephes minimax polynomials on [-pi/4, pi/4]:
//   sin r = r + r^3 (-1.6666654611e-1 + r^2 (8.3321608736e-3 - r^2 1.9515295891e-4))
//   cos r = 1 - r^2/2 + r^4 (4.166664568298827e-2 + r^2 (-1.388731625493765e-3
//                                                  + r^2 2.443315711809948e-5))
// and the quadrant k mod 4 selects sign and function. C1 has 9 significant
// bits, so k*C1 is exact for |k| < 2^15: the fast path serves |x| <= 32768;
// larger arguments (rare on the chaos game's attractors) take the library's
// full-range sincosf in a separate non-inlined function, keeping its
// Payne-Hanek reduction out of the hot loop's registers.
//
// The same text compiles as C (for a host accuracy check) and as HIP.
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
// The square root is the hardware's 1-ulp v_sqrt_f32 (WGSL's own sqrt bound
// is 2.5 ulp). The host build (accuracy checks) uses the IEEE operations.
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

FFM_FM_NOINLINE void ffm_sincos_wide(float x, float* s, float* c) {
#ifdef __HIPCC__
  sincosf(x, s, c);
#else
  *s = sinf(x);
  *c = cosf(x);
#endif
}

FFM_FM_INLINE void ffm_sincos_reduced(float x, float* s, float* c) {
  const float k = rintf(x * 0.636619772367581343f);
  float r = fmaf(-k, 1.5703125f, x);
  r = fmaf(-k, 4.837512969970703125e-4f, r);
  r = fmaf(-k, 7.54978995489188216e-8f, r);
  const float r2 = r * r;
  const float sp = fmaf(fmaf(fmaf(-1.9515295891e-4f, r2, 8.3321608736e-3f), r2,
                             -1.6666654611e-1f),
                        r2 * r, r);
  const float cp = fmaf(fmaf(fmaf(2.443315711809948e-5f, r2, -1.388731625493765e-3f), r2,
                             4.166664568298827e-2f),
                        r2 * r2, fmaf(-0.5f, r2, 1.0f));
  const int q = (int)k & 3;
  const float ss = (q & 1) ? cp : sp;
  const float cc = (q & 1) ? sp : cp;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

FFM_FM_INLINE void ffm_sincosf(float x, float* s, float* c) {
  if (fabsf(x) <= 32768.0f) {
    ffm_sincos_reduced(x, s, c);
  } else {
    ffm_sincos_wide(x, s, c);
  }
}

// atan2 by reduction to a = min(|x|,|y|) / max(|x|,|y|) in [0, 1] (1.5-ulp
// quotients), then
// a -> (a - 1) / (a + 1) above tan(pi/8) (adding pi/4), and Cephes atanf's
// minimax polynomial on [-tan(pi/8), tan(pi/8)]:
//   atan z = z + z^3 (-3.33329491539e-1 + z^2 (1.99777106478e-1
//                     + z^2 (-1.38776856032e-1 + z^2 8.05374449538e-2)))
// Octant fix-ups use the operands' sign bits, so the signed-zero cases
// follow IEEE atan2: atan2(+-0, +0) = +-0, atan2(+-0, -0) = +-pi.
FFM_FM_INLINE float ffm_atan2f(float y, float x) {
  // pi/4, pi/2, pi as float + remainder (two-part constants)
  const float PIO4_HI = 0.785398185253143310546875f, PIO4_LO = -2.18556941e-8f;
  const float PIO2_HI = 1.57079637050628662109375f, PIO2_LO = -4.37113883e-8f;
  const float PI_HI = 3.1415927410125732421875f, PI_LO = -8.74227766e-8f;
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  float a = mx > 0.0f ? ffm_fdiv(mn, mx) : 0.0f;
  const int reduced = a > 0.41421356237309503f;
  if (reduced) a = ffm_fdiv(a - 1.0f, a + 1.0f);
  const float z = a * a;
  const float p = fmaf(fmaf(fmaf(8.05374449538e-2f, z, -1.38776856032e-1f), z, 1.99777106478e-1f),
                       z, -3.33329491539e-1f);
  const float q = p * z * a;  // atan(a) = a + q
  float t = reduced ? PIO4_HI + (a + (q + PIO4_LO)) : a + q;
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
