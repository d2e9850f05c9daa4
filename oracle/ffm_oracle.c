/*
 * ffm_oracle.c -- CPU restatement of the reference's flame path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg. The product
 * library (libffm.so) never links, loads or calls it.
 *
 * Build: -O2 -ffp-contract=off (no FMA contraction, no fast-math), so each
 * FP32 operation rounds once, in the order the reference's WGSL writes it.
 * Transcendentals: the reference's are whatever the GPU driver compiles,
 * which no reference test pins. The variation bodies' sin, cos and atan2
 * are a fixed method (below, at orc_use_libm), itself pinned against
 * double-precision libm; everything else is the host libm's (powf, expf,
 * logf, and sinf/cosf/atan2f under orc_use_libm(1)).
 *
 * Pinning (what checks each function against something other than itself):
 *  - RNG: published PCG-RXS-M-XS constants plus an independent numpy
 *    restatement and its period/bijectivity properties (tests).
 *  - select / affine / colour / world_to_pixel / splat / accumulate /
 *    OrbitFuses: the reference's known-answer tests
 *    (src/renderer/compute_kernel.rs:6056-6144,6404-6508,6997-7120) and
 *    brute-force/closed-form checks.
 *  - variations: closed forms evaluated independently in FP64 (tests);
 *    their sin/cos/atan2 method: a sweep against double-precision libm.
 *  - tonemap: the reference's own rendered baseline
 *    tests/visual/baseline/2d-simple-linear.png (golden fixture) through
 *    the full render.
 *  - density estimation: no reference implementation exists
 *    (src/flame_xml.rs:1488-1490); pinned by mass conservation, constant
 *    fields and brute force only.
 */
#include "ffm_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define F32(x) ((float)(x))

/* ------------------------------------------------------------------ RNG */

/* shaders/core/rng.wgsl:22-27 */
uint32_t orc_pcg_hash(uint32_t input) {
  uint32_t state = input * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

/* shaders/core/rng.wgsl:29-37 */
void orc_rng_init(orc_rng* r, uint32_t thread_id, uint32_t seed) {
  r->state = orc_pcg_hash(thread_id ^ seed);
  r->inc = (orc_pcg_hash(orc_pcg_hash(thread_id) + seed) << 1u) | 1u;
}

/* shaders/core/rng.wgsl:39-45 */
uint32_t orc_rng_next(orc_rng* r) {
  uint32_t old = r->state;
  r->state = old * 747796405u + r->inc;
  uint32_t xs = ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
  return (xs >> 22u) ^ xs;
}

/* shaders/core/rng.wgsl:53-55: top 24 bits times 2^-24 */
float orc_rng_nextf(orc_rng* r) { return (float)(orc_rng_next(r) >> 8u) * (1.0f / 16777216.0f); }

void orc_rng_stream(uint32_t tid, uint32_t seed, uint32_t n, uint32_t* out) {
  orc_rng r;
  orc_rng_init(&r, tid, seed);
  for (uint32_t i = 0; i < n; i++) out[i] = orc_rng_next(&r);
}

/* ------------------------------------------------------------ selection */

/* shaders/core/utilities.wgsl:35-55 select_transform_const: sequential f32
 * sums, first i with target <= cumulative, else the last. */
int orc_select(const ffm_flame_desc* f, float rand_val) {
  float total = 0.0f;
  for (int i = 0; i < f->n_xforms; i++) total += f->xforms[i].weight;
  float target = rand_val * total;
  float cum = 0.0f;
  for (int i = 0; i < f->n_xforms; i++) {
    cum += f->xforms[i].weight;
    if (target <= cum) return i;
  }
  return f->n_xforms - 1;
}

/* shaders/core/utilities.wgsl:107-140 select_transform_xaos: the total and
 * the running sum of weight[i] * xaos[prev][i], each product rounded, in
 * order; first i with rand * total <= cumulative, else the last. */
int orc_select_xaos(const ffm_flame_desc* f, float rand_val, int prev) {
  float total = 0.0f;
  for (int i = 0; i < f->n_xforms; i++) total += f->xforms[i].weight * f->xaos[prev][i];
  float threshold = rand_val * total;
  float cum = 0.0f;
  for (int i = 0; i < f->n_xforms; i++) {
    cum += f->xforms[i].weight * f->xaos[prev][i];
    if (threshold <= cum) return i;
  }
  return f->n_xforms - 1;
}

/* src/scene/xaos.rs:64-124: transitions i -> j where weight_j * xaos[i][j]
 * > 0 (none: to 0); reachability by depth-first search; the live set
 * reachable from a first pick (weight > 0); closed groups (members reach
 * only what reaches them back), each named by its lowest member; a split
 * when some first pick reaches two groups. */
uint32_t orc_xaos_refresh_period(const ffm_flame_desc* f) {
  int n = f->n_xforms;
  if (!f->has_xaos || n <= 0) return 0;
  char reach[FFM_MAX_XAOS][FFM_MAX_XAOS];
  memset(reach, 0, sizeof reach);
  for (int s0 = 0; s0 < n; s0++) {
    int stack[FFM_MAX_XAOS * FFM_MAX_XAOS], top = 0;
    stack[top++] = s0;
    reach[s0][s0] = 1;
    while (top) {
      int i = stack[--top];
      int any = 0;
      for (int j = 0; j < n; j++) {
        float wj = f->xforms[j].weight > 0.0f ? f->xforms[j].weight : 0.0f;
        if (wj * f->xaos[i][j] > 0.0f) {
          any = 1;
          if (!reach[s0][j]) {
            reach[s0][j] = 1;
            stack[top++] = j;
          }
        }
      }
      if (!any && !reach[s0][0]) {
        reach[s0][0] = 1;
        stack[top++] = 0;
      }
    }
  }
  char live[FFM_MAX_XAOS] = {0};
  for (int i = 0; i < n; i++)
    if (f->xforms[i].weight > 0.0f)
      for (int j = 0; j < n; j++) live[j] |= reach[i][j];
  int closed[FFM_MAX_XAOS], nc = 0;
  for (int i = 0; i < n; i++) {
    if (!live[i]) continue;
    int ok = 1;
    for (int j = 0; j < n && ok; j++) {
      if (reach[i][j] && !reach[j][i]) ok = 0;
      if (j < i && reach[i][j] && reach[j][i]) ok = 0;
    }
    if (ok) closed[nc++] = i;
  }
  if (nc < 2) return 0;
  for (int i = 0; i < n; i++) {
    if (!(f->xforms[i].weight > 0.0f)) continue;
    int hit = 0;
    for (int g = 0; g < nc; g++) hit += reach[i][closed[g]];
    if (hit > 1) return 4;
  }
  return 0;
}

/* ----------------------------------------------------------- variations */

static const float PI_F = F32(3.14159265359);  /* the WGSL literal */

/* shaders/core/utilities.wgsl:604-611 ff_atan2: signed-zero origin handled
 * by sign bits, else atan2. */
static float v_atan2(float y, float x);
static float ff_atan2(float y, float x) {
  if (y == 0.0f && x == 0.0f) {
    const float pi = F32(3.14159265358979);
    uint32_t bx, by;
    memcpy(&bx, &x, 4);
    memcpy(&by, &y, 4);
    float mag = (bx & 0x80000000u) ? pi : 0.0f;
    return (by & 0x80000000u) ? -mag : mag;
  }
  return v_atan2(y, x);
}

static float len2(float x, float y) { return x * x + y * y; }

/* sin, cos and atan2 of the variation bodies. WGSL leaves their last bits
 * to the platform (the reference notes vendor differences,
 * src/probe/mod.rs:17-34), so the oracle fixes a method: the device's
 * (ffm_fastmath.h; restated here, and both pinned against double-precision
 * libm by tests/test_fastmath.py): Cody-Waite reduction by a three-float
 * pi/2 and the Cephes polynomials for sin/cos (orc_fx_sincos, which the
 * colour effects use too), and for atan2 the quotient min/max with its
 * residual and one degree-7 odd polynomial on [0, 1]. orc_use_libm(1)
 * switches the bodies to the host libm's sinf/cosf/atan2f instead. */
static int g_libm;
void orc_use_libm(int on) { g_libm = on; }

float orc_fm_atan2(float y, float x) {
  const float PIO2_HI = 1.57079637050628662109375f, PIO2_LO = -4.37113883e-8f;
  const float PI_HI = 3.1415927410125732421875f, PI_LO = -8.74227766e-8f;
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  float a = 0.0f, a_lo = 0.0f;
  if (mx > 0.0f) {
    a = mn / mx;
    a_lo = fmaf(-mx, a, mn) / mx;
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
  if (signbit(x)) t = (PI_HI - t) + PI_LO;
  return signbit(y) ? -t : t;
}

static float v_sin(float x) {
  if (g_libm) return sinf(x);
  float s, c;
  orc_fx_sincos(x, &s, &c);
  return s;
}
static float v_cos(float x) {
  if (g_libm) return cosf(x);
  float s, c;
  orc_fx_sincos(x, &s, &c);
  return c;
}
static float v_atan2(float y, float x) { return g_libm ? atan2f(y, x) : orc_fm_atan2(y, x); }

/* The 13 bodies. Sources: src/variations/defs/basic.rs (linear :37,
 * sinusoidal :100, spherical :130, swirl :162, horseshoe :198) and
 * src/variations/defs/advanced.rs (polar :32, disc :134, hyperbolic :213,
 * julia :329, waves :414, julian :478, bubble :639, curl :1686). */
void orc_variation(int var_id, const float* p, const ffm_xform* xf, float x, float y,
                   orc_rng* rng, float* ox, float* oy) {
  switch (var_id) {
    case FFM_VAR_LINEAR:
      *ox = x;
      *oy = y;
      return;
    case FFM_VAR_SINUSOIDAL:
      *ox = v_sin(x);
      *oy = v_sin(y);
      return;
    case FFM_VAR_SPHERICAL: {
      float r2 = len2(x, y) + 1e-6f;
      *ox = x / r2;
      *oy = y / r2;
      return;
    }
    case FFM_VAR_SWIRL: {
      float r2 = len2(x, y);
      float s = v_sin(r2), c = v_cos(r2);
      *ox = x * s - y * c;
      *oy = x * c + y * s;
      return;
    }
    case FFM_VAR_HORSESHOE: {
      float r = sqrtf(len2(x, y)) + 1e-6f;
      float r_inv = 1.0f / r;
      *ox = (x - y) * (x + y) * r_inv;
      *oy = 2.0f * x * y * r_inv;
      return;
    }
    case FFM_VAR_POLAR: {
      float theta = v_atan2(x, y);
      float r = sqrtf(len2(x, y));
      *ox = theta / PI_F;
      *oy = r - 1.0f;
      return;
    }
    case FFM_VAR_DISC: {
      float theta = v_atan2(x, y);
      float r = sqrtf(len2(x, y));
      float theta_pi = theta / PI_F;
      float pi_r = PI_F * r;
      *ox = theta_pi * v_sin(pi_r);
      *oy = theta_pi * v_cos(pi_r);
      return;
    }
    case FFM_VAR_HYPERBOLIC: {
      float r2 = len2(x, y) + 1e-6f;
      *ox = x / r2;
      *oy = y;
      return;
    }
    case FFM_VAR_JULIA: {
      float r = sqrtf(len2(x, y));
      float theta = v_atan2(y, x);
      float sqrt_r = sqrtf(r);
      float omega = (orc_rng_nextf(rng) < 0.5f) ? PI_F : 0.0f;
      float half_theta = theta / 2.0f + omega;
      *ox = sqrt_r * v_cos(half_theta);
      *oy = sqrt_r * v_sin(half_theta);
      return;
    }
    case FFM_VAR_WAVES:
      *ox = x + xf->b * v_sin(y / (xf->e * xf->e + 1e-6f));
      *oy = y + xf->d * v_sin(x / (xf->f * xf->f + 1e-6f));
      return;
    case FFM_VAR_JULIAN: {
      float power = p[0];
      float cpower = p[2];
      float r2 = len2(x, y);
      float r = powf(r2, cpower);
      float theta = ff_atan2(y, x);
      float trunc_val = floorf(fabsf(power) * orc_rng_nextf(rng));
      float t = (theta + F32(6.28318530718) * trunc_val) / power;
      *ox = r * v_cos(t);
      *oy = r * v_sin(t);
      return;
    }
    case FFM_VAR_BUBBLE: {
      float r2 = len2(x, y);
      float r = r2 / 4.0f + 1.0f;
      float scale = 1.0f / r;
      *ox = scale * x;
      *oy = scale * y;
      return;
    }
    case FFM_VAR_CURL: {
      float c1 = p[0], c2 = p[1];
      float re = 1.0f + c1 * x + c2 * (x * x - y * y);
      float im = c1 * y + 2.0f * c2 * x * y;
      float r = 1.0f / (re * re + im * im);
      *ox = (x * re + y * im) * r;
      *oy = (y * re - x * im) * r;
      return;
    }
    case FFM_VAR_BLOB: {
      /* advanced.rs BLOB: r = length(p), theta = ff_atan2(p.x, p.y) */
      float high = p[0], low = p[1], waves = p[2];
      float r = sqrtf(len2(x, y));
      float theta = ff_atan2(x, y);
      float scale = r * (low + ((high - low) / 2.0f) * (v_sin(waves * theta) + 1.0f));
      *ox = scale * v_cos(theta);
      *oy = scale * v_sin(theta);
      return;
    }
    case FFM_VAR_NOISE: {
      /* WGSL evaluates the draws left to right: theta first, then r */
      float theta = orc_rng_nextf(rng) * F32(6.28318530718);
      float r = orc_rng_nextf(rng);
      *ox = x * r * v_cos(theta);
      *oy = y * r * v_sin(theta);
      return;
    }
    case FFM_VAR_BLUR: {
      float theta = orc_rng_nextf(rng) * F32(6.28318530718);
      float r = orc_rng_nextf(rng);
      *ox = r * v_cos(theta);
      *oy = r * v_sin(theta);
      return;
    }
    case FFM_VAR_GAUSSIAN_BLUR: {
      float theta = orc_rng_nextf(rng) * F32(6.28318530718);
      float d1 = orc_rng_nextf(rng);
      float d2 = orc_rng_nextf(rng);
      float d3 = orc_rng_nextf(rng);
      float d4 = orc_rng_nextf(rng);
      float r = d1 + d2 + d3 + d4 - 2.0f;
      *ox = r * v_cos(theta);
      *oy = r * v_sin(theta);
      return;
    }
    default:
      *ox = 0.0f;
      *oy = 0.0f;
      return;
  }
}

/* init_julian (src/variations/defs/advanced.rs:467-474):
 * cpower = dist / max(|power|, 1e-30) / 2. */
void orc_flame_prepare(ffm_flame_desc* f) {
  for (int k = 0; k < f->n_vars; k++) {
    if (f->var_id[k] != FFM_VAR_JULIAN) continue;
    for (int i = 0; i < f->n_xforms + f->n_finals; i++) {
      ffm_xform* xf = i < f->n_xforms ? &f->xforms[i] : &f->finals[i - f->n_xforms];
      float power = xf->var_param[k][0], dist = xf->var_param[k][1];
      xf->var_param[k][2] = dist / fmaxf(fabsf(power), 1e-30f) / 2.0f;
    }
  }
}

/* The largest magnitude among the terms one step sums (the walker's point,
 * the affine point, each weighted variation output, the post-affine
 * products), recorded while orc_step_from runs: the scale at which the
 * step's FP32 sums round, for per-step error bars. */
static _Thread_local int g_track;
static _Thread_local float g_term;
static void track(float v) {
  if (g_track && fabsf(v) > g_term) g_term = fabsf(v);
}
float orc_step_term_scale(void) { return g_term; }

/* affine (shaders/core/affine.wgsl:4-19) -> variations in local-index
 * order, weight-gated (src/shader_builder_v2.rs:2525-2806, normal phase
 * `result += w * V(temp)`) -> post-affine if enabled
 * (shaders/core/main_template.wgsl:355-384). */
void orc_apply_xform(const ffm_flame_desc* f, const ffm_xform* xf, float* px, float* py,
                     orc_rng* rng) {
  float x = *px, y = *py;
  float ax = xf->a * x + xf->b * y + xf->e;
  float ay = xf->c * x + xf->d * y + xf->f;
  float rx = 0.0f, ry = 0.0f;
  track(x);
  track(y);
  track(xf->a * x);
  track(xf->b * y);
  track(xf->c * x);
  track(xf->d * y);
  track(ax);
  track(ay);
  /* phase 1, pre variations in local-index order, each replacing the point
   * the normal phase reads (src/shader_builder_v2.rs:2596-2650): pre_blur
   * (src/variations/defs/blur.rs PRE_BLUR) adds a Gaussian-ish offset of
   * its own weight times (six draws summed - 3) at angle 2 pi * draw */
  for (int k = 0; k < f->n_vars; k++) {
    float w = xf->var_weight[k];
    if (f->var_id[k] != FFM_VAR_PRE_BLUR || w == 0.0f) continue;
    float g1 = orc_rng_nextf(rng), g2 = orc_rng_nextf(rng), g3 = orc_rng_nextf(rng);
    float g4 = orc_rng_nextf(rng), g5 = orc_rng_nextf(rng), g6 = orc_rng_nextf(rng);
    float rnd_g = w * (g1 + g2 + g3 + g4 + g5 + g6 - 3.0f);
    float rnd_a = orc_rng_nextf(rng) * F32(6.28318530717959);
    ax = ax + rnd_g * v_cos(rnd_a);
    ay = ay + rnd_g * v_sin(rnd_a);
    track(ax);
    track(ay);
  }
  /* phase 3, normal variations */
  for (int k = 0; k < f->n_vars; k++) {
    float w = xf->var_weight[k];
    if (f->var_id[k] == FFM_VAR_PRE_BLUR) continue;
    if (w != 0.0f) {
      float vx, vy;
      orc_variation(f->var_id[k], xf->var_param[k], xf, ax, ay, rng, &vx, &vy);
      rx += w * vx;
      ry += w * vy;
      track(w * vx);
      track(w * vy);
      track(rx);
      track(ry);
    }
  }
  if (xf->post_enabled) {
    float qx = xf->post_a * rx + xf->post_b * ry + xf->post_e;
    float qy = xf->post_c * rx + xf->post_d * ry + xf->post_f;
    track(xf->post_a * rx);
    track(xf->post_b * ry);
    track(xf->post_c * rx);
    track(xf->post_d * ry);
    track(xf->post_e);
    track(xf->post_f);
    rx = qx;
    ry = qy;
  }
  track(rx);
  track(ry);
  *px = rx;
  *py = ry;
}

/* -------------------------------------------------------------- palette */

/* src/scene/palette.rs:310-360 sample_color. */
void orc_palette_sample(const float* s, int n, float t, float out[3]) {
  if (n <= 0) {
    out[0] = out[1] = out[2] = 1.0f;
    return;
  }
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  if (t < s[0]) {
    out[0] = s[1];
    out[1] = s[2];
    out[2] = s[3];
    return;
  }
  int idx = 0;
  for (int i = 0; i < n; i++) {
    if (s[4 * i] > t) break;
    idx = i;
  }
  if (idx == n - 1) {
    out[0] = s[4 * idx + 1];
    out[1] = s[4 * idx + 2];
    out[2] = s[4 * idx + 3];
    return;
  }
  const float* a = s + 4 * idx;
  const float* b = s + 4 * (idx + 1);
  float range = b[0] - a[0];
  if (range < 1e-6f) {
    out[0] = a[1];
    out[1] = a[2];
    out[2] = a[3];
    return;
  }
  float lt = (t - a[0]) / range;
  for (int c = 0; c < 3; c++) out[c] = a[1 + c] + (b[1 + c] - a[1 + c]) * lt;
}

/* IEEE binary16, round to nearest even (half::f16::from_f32). */
uint16_t orc_f32_to_f16(float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  uint32_t sign = (u >> 16) & 0x8000u;
  uint32_t exp = (u >> 23) & 0xFFu;
  uint32_t man = u & 0x7FFFFFu;
  if (exp == 0xFFu) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u : 0u));
  int e = (int)exp - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7C00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    man |= 0x800000u;
    int shift = 14 - e;
    uint32_t half = man >> shift;
    uint32_t rem = man & ((1u << shift) - 1u);
    uint32_t mid = 1u << (shift - 1);
    if (rem > mid || (rem == mid && (half & 1u))) half++;
    return (uint16_t)(sign | half);
  }
  uint32_t half = ((uint32_t)e << 10) | (man >> 13);
  uint32_t rem = man & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (half & 1u))) half++;
  return (uint16_t)(sign | half);
}

float orc_f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t exp = (h >> 10) & 0x1Fu;
  uint32_t man = h & 0x3FFu;
  uint32_t u;
  if (exp == 0) {
    if (man == 0) {
      u = sign;
    } else {
      int e = -1;
      do {
        e++;
        man <<= 1;
      } while (!(man & 0x400u));
      u = sign | ((uint32_t)(127 - 15 - e) << 23) | ((man & 0x3FFu) << 13);
    }
  } else if (exp == 31) {
    u = sign | 0x7F800000u | (man << 13);
  } else {
    u = sign | ((exp + 127 - 15) << 23) | (man << 13);
  }
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* Texel table: t = i/(size-1) (src/scene/palette.rs:265-278), clamp to
 * [0,1] and f16 (src/gpu/buffers.rs:3215-3221), then srgb_to_linear
 * pow(max(c,0), 2.2) (shaders/core/utilities.wgsl:143-146). */
void orc_palette_linear(const float* stops, int n, int size, float* lin) {
  for (int i = 0; i < size; i++) {
    float t = (float)i / (float)(size - 1);
    float c[3];
    orc_palette_sample(stops, n, t, c);
    for (int k = 0; k < 3; k++) {
      float v = c[k] < 0.0f ? 0.0f : (c[k] > 1.0f ? 1.0f : c[k]);
      float h = orc_f16_to_f32(orc_f32_to_f16(v));
      lin[3 * i + k] = powf(fmaxf(h, 0.0f), F32(2.2));
    }
  }
}

/* transform_lookup_table (src/scene/palette.rs:703-800), stage by stage on
 * a table of texel numbers: squeeze (linear repeats or geometric octaves,
 * palette.rs:607-640), exponential redistribution (palette.rs:653-671),
 * cyclic rotation, reverse. src[i] is the sampled texel entry i shows. */
static size_t orc_as_index(float v, int size) { /* Rust `as usize`, clamped */
  if (!(v > 0.0f)) return 0;
  size_t k = v >= 4294967295.0f ? (size_t)size : (size_t)v;
  return k < (size_t)size - 1 ? k : (size_t)size - 1;
}

static float orc_geometric(float t, float r) {
  if (r >= 0.999f) return t;
  if (r <= 0.001f) {
    float q = t / fmaxf(r, 1e-6f);
    return q < 0.0f ? 0.0f : (q > 1.0f ? 1.0f : q);
  }
  if (t <= 0.0f) return 0.0f;
  float start = 0.0f, rn = 1.0f;
  for (int n = 0; n < 64; n++) {
    float w = rn * (1.0f - r);
    float end = start + w;
    if (t < end) {
      if (w <= 0.0f) return 0.0f;
      float q = (t - start) / w;
      return q < 0.0f ? 0.0f : (q > 1.0f ? 1.0f : q);
    }
    start = end;
    rn *= r;
    if (rn < 1e-10f) return 1.0f;
  }
  return 1.0f;
}

static float orc_redistribute(float t, float s) {
  if (fabsf(s) < 1.1920929e-7f) return t;
  float q;
  if (s > 0.0f) {
    float d = expm1f(s);
    if (d == 0.0f) return t;
    q = expm1f(s * t) / d;
  } else {
    float d = expm1f(-s);
    if (d == 0.0f) return t;
    q = 1.0f - expm1f(-s * (1.0f - t)) / d;
  }
  return q < 0.0f ? 0.0f : (q > 1.0f ? 1.0f : q);
}

void orc_palette_index_map(int size, const ffm_palette_xf* xf, int32_t* src) {
  int32_t* a = (int32_t*)malloc(sizeof(int32_t) * (size_t)size);
  int32_t* b = (int32_t*)malloc(sizeof(int32_t) * (size_t)size);
  for (int i = 0; i < size; i++) a[i] = i;
  if (xf) {
    /* stage 1: squeeze */
    if (xf->squeeze_mode == 0 && xf->squeeze_factor != 1.0f) {
      for (int i = 0; i < size; i++) {
        float t = (float)i / (float)size;
        float x = t * xf->squeeze_factor;
        float st = x - truncf(x);
        b[i] = a[orc_as_index(st * (float)size, size)];
      }
      memcpy(a, b, sizeof(int32_t) * (size_t)size);
    } else if (xf->squeeze_mode == 1) {
      for (int i = 0; i < size; i++) {
        float st = orc_geometric((float)i / (float)size, xf->squeeze_factor);
        b[i] = a[orc_as_index(st * (float)size, size)];
      }
      memcpy(a, b, sizeof(int32_t) * (size_t)size);
    }
    /* stage 2: logarithmic redistribution */
    if (!(fabsf(xf->log_strength) < 1.1920929e-7f)) {
      for (int i = 0; i < size; i++) {
        float st = orc_redistribute((float)i / (float)size, xf->log_strength);
        b[i] = a[orc_as_index(st * (float)size, size)];
      }
      memcpy(a, b, sizeof(int32_t) * (size_t)size);
    }
    /* stage 3: rotation by round(rotation * size) entries (rem_euclid) */
    if (xf->rotation != 0.0f) {
      int amount = (int)roundf(xf->rotation * (float)size);
      for (int i = 0; i < size; i++) {
        int k = (i + amount) % size;
        b[i] = a[k < 0 ? k + size : k];
      }
      memcpy(a, b, sizeof(int32_t) * (size_t)size);
    }
    /* stage 4: reverse */
    if (xf->reverse) {
      for (int i = 0; i < size; i++) b[i] = a[size - 1 - i];
      memcpy(a, b, sizeof(int32_t) * (size_t)size);
    }
  }
  memcpy(src, a, sizeof(int32_t) * (size_t)size);
  free(a);
  free(b);
}

void orc_palette_linear_xf(const float* stops, int n, int size, const ffm_palette_xf* xf,
                           float* lin) {
  float* base = (float*)malloc(sizeof(float) * 3 * (size_t)size);
  int32_t* src = (int32_t*)malloc(sizeof(int32_t) * (size_t)size);
  orc_palette_linear(stops, n, size, base);
  orc_palette_index_map(size, xf, src);
  for (int i = 0; i < size; i++)
    for (int k = 0; k < 3; k++) lin[3 * i + k] = base[3 * src[i] + k];
  free(base);
  free(src);
}

/* Nearest texel of a clamp-to-edge sampler at u = colour index. */
int orc_palette_texel(float ci, int size) {
  float u = ci * (float)size;
  if (!(u >= 0.0f)) return 0;
  if (u >= (float)size) return size - 1;
  int i = (int)floorf(u);
  return i > size - 1 ? size - 1 : i;
}

/* ----------------------------------------------------------------- view */

/* shaders/core/utilities.wgsl:417-445 (2D, not cylinder-relative):
 * pan, rotate, zoom, scale by min(W,H)/4 about the centre, floor. */
void orc_world_to_pixel(const ffm_view* v, float x, float y, int32_t* px, int32_t* py) {
  float tx = x - v->pan_x, ty = y - v->pan_y;
  float cr = cosf(v->rotation), sr = sinf(v->rotation);
  float rx = tx * cr - ty * sr;
  float ry = tx * sr + ty * cr;
  rx = rx * v->zoom;
  ry = ry * v->zoom;
  float scale = (float)(v->width < v->height ? v->width : v->height) * 0.25f;
  float cx = (float)v->width * 0.5f, cy = (float)v->height * 0.5f;
  float pxf = cx + rx * scale;
  float pyf = cy + ry * scale;
  float fx = floorf(pxf), fy = floorf(pyf);
  /* i32(f32) saturates on GPUs (WGSL); far points are off-frame either way */
  *px = fx >= 2147483520.0f ? INT32_MAX : (fx <= -2147483648.0f ? INT32_MIN : (fx != fx ? 0 : (int32_t)fx));
  *py = fy >= 2147483520.0f ? INT32_MAX : (fy <= -2147483648.0f ? INT32_MIN : (fy != fy ? 0 : (int32_t)fy));
}

/* ----------------------------------------------------------- chaos game */

typedef struct {
  const ffm_flame_desc* f;
  const float* pal;
  int pal_size;
  const ffm_view* v;
  uint32_t burn_in;
} step_ctx;

enum { ST_PLOT = 1, ST_RESPAWN = 2, ST_BURN = 4 };

/* One iteration of the loop body, shaders/core/main_template.wgsl
 * :193-1383 in its 2D / palette colour / direct-histogram form
 * (tests/shader_dumps/in-app-2d.wgsl:1163-1382). Returns ST_* flags;
 * *pix gets the pixel index when plotted in frame and *rgb the splat. */
static int step(const step_ctx* c, orc_walker* w, orc_rng* rng, int32_t* ppx, int32_t* ppy,
                int* xf_out, uint32_t rgb[3]) {
  const ffm_flame_desc* f = c->f;
  float rand_val = orc_rng_nextf(rng);
  int xi;
  if (f->has_xaos) {
    xi = orc_select_xaos(f, rand_val, (int)w->prev);
    w->prev = (uint32_t)xi;
  } else {
    xi = orc_select(f, rand_val);
  }
  *xf_out = xi;
  const ffm_xform* xf = &f->xforms[xi];
  int should_plot = orc_rng_nextf(rng) < xf->opacity;
  float x = w->x, y = w->y;
  orc_apply_xform(f, xf, &x, &y, rng);
  /* bad-value respawn (main_template.wgsl:397-479) */
  if (!(fabsf(x) <= 1e32f) || !(fabsf(y) <= 1e32f)) {
    w->x = orc_rng_nextf(rng) * 2.0f - 1.0f;
    w->y = orc_rng_nextf(rng) * 2.0f - 1.0f;
    w->fuse = c->burn_in;
    *ppx = *ppy = -1;
    return ST_RESPAWN;
  }
  w->x = x;
  w->y = y;
  /* colour-index blend (main_template.wgsl:496-503) */
  float sym = xf->color_speed;
  float c1 = (1.0f + sym) / 2.0f;
  float c2 = xf->color * (1.0f - sym) / 2.0f;
  w->color_index = w->color_index * c1 + c2;
  if (w->fuse > 0u) {
    w->fuse -= 1u;
    *ppx = *ppy = -1;
    return ST_BURN;
  }
  /* final chain (main_template.wgsl:663-717): reshapes the plot only; the
   * finals this transform attaches (its AttachmentList,
   * shaders/core/header.wgsl:173-178), or every final in order */
  g_track = 0; /* the final chain reshapes the plot, not the walker */
  float fx = x, fy = y;
  int nc = f->per_transform_finals ? xf->n_final_attach : f->n_finals;
  for (int j = 0; j < nc; j++) {
    int i = f->per_transform_finals ? xf->final_attach[j] : j;
    orc_apply_xform(f, &f->finals[i], &fx, &fy, rng);
  }
  int32_t px, py;
  orc_world_to_pixel(c->v, fx, fy, &px, &py);
  *ppx = px;
  *ppy = py;
  if (px >= 0 && px < c->v->width && py >= 0 && py < c->v->height && should_plot) {
    int t = orc_palette_texel(w->color_index, c->pal_size);
    for (int k = 0; k < 3; k++) {
      float v = c->pal[3 * t + k];
      v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
      rgb[k] = (uint32_t)(v * 100.0f); /* main_template.wgsl:1273-1320 */
    }
    return ST_PLOT;
  }
  return 0;
}

static void walker_start(const ffm_flame_desc* f, orc_walker* w, orc_rng* rng,
                         uint32_t burn_in) {
  /* main_template.wgsl:25-56: two draws for the start point, always */
  float x = orc_rng_nextf(rng) * 2.0f - 1.0f;
  float y = orc_rng_nextf(rng) * 2.0f - 1.0f;
  w->x = x;
  w->y = y;
  w->color_index = 0.0f;
  w->fuse = burn_in;
  /* main_template.wgsl:63-71: with xaos, an unconditioned pick seeds the
   * previous transform (one more draw) */
  w->prev = f->has_xaos ? (uint32_t)orc_select(f, orc_rng_nextf(rng)) : 0u;
}

/* main_template.wgsl:125-181: a persistent orbit resumes unless the
 * dispatch starts it (fresh, or the refresh period picks it); one that
 * starts with xaos picks its first transform at the Kronecker point
 * (thread * 2654435769 + seed) >> 8, over 2^24. */
static void walker_resume(const ffm_flame_desc* f, orc_walker* w, const orc_walker* stored,
                          uint32_t tid, uint32_t seed, int fresh, uint32_t period) {
  int restart = fresh || (period != 0 && (tid + seed) % period == 0);
  if (!restart) {
    *w = *stored;
  } else if (f->has_xaos) {
    uint32_t u = tid * 2654435769u + seed;
    w->prev = (uint32_t)orc_select(f, (float)(u >> 8u) / 16777216.0f);
  }
}

uint64_t orc_iterate(const ffm_flame_desc* f, const float* pal, int pal_size, const ffm_view* v,
                     uint32_t walker0, uint32_t n, uint32_t ipt, uint32_t burn_in,
                     uint32_t seed, orc_walker* orbits, int fresh, uint32_t period,
                     uint32_t* hist, uint64_t* inframe) {
  step_ctx c = {f, pal, pal_size, v, burn_in};
  uint64_t plotted = 0, in = 0;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t tid = walker0 + i;
    orc_rng rng;
    orc_rng_init(&rng, tid, seed);
    orc_walker w;
    walker_start(f, &w, &rng, burn_in);
    if (orbits) walker_resume(f, &w, &orbits[i], tid, seed, fresh, period);
    for (uint32_t it = 0; it < ipt; it++) {
      int32_t px, py;
      int xi;
      uint32_t rgb[3];
      int st = step(&c, &w, &rng, &px, &py, &xi, rgb);
      if (!(st & (ST_RESPAWN | ST_BURN))) plotted++;
      if (st & ST_PLOT) {
        in++;
        if (hist) {
          uint32_t* h = hist + 4 * ((size_t)py * (size_t)v->width + (size_t)px);
          h[0] += rgb[0];
          h[1] += rgb[1];
          h[2] += rgb[2];
          h[3] += 100u;
        }
      }
    }
    if (orbits) orbits[i] = w;
  }
  if (inframe) *inframe = in;
  return plotted;
}

typedef struct {
  const ffm_flame_desc* f;
  const float* pal;
  int pal_size;
  const ffm_view* v;
  uint32_t w0, n, ipt, burn_in, seed;
  uint32_t* hist;
  orc_walker* orbits;
  int fresh;
  uint32_t period;
  uint64_t plotted;
} mt_job;

static void* mt_run(void* p) {
  mt_job* j = (mt_job*)p;
  j->plotted = orc_iterate(j->f, j->pal, j->pal_size, j->v, j->w0, j->n, j->ipt, j->burn_in,
                           j->seed, j->orbits, j->fresh, j->period, j->hist, NULL);
  return NULL;
}

static uint64_t iterate_mt(const ffm_flame_desc* f, const float* pal, int pal_size,
                           const ffm_view* v, uint32_t n, uint32_t ipt, uint32_t burn_in,
                           uint32_t seed, orc_walker* orbits, int fresh, uint32_t period,
                           uint32_t* hist, int threads) {
  if (threads < 1) threads = 1;
  size_t words = (size_t)v->width * (size_t)v->height * 4;
  mt_job* jobs = (mt_job*)calloc((size_t)threads, sizeof(mt_job));
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  uint32_t per = n / (uint32_t)threads, rem = n % (uint32_t)threads, w0 = 0;
  for (int t = 0; t < threads; t++) {
    uint32_t cnt = per + ((uint32_t)t < rem ? 1u : 0u);
    jobs[t] = (mt_job){f, pal, pal_size, v, w0, cnt, ipt, burn_in, seed, NULL,
                       orbits ? orbits + w0 : NULL, fresh, period, 0};
    jobs[t].hist = t == 0 ? hist : (uint32_t*)calloc(words, 4);
    w0 += cnt;
    pthread_create(&th[t], NULL, mt_run, &jobs[t]);
  }
  uint64_t plotted = 0;
  for (int t = 0; t < threads; t++) {
    pthread_join(th[t], NULL);
    plotted += jobs[t].plotted;
  }
  for (int t = 1; t < threads; t++) {
    for (size_t i = 0; i < words; i++) hist[i] += jobs[t].hist[i];
    free(jobs[t].hist);
  }
  free(jobs);
  free(th);
  return plotted;
}

uint64_t orc_iterate_mt(const ffm_flame_desc* f, const float* pal, int pal_size,
                        const ffm_view* v, uint32_t n, uint32_t ipt, uint32_t burn_in,
                        uint32_t seed, uint32_t* hist, int threads) {
  return iterate_mt(f, pal, pal_size, v, n, ipt, burn_in, seed, NULL, 1, 0, hist, threads);
}

void orc_trace(const ffm_flame_desc* f, const float* pal, int pal_size, const ffm_view* v,
               uint32_t walker, uint32_t seed, uint32_t burn_in, uint32_t steps, uint32_t* out) {
  step_ctx c = {f, pal, pal_size, v, burn_in};
  orc_rng rng;
  orc_rng_init(&rng, walker, seed);
  orc_walker w;
  walker_start(f, &w, &rng, burn_in);
  for (uint32_t s = 0; s < steps; s++) {
    int32_t px, py;
    int xi;
    uint32_t rgb[3];
    int st = step(&c, &w, &rng, &px, &py, &xi, rgb);
    uint32_t* o = out + 8 * (size_t)s;
    o[0] = rng.state;
    o[1] = (uint32_t)xi;
    memcpy(&o[2], &w.x, 4);
    memcpy(&o[3], &w.y, 4);
    memcpy(&o[4], &w.color_index, 4);
    o[5] = (uint32_t)px;
    o[6] = (uint32_t)py;
    o[7] = (uint32_t)st;
  }
}

/* One step from a given RNG and walker state (per-step parity): writes the
 * same 8-word record as orc_trace and advances *rng and *w. */
int orc_step_from(const ffm_flame_desc* f, const float* pal, int pal_size, const ffm_view* v,
                  uint32_t burn_in, orc_rng* rng, orc_walker* w, uint32_t* o) {
  step_ctx c = {f, pal, pal_size, v, burn_in};
  int32_t px, py;
  int xi;
  uint32_t rgb[3];
  g_track = 1;
  g_term = 0.0f;
  int st = step(&c, w, rng, &px, &py, &xi, rgb);
  g_track = 0;
  o[0] = rng->state;
  o[1] = (uint32_t)xi;
  memcpy(&o[2], &w->x, 4);
  memcpy(&o[3], &w->y, 4);
  memcpy(&o[4], &w->color_index, 4);
  o[5] = (uint32_t)px;
  o[6] = (uint32_t)py;
  o[7] = (uint32_t)st;
  return st;
}

/* ---------------------------------------------------------- OrbitFuses */

/* src/renderer/compute_kernel.rs:73-102 */
uint64_t orc_orbit_fuses_dispatch(orc_orbit_fuses* c, uint32_t threads, uint32_t ipt,
                                  uint32_t burn_in, uint32_t period, uint32_t seed) {
  if (c->len < threads) {
    c->left = (uint32_t*)realloc(c->left, (size_t)threads * 4);
    for (uint32_t i = c->len; i < threads; i++) c->left[i] = burn_in;
    c->len = threads;
  }
  uint64_t plotted = 0;
  for (uint32_t tid = 0; tid < threads; tid++) {
    if (period != 0 && (tid + seed) % period == 0) c->left[tid] = burn_in;
    uint32_t burnt = c->left[tid] < ipt ? c->left[tid] : ipt;
    plotted += (uint64_t)(ipt - burnt);
    c->left[tid] -= burnt;
  }
  return plotted;
}

void orc_orbit_fuses_free(orc_orbit_fuses* c) {
  free(c->left);
  c->left = NULL;
  c->len = 0;
}

/* ------------------------------------------------------------- stencils */

/* shaders/histogram_blur.wgsl:39-119, one pass. `center_src` is the buffer
 * the centre density comes from (the original primary in both passes). */
void orc_blur_pass(const uint32_t* in, uint32_t* out, const uint32_t* center_src, int W, int H,
                   float radius_pixels, float density_sigma, int direction) {
  const int KMAX = 8;
  float sigma = fmaxf(radius_pixels, 0.0001f);
  int kh = (int)ceilf(3.0f * sigma);
  if (kh > KMAX) kh = KMAX;
  float inv_2s2 = 1.0f / (2.0f * sigma * sigma);
  float sigma_d = fmaxf(density_sigma, 1.0f);
  float inv_2sd2 = 1.0f / (2.0f * sigma_d * sigma_d);
  for (int y = 0; y < H; y++) {
    for (int x = 0; x < W; x++) {
      size_t ci = ((size_t)y * W + x) * 4;
      float cd = (float)center_src[ci + 3];
      float sr = 0.0f, sg = 0.0f, sb = 0.0f, sd = 0.0f, wt = 0.0f;
      for (int k = -KMAX; k <= KMAX; k++) {
        if (abs(k) > kh) continue;
        int nx = x, ny = y;
        if (direction == 0) {
          nx = x + k < 0 ? 0 : (x + k > W - 1 ? W - 1 : x + k);
        } else {
          ny = y + k < 0 ? 0 : (y + k > H - 1 ? H - 1 : y + k);
        }
        size_t ni = ((size_t)ny * W + nx) * 4;
        float tr = (float)in[ni], tg = (float)in[ni + 1], tb = (float)in[ni + 2],
              td = (float)in[ni + 3];
        float fk = (float)k;
        float w_sp = expf(-fk * fk * inv_2s2);
        float delta = td - cd;
        float w_d = expf(-delta * delta * inv_2sd2);
        float w = w_sp * w_d;
        sr = sr + tr * w;
        sg = sg + tg * w;
        sb = sb + tb * w;
        sd = sd + td * w;
        wt = wt + w;
      }
      float inv_w = 1.0f / fmaxf(wt, 1e-6f);
      out[ci] = (uint32_t)(sr * inv_w);
      out[ci + 1] = (uint32_t)(sg * inv_w);
      out[ci + 2] = (uint32_t)(sb * inv_w);
      out[ci + 3] = (uint32_t)(sd * inv_w);
    }
  }
}

/* H pass primary -> scratch, V pass scratch -> primary with the centre
 * density read from the primary before it is written
 * (histogram_blur.wgsl:57-70): the V pass writes a copy, then swaps. */
void orc_blur(uint32_t* hist, uint32_t* scratch, int W, int H, float sigma, float density_sigma) {
  size_t words = (size_t)W * H * 4;
  orc_blur_pass(hist, scratch, hist, W, H, sigma, density_sigma, 0);
  uint32_t* tmp = (uint32_t*)malloc(words * 4);
  orc_blur_pass(scratch, tmp, hist, W, H, sigma, density_sigma, 1);
  memcpy(hist, tmp, words * 4);
  free(tmp);
}

/* src/renderer/compute_kernel.rs:1688-1691:
 * sigma_d = (samples/pixels)*100 * 100^(2*edges-1). */
float orc_blur_density_sigma(double samples, int W, int H, float blur_edges) {
  float pixel_count = (float)W * (float)H;
  float mean_scaled = ((float)samples / pixel_count) * 100.0f;
  return mean_scaled * powf(100.0f, 2.0f * blur_edges - 1.0f);
}

/* shaders/accumulate.wgsl:115-230 (no depth). mode 0 cumulative, 1 EMA,
 * 2 overwrite; count != NULL is the exact 64-bit count. */
void orc_accumulate(const void* hist_v, int hist_is_f32, const float* prev_buf, float* out,
                    uint32_t* count, int W, int H, int mode, float blend) {
  const float SCALE = 100.0f;
  for (size_t p = 0; p < (size_t)W * H; p++) {
    float prev[4];
    for (int k = 0; k < 4; k++) prev[k] = mode == 2 ? 0.0f : prev_buf[4 * p + k];
    float h[4];
    uint32_t hd = 0;
    if (hist_is_f32) {
      const float* hf = (const float*)hist_v;
      for (int k = 0; k < 4; k++) h[k] = hf[4 * p + k];
      hd = (uint32_t)h[3];
    } else {
      const uint32_t* hu = (const uint32_t*)hist_v;
      for (int k = 0; k < 4; k++) h[k] = (float)hu[4 * p + k];
      hd = hu[4 * p + 3];
    }
    float new_density = h[3] / SCALE;
    float total = prev[3] + new_density;
    if (count) {
      uint32_t lo = count[2 * p], hi = count[2 * p + 1];
      if (prev[3] == 0.0f) {
        lo = hi = 0;
      } else if (lo == 0 && hi == 0) {
        float seed = prev[3] * SCALE;
        float fhi = floorf(seed / 4294967296.0f);
        lo = (uint32_t)(seed - fhi * 4294967296.0f);
        hi = (uint32_t)fhi;
      }
      uint32_t nlo = lo + hd;
      hi = hi + (nlo < lo ? 1u : 0u);
      lo = nlo;
      count[2 * p] = lo;
      count[2 * p + 1] = hi;
      total = ((float)hi * 4294967296.0f + (float)lo) / SCALE;
      prev[3] = total;
    }
    if (new_density <= 0.0f) {
      for (int k = 0; k < 4; k++) out[4 * p + k] = prev[k];
      continue;
    }
    float dn = fmaxf(h[3], 1.0f);
    float nc[3];
    for (int k = 0; k < 3; k++) {
      float v = h[k] / dn;
      nc[k] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    }
    float cum_alpha = new_density / total;
    float alpha = mode == 1 ? blend : cum_alpha;
    for (int k = 0; k < 3; k++) out[4 * p + k] = prev[k] + (nc[k] - prev[k]) * alpha;
    out[4 * p + 3] = total;
  }
}

/* ------------------------------------------------- density estimation */

/* flam3's estimator as Draves & Reckase describe it ("The Fractal Flame
 * Algorithm", 2003/2008, section on density estimation) and flam3 builds
 * it: a table of Gaussian discs, filter i of half-width
 * h_i = (max_r+1) / (i+1)^curve (indices past 100 spaced by
 * (i-100)^(1/curve)), floored at min_r+1; a pixel of count c uses
 * filter ceil(c)-1 up to 100, then 100 + floor((c-100)^curve). Tap (dx,dy)
 * of filter h weighs exp(-2 (1.5 d/h)^2) for d = |(dx,dy)| <= h, normalised
 * over the disc; each pixel scatters its colour sums and count through its
 * filter. The reference has no estimator, so nothing here is pinned by it. */
typedef struct {
  int n;
  double* h;
  double* inv_norm;
} de_table;

static void de_build(de_table* t, float max_r, float min_r, float curve) {
  double comp_max = (double)max_r + 1.0, comp_min = (double)min_r + 1.0;
  int cap = 4096;
  t->h = (double*)malloc(sizeof(double) * cap);
  t->inv_norm = (double*)malloc(sizeof(double) * cap);
  t->n = 0;
  const int keep = 100;
  for (int i = 0; i < cap; i++) {
    double adj = i < keep ? (double)i : pow((double)(i - keep), 1.0 / curve) + keep;
    double h = comp_max / pow(adj + 1.0, curve);
    int last = 0;
    if (h <= comp_min) {
      h = comp_min;
      last = 1;
    }
    int R = (int)ceil(h);
    double sum = 0.0;
    for (int dy = -R; dy <= R; dy++)
      for (int dx = -R; dx <= R; dx++) {
        double d = sqrt((double)(dx * dx + dy * dy)) / h;
        if (d <= 1.0) sum += exp(-2.0 * (1.5 * d) * (1.5 * d));
      }
    t->h[t->n] = h;
    t->inv_norm[t->n] = 1.0 / sum;
    t->n++;
    if (last) break;
  }
}

static int de_index(const de_table* t, double c, float curve) {
  int idx;
  if (c <= 100.0)
    idx = (int)ceil(c) - 1;
  else
    idx = 100 + (int)floor(pow(c - 100.0, curve));
  if (idx < 0) idx = 0;
  if (idx > t->n - 1) idx = t->n - 1;
  return idx;
}

void orc_density_estimate(const float* in, float* out, int W, int H, float max_r, float min_r,
                          float curve) {
  de_table t;
  de_build(&t, max_r, min_r, curve);
  size_t np = (size_t)W * H;
  double* acc = (double*)calloc(np * 4, sizeof(double));
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      size_t p = (size_t)y * W + x;
      double c = in[4 * p + 3];
      if (!(c > 0.0)) continue;
      int idx = de_index(&t, c, curve);
      double h = t.h[idx], inv = t.inv_norm[idx];
      int R = (int)ceil(h);
      double v[4] = {in[4 * p] * c, in[4 * p + 1] * c, in[4 * p + 2] * c, c};
      for (int dy = -R; dy <= R; dy++) {
        int yy = y + dy;
        if (yy < 0 || yy >= H) continue;
        for (int dx = -R; dx <= R; dx++) {
          int xx = x + dx;
          if (xx < 0 || xx >= W) continue;
          double d = sqrt((double)(dx * dx + dy * dy)) / h;
          if (d > 1.0) continue;
          double wgt = exp(-2.0 * (1.5 * d) * (1.5 * d)) * inv;
          size_t q = (size_t)yy * W + xx;
          for (int k = 0; k < 4; k++) acc[4 * q + k] += v[k] * wgt;
        }
      }
    }
  for (size_t p = 0; p < np; p++) {
    double c = acc[4 * p + 3];
    if (c > 0.0) {
      for (int k = 0; k < 3; k++) out[4 * p + k] = (float)(acc[4 * p + k] / c);
    } else {
      for (int k = 0; k < 3; k++) out[4 * p + k] = 0.0f;
    }
    out[4 * p + 3] = (float)c;
  }
  free(acc);
  free(t.h);
  free(t.inv_norm);
}

/* -------------------------------------------------------------- tonemap */

void orc_tonemap_defaults(ffm_tonemap_params* tp) {
  /* src/config/defaults.rs:19-47,100-101 */
  memset(tp, 0, sizeof(*tp));
  tp->tonemap_mode = 1;
  tp->highlight_mode = 0;
  tp->exposure = 1.0f;
  tp->gamma = 4.0f;
  tp->gamma_threshold = 5.0f;
  tp->brightness = 4.0f;
  tp->vibrancy = 1.0f;
  tp->white_level = 200.0f;
  tp->density_scale = 1.0f;
  tp->saturation = 1.0f;
  tp->hue_shift = 0.0f;
  tp->alpha_blend_low = 0.3f;
  tp->alpha_blend_high = 0.8f;
  tp->levels_enabled = 1;
  tp->use_curve = 1; /* fractal_config.rs:475-476, the linear ToneCurve */
  tp->levels_low = 0.0f;
  tp->levels_high = 10.0f;
  tp->levels_gamma = 0.5f;
}

static float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* shaders/tonemap.wgsl:153-211 */
static void rgb_to_hsv(const float c[3], float o[3]) {
  float r = c[0], g = c[1], b = c[2];
  float mx = fmaxf(fmaxf(r, g), b), mn = fminf(fminf(r, g), b);
  float delta = mx - mn;
  float h = 0.0f, s = 0.0f;
  if (delta > 0.00001f) {
    s = delta / mx;
    if (r >= mx)
      h = (g - b) / delta;
    else if (g >= mx)
      h = 2.0f + (b - r) / delta;
    else
      h = 4.0f + (r - g) / delta;
    h = h * 60.0f;
    if (h < 0.0f) h = h + 360.0f;
  }
  o[0] = h;
  o[1] = s;
  o[2] = mx;
}

static void hsv_to_rgb(const float hsv[3], float o[3]) {
  float h = hsv[0], s = hsv[1], v = hsv[2];
  if (s <= 0.0f) {
    o[0] = o[1] = o[2] = v;
    return;
  }
  float hh = h;
  if (hh >= 360.0f) hh = 0.0f;
  hh = hh / 60.0f;
  uint32_t i = (uint32_t)hh;
  float ff = hh - (float)i;
  float p = v * (1.0f - s);
  float q = v * (1.0f - (s * ff));
  float t = v * (1.0f - (s * (1.0f - ff)));
  switch (i) {
    case 0: o[0] = v; o[1] = t; o[2] = p; break;
    case 1: o[0] = q; o[1] = v; o[2] = p; break;
    case 2: o[0] = p; o[1] = v; o[2] = t; break;
    case 3: o[0] = p; o[1] = q; o[2] = v; break;
    case 4: o[0] = t; o[1] = p; o[2] = v; break;
    default: o[0] = v; o[1] = p; o[2] = q; break;
  }
}

static uint8_t unorm8(float v) {
  v = v != v ? 0.0f : clampf(v, 0.0f, 1.0f);
  return (uint8_t)floorf(v * 255.0f + 0.5f);
}

/* shaders/tonemap.wgsl:224-590 fs_main, with TonemapParams as
 * update_tonemap fills them (src/renderer/compute_kernel.rs:3813-3918). */
/* f16::from_f32 (round to nearest, ties to even; overflow to inf, 2^-24
 * subnormals) taken back to f32, by the bit fields of the f32 input. */
float orc_round_f16(float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  uint32_t sign = u & 0x80000000u, a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return v;                   /* NaN */
  int e = (int)(a >> 23) - 127;                    /* unbiased exponent */
  int drop;                                        /* f32 fraction bits f16 lacks */
  if (e < -25) { a = 0; drop = 0; }                /* below half the smallest subnormal */
  else drop = e < -14 ? 13 + (-14 - e) : 13;       /* subnormal f16 keeps fewer bits */
  if (a) {
    uint32_t m = (a & 0x7fffffu) | (e > -127 ? 0x800000u : 0);  /* significand, implicit 1 */
    uint32_t keep = drop < 32 ? m >> drop : 0;
    uint32_t rest = drop < 32 ? m & ((1u << drop) - 1u) : m, half = 1u << (drop - 1);
    if (rest > half || (rest == half && (keep & 1u))) keep++;
    /* back to a value: keep x 2^(e - 23 + drop) */
    double val = ldexp((double)keep, e - 23 + drop);
    if (val >= 65520.0) val = INFINITY;
    float f = (float)val;
    memcpy(&a, &f, 4);
  }
  a |= sign;
  memcpy(&v, &a, 4);
  return v;
}

/* ToneCurve::evaluate (src/scene/tonemap.rs:209-286): monotone cubic
 * Hermite through the points in f32, tangents the mean of the neighbouring
 * secants (the segment's own at the ends) limited by Fritsch-Carlson. */
float orc_curve_eval(const float* px, const float* py, int n, float x) {
  if (x < 0.0f) x = 0.0f;
  if (x > 1.0f) x = 1.0f;
  for (int i = 0; i < n - 1; i++) {
    if (!(x >= px[i] && x <= px[i + 1])) continue;
    float h = px[i + 1] - px[i];
    float delta = (py[i + 1] - py[i]) / h;
    float m0 = delta, m1 = delta;
    if (i > 0) m0 = ((py[i] - py[i - 1]) / (px[i] - px[i - 1]) + delta) * 0.5f;
    if (i < n - 2) m1 = (delta + (py[i + 2] - py[i + 1]) / (px[i + 2] - px[i + 1])) * 0.5f;
    float g0, g1;
    if (fabsf(delta) < 1e-6f) {
      g0 = 0.0f;
      g1 = 0.0f;
    } else {
      float al = m0 / delta, be = m1 / delta;
      if (al < 0.0f || be < 0.0f) {
        g0 = delta;
        g1 = delta;
      } else if (al * al + be * be > 9.0f) {
        float tau = 3.0f / sqrtf(al * al + be * be);
        g0 = tau * al * delta;
        g1 = tau * be * delta;
      } else {
        g0 = m0;
        g1 = m1;
      }
    }
    float t = (x - px[i]) / h, t2 = t * t, t3 = t2 * t;
    float h00 = 2.0f * t3 - 3.0f * t2 + 1.0f, h10 = t3 - 2.0f * t2 + t;
    float h01 = -2.0f * t3 + 3.0f * t2, h11 = t3 - t2;
    return h00 * py[i] + h10 * h * g0 + h01 * py[i + 1] + h11 * h * g1;
  }
  return x;
}

/* ToneCurve::generate_lut (src/scene/tonemap.rs:289-303): texel i holds
 * f16(curve(i/255)); curve_n 0 is the linear curve. */
void orc_curve_lut(const ffm_tonemap_params* tp, float* lut) {
  static const float lin[2] = {0.0f, 1.0f};
  int n = tp->curve_n ? tp->curve_n : 2;
  const float* px = tp->curve_n ? tp->curve_x : lin;
  const float* py = tp->curve_n ? tp->curve_y : lin;
  for (int i = 0; i < 256; i++) lut[i] = orc_round_f16(orc_curve_eval(px, py, n, (float)i / 255.0f));
}

/* A linear-filtering, clamp-to-edge read of the 256-texel table at
 * texel-space position s = 255c: lut[i] (1 - f) + lut[i + 1] f for
 * i = floor(s), f = s - i; s <= 0 or NaN reads texel 0, s >= 255 texel 255.
 * Written the way a texture unit filters (texel centres at (i + 0.5)/256). */
float orc_curve_sample(const float* lut, float c) {
  float s = c * 255.0f;
  if (!(s > 0.0f)) return lut[0];
  if (s >= 255.0f) return lut[255];
  int i = (int)floorf(s);
  float f = s - (float)i;
  return lut[i] * (1.0f - f) + lut[i + 1] * f;
}

/* downsample_2x_firefly (src/export/supersample.rs:117-187): RGBA8 of
 * (2 ow) x (2 oh) boxed 2x2 down to ow x oh. Samples in the order (0,0),
 * (1,0), (0,1), (1,1); luminance proxy r + 2g + b; the median of the four
 * is the mean of the two middle values (sorted); a sample above
 * cap = 4 median + 64 is scaled by cap / lum; alpha is averaged unclamped;
 * (acc / 4) rounds half away from zero, clamped to 0..255. */
void orc_downsample2x(const uint8_t* in, uint8_t* out, int ow, int oh) {
  const int w2 = 2 * ow;
  for (int oy = 0; oy < oh; oy++)
    for (int ox = 0; ox < ow; ox++) {
      static const int dxy[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
      float smp[4][4], lum[4], srt[4];
      for (int i = 0; i < 4; i++) {
        const uint8_t* q = in + 4 * ((size_t)(2 * oy + dxy[i][1]) * w2 + 2 * ox + dxy[i][0]);
        for (int c = 0; c < 4; c++) smp[i][c] = (float)q[c];
        lum[i] = smp[i][0] + 2.0f * smp[i][1] + smp[i][2];
        srt[i] = lum[i];
      }
      for (int i = 1; i < 4; i++) /* insertion sort of four */
        for (int j = i; j > 0 && srt[j - 1] > srt[j]; j--) {
          float t = srt[j];
          srt[j] = srt[j - 1];
          srt[j - 1] = t;
        }
      const float median = 0.5f * (srt[1] + srt[2]);
      const float cap = median * 4.0f + 64.0f;
      float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      for (int i = 0; i < 4; i++) {
        const float scale = lum[i] > cap ? cap / lum[i] : 1.0f;
        acc[0] += smp[i][0] * scale;
        acc[1] += smp[i][1] * scale;
        acc[2] += smp[i][2] * scale;
        acc[3] += smp[i][3];
      }
      uint8_t* o = out + 4 * ((size_t)oy * ow + ox);
      for (int c = 0; c < 4; c++) {
        float v = roundf(acc[c] * 0.25f);
        o[c] = (uint8_t)(v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v));
      }
    }
}

void orc_tonemap(const float* accum, uint8_t* rgba, int W, int H, float zoom,
                 double samples_in_buffer, const ffm_tonemap_params* tp) {
  float lut[256];
  if (tp->use_curve) orc_curve_lut(tp, lut);
  const float BRIGHT_ADJUST = 2.3f, PREFILTER_WHITE = 67108864.0f;
  float apo_zoom = log2f(zoom);
  float ppu = (float)(W < H ? W : H) * 0.25f * powf(2.0f, apo_zoom);
  float area = (float)((uint32_t)W * (uint32_t)H) / (ppu * ppu);
  float total_pixels = fmaxf((float)W * (float)H, 1.0f);
  float sample_density = fmaxf((float)samples_in_buffer / total_pixels, 1e-6f);
  float white = tp->white_level;
  float k1 = 1.0f * BRIGHT_ADJUST * tp->brightness * 268.0f * PREFILTER_WHITE / 256.0f;
  float k2 = 1.0f / (1.0f * area * white * sample_density);
  float gamma = tp->gamma == 0.0f ? tp->gamma : 1.0f / tp->gamma;
  float funcval = tp->gamma_threshold != 0.0f ? powf(tp->gamma_threshold, gamma - 1.0f) : 0.0f;
  float vib = rintf(tp->vibrancy * 256.0f);
  float notvib = 256.0f - vib;
  float bg_lin[3];
  for (int k = 0; k < 3; k++) bg_lin[k] = powf(fmaxf(tp->background[k], 0.0f), F32(2.2));
  const float inv22 = F32(1.0 / 2.2);
  for (size_t p = 0; p < (size_t)W * H; p++) {
    const float* a = accum + 4 * p;
    float bucket_count = a[3] * 100.0f;
    float color[3];
    float alpha = 0.0f;
    if (tp->tonemap_mode == 0) {
      float g = tp->gamma == 0.0f ? tp->gamma : 1.0f / tp->gamma;
      for (int k = 0; k < 3; k++) color[k] = powf(a[k] * tp->exposure, g);
      alpha = clampf(bucket_count * 0.01f * tp->density_scale, 0.0f, 1.0f);
    } else if (tp->tonemap_mode == 2) {
      float nd = clampf(bucket_count * 0.01f * tp->exposure, 0.0f, 1.0f);
      float g = tp->gamma == 0.0f ? tp->gamma : 1.0f / tp->gamma;
      float gd = powf(nd, g);
      color[0] = color[1] = color[2] = gd;
      alpha = nd;
    } else {
      float bs;
      if (bucket_count < 0.001f) {
        bs = 0.0f;
      } else {
        float l10 = logf(1.0f + white * bucket_count * k2) / logf(10.0f);
        bs = (k1 * l10) / (white * bucket_count);
      }
      float ls = bs / PREFILTER_WHITE;
      float fp0 = ls * (a[0] * bucket_count);
      float fp1 = ls * (a[1] * bucket_count);
      float fp2 = ls * (a[2] * bucket_count);
      float fp3 = ls * bucket_count * white;
      if (fp3 > 0.0f) {
        if (fp3 <= tp->gamma_threshold) {
          float frac = fp3 / tp->gamma_threshold;
          alpha = (1.0f - frac) * fp3 * funcval + frac * powf(fp3, gamma);
        } else {
          alpha = powf(fp3, gamma);
        }
      }
      ls = fp3 > 0.0f ? vib * alpha / fp3 : 0.0f;
      if (notvib > 0.0f) {
        color[0] = ls * fp0 + notvib * powf(fp0, gamma);
        color[1] = ls * fp1 + notvib * powf(fp1, gamma);
        color[2] = ls * fp2 + notvib * powf(fp2, gamma);
      } else {
        color[0] = ls * fp0;
        color[1] = ls * fp1;
        color[2] = ls * fp2;
      }
    }
    if (tp->saturation != 1.0f || tp->hue_shift != 0.0f) {
      float hsv[3];
      rgb_to_hsv(color, hsv);
      if (tp->hue_shift != 0.0f) {
        hsv[0] = hsv[0] + tp->hue_shift;
        if (hsv[0] < 0.0f)
          hsv[0] = hsv[0] + 360.0f;
        else if (hsv[0] >= 360.0f)
          hsv[0] = hsv[0] - 360.0f;
      }
      if (tp->saturation != 1.0f) hsv[1] = clampf(hsv[1] * tp->saturation, 0.0f, 1.0f);
      hsv_to_rgb(hsv, color);
    }
    if (tp->tonemap_mode == 1)
      for (int k = 0; k < 3; k++) color[k] = color[k] * tp->exposure;
    for (int k = 0; k < 3; k++) color[k] = fmaxf(color[k], 0.0f);
    if (tp->highlight_mode == 1) {
      float m = fmaxf(color[0], fmaxf(color[1], color[2]));
      if (m > 1.0f)
        for (int k = 0; k < 3; k++) color[k] = color[k] / m;
    } else if (tp->highlight_mode == 2) {
      float lum = color[0] * 0.2126f + color[1] * 0.7152f + color[2] * 0.0722f;
      if (lum > 1e-6f) {
        float lm = lum / (1.0f + lum);
        for (int k = 0; k < 3; k++) color[k] = color[k] * (lm / lum);
      }
      for (int k = 0; k < 3; k++) color[k] = fminf(color[k], 1.0f);
    } else if (tp->highlight_mode == 3) {
      for (int k = 0; k < 3; k++) {
        float c = color[k];
        c = (c * (2.51f * c + 0.03f)) / (c * (2.43f * c + 0.59f) + 0.14f);
        color[k] = clampf(c, 0.0f, 1.0f);
      }
    } else {
      for (int k = 0; k < 3; k++) color[k] = fminf(color[k], 1.0f);
    }
    /* the tone curve (tonemap.wgsl:479-492): the 256-texel table read at
     * u = (255c + 0.5)/256 by a linear, clamp-to-edge sampler
     * (src/gpu/buffers.rs:1877-1886), i.e. at texel-space position 255c */
    if (tp->use_curve && bucket_count > 0.001f) {
      for (int k = 0; k < 3; k++) color[k] = orc_curve_sample(lut, color[k]);
    }
    /* apply_levels (tonemap.wgsl:125-151) on accum.a, mean = sample_density */
    float leveled = 1.0f;
    {
      float low = tp->levels_low, high = tp->levels_high, lg = tp->levels_gamma;
      float mean = sample_density;
      if (mean > 0.0f) {
        float nd = a[3] / mean;
        if (high <= low) {
          leveled = nd > low ? 1.0f : 0.0f;
        } else {
          float nrm = (nd - low) / (high - low);
          float cl = clampf(nrm, 0.0f, 1.0f);
          leveled = (lg != 1.0f && lg > 0.0f) ? powf(cl, lg) : cl;
        }
      }
    }
    float linear_alpha = clampf(bucket_count * 0.01f * tp->density_scale, 0.0f, 1.0f);
    float gamma_alpha = clampf(alpha, 0.0f, 1.0f);
    float t = clampf((gamma_alpha - tp->alpha_blend_low) / (tp->alpha_blend_high - tp->alpha_blend_low),
                     0.0f, 1.0f);
    float blend_t = t * t * (3.0f - 2.0f * t);
    float base_alpha = gamma_alpha * (1.0f - blend_t) + linear_alpha * blend_t;
    float fa = tp->levels_enabled ? fminf(base_alpha, leveled) : base_alpha;
    if (tp->transparent_mode != 0) {
      /* tonemap.wgsl:531-576: over_black = pow(max(colour * alpha, 0), 1/2.2) */
      float ob[3];
      for (int k = 0; k < 3; k++) ob[k] = powf(fmaxf(color[k] * fa, 0.0f), inv22);
      if (tp->transparent_mode == 2) { /* premultiplied */
        for (int k = 0; k < 3; k++) rgba[4 * p + k] = unorm8(ob[k]);
        rgba[4 * p + 3] = unorm8(fa);
      } else { /* straight alpha with coverage = max channel */
        float cov = fmaxf(ob[0], fmaxf(ob[1], ob[2]));
        for (int k = 0; k < 3; k++) rgba[4 * p + k] = unorm8(cov > 0.0f ? ob[k] / cov : 0.0f);
        rgba[4 * p + 3] = unorm8(cov);
      }
      continue;
    }
    for (int k = 0; k < 3; k++) {
      float fin = bg_lin[k] * (1.0f - fa) + color[k] * fa;
      rgba[4 * p + k] = unorm8(powf(fin, inv22));
    }
    rgba[4 * p + 3] = 255;
  }
}

/* ---------------------------------------------------------- the render */

/* render_with's loop (src/renderer/render.rs:624-710): dispatch, count
 * with OrbitFuses (compute_kernel.rs:1402-1431), accumulate every
 * batch_size dispatches and once more for a partial batch, stop at the
 * target or after burn_in/ipt+4 empty dispatches; then tonemap. */
int orc_render(const ffm_flame_desc* f_in, const float* stops, int n_stops, int pal_size,
               const ffm_palette_xf* pal_xf, const ffm_job* job, uint64_t exact_after,
               float* accum_out, uint8_t* rgba_out, uint64_t* total_out, int threads,
               const ffm_effects* fx) {
  ffm_flame_desc* f = (ffm_flame_desc*)malloc(sizeof(ffm_flame_desc));
  memcpy(f, f_in, sizeof(*f));
  orc_flame_prepare(f);
  const ffm_view* v = &job->view;
  int W = v->width, H = v->height;
  size_t np = (size_t)W * H;
  float* pal = (float*)malloc(sizeof(float) * 3 * (size_t)pal_size);
  orc_palette_linear_xf(stops, n_stops, pal_size, pal_xf, pal);
  uint32_t* hist = (uint32_t*)calloc(np * 4, 4);
  uint32_t* scratch = (uint32_t*)calloc(np * 4, 4);
  float* acc = (float*)calloc(np * 4, 4);
  float* acc2 = (float*)calloc(np * 4, 4);
  uint32_t nw = (uint32_t)job->walkers;
  orc_walker* orbits = job->persistent_orbits ? (orc_walker*)calloc(nw, sizeof(orc_walker)) : NULL;
  orc_orbit_fuses fuses = {NULL, 0};
  const uint32_t period = orbits ? orc_xaos_refresh_period(f) : 0u;
  /* the exact-count switch (src/renderer/compute_kernel.rs:1562-1575):
   * 64-bit counts from accumulate pass exact_after on */
  uint32_t* count = NULL;
  uint64_t passes = 0;
  uint64_t total = 0, batch_samples = 0, samples_in_buffer = 0;
  uint32_t batch_count = 0, empty = 0, k = 0;
  uint32_t stall_limit = job->burn_in / (job->ipt ? job->ipt : 1) + 4;
  uint32_t bs = job->batch_size ? job->batch_size : 4;
  for (;;) {
    uint32_t seed = job->seed_base + k * job->seed_stride;
    if (batch_count == 0) memset(hist, 0, np * 16);
    if (threads > 1) {
      uint32_t* h2 = (uint32_t*)calloc(np * 4, 4);
      iterate_mt(f, pal, pal_size, v, nw, job->ipt, job->burn_in, seed, orbits, k == 0, period,
                 h2, threads);
      for (size_t i = 0; i < np * 4; i++) hist[i] += h2[i];
      free(h2);
    } else {
      orc_iterate(f, pal, pal_size, v, 0, nw, job->ipt, job->burn_in, seed, orbits, k == 0,
                  period, hist, NULL);
    }
    uint64_t plotted;
    if (orbits) {
      plotted = orc_orbit_fuses_dispatch(&fuses, nw, job->ipt, job->burn_in, period, seed);
    } else {
      uint32_t per = job->ipt > job->burn_in ? job->ipt - job->burn_in : 0;
      plotted = (uint64_t)nw * per;
    }
    k++;
    total += plotted;
    batch_samples += plotted;
    batch_count++;
    empty = plotted == 0 ? empty + 1 : 0;
    int stalled = empty >= stall_limit;
    int done = total >= job->target_samples || stalled;
    if (batch_count >= bs || done) {
      if (job->filter_radius > 0.0f) {
        float dsig = orc_blur_density_sigma((double)batch_samples, W, H, job->filter_blur_edges);
        orc_blur(hist, scratch, W, H, job->filter_radius, dsig);
      }
      if (++passes >= exact_after && !count) count = (uint32_t*)calloc(np * 2, 4);
      orc_accumulate(hist, 0, acc, acc2, count, W, H, 0, 0.0f);
      float* t = acc;
      acc = acc2;
      acc2 = t;
      samples_in_buffer += batch_samples;
      batch_count = 0;
      batch_samples = 0;
    }
    if (done) break;
  }
  if (job->de_max_radius > 0.0f) {
    orc_density_estimate(acc, acc2, W, H, job->de_max_radius, job->de_min_radius, job->de_curve);
    float* t = acc;
    acc = acc2;
    acc2 = t;
  }
  /* the tail: density effects (their mean density is the tone map's
   * sample density), the tone map, colour effects (src/renderer/tail.rs) */
  const float density_mean =
      fmaxf((float)samples_in_buffer / fmaxf((float)W * (float)H, 1.0f), 1e-6f);
  orc_density_chain(fx, acc, acc2, W, H, density_mean);
  if (accum_out) memcpy(accum_out, acc, np * 16);
  if (rgba_out) {
    orc_tonemap(acc, rgba_out, W, H, v->zoom, (double)samples_in_buffer, &job->tonemap);
    if (fx && fx->n_color > 0) {
      uint8_t* tmp = (uint8_t*)malloc(np * 4);
      orc_color_chain(fx, rgba_out, tmp, W, H);
      free(tmp);
    }
  }
  if (total_out) *total_out = total;
  orc_orbit_fuses_free(&fuses);
  free(orbits);
  free(count);
  free(hist);
  free(scratch);
  free(acc);
  free(acc2);
  free(pal);
  free(f);
  return 0;
}
