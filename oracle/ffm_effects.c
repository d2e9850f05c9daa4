/* The render tail's effects, restated on the CPU.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP effect kernels are held to.
 * Nothing of the product links or calls this file.
 *
 * Follows, as text:
 *   texture read: uf_texture, src/lang/wgsl.rs:184-200 (bilinear between
 *     texel centres in f32, clamped at the edges);
 *   density effects: assets/effects/density_blur.rhai, bilateral_blur.rhai,
 *     sharpen.rhai (their WGSL predecessors under shaders/effects/density fix the same pixels);
 *   colour effects: assets/effects/vignette.rhai, film_grain.rhai,
 *     chromatic_aberration.rhai, hue_shift.rhai, sobel_edges.rhai, their
 *     colour blended by apply_blend (shaders/effects/common/
 *     blend_modes.wgsl) over the input, the input's alpha kept
 *     (src/effects/custom.rs:49-60), stored as RGBA8;
 *   the chains: EffectChainRunner::run_density_effects /
 *     run_color_effects (src/renderer/effect_chain.rs:737-962): each
 *     enabled effect reads the previous one's output.
 * A pixel's uv is its centre, ((x + 0.5) / W, (y + 0.5) / H); its own
 * colour `c` is its texel. All arithmetic is f32 in source order
 * (-ffp-contract=off); exp, sqrt and sin are the C library's.
 */
#include <math.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>

#include "ffm_oracle.h"

typedef struct {
  float v[4];
} px4;

typedef struct {
  const float* f32; /* W*H*4, or NULL */
  const uint8_t* u8;
  int W, H;
} img_t;

static px4 img_at(const img_t* im, int x, int y) {
  px4 r;
  const size_t i = 4 * ((size_t)y * im->W + x);
  for (int c = 0; c < 4; c++)
    r.v[c] = im->f32 ? im->f32[i + c] : (float)im->u8[i + c] / 255.0f;
  return r;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
static float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

/* uf_texture */
static px4 tex(const img_t* im, float u, float v) {
  float px = u * (float)im->W - 0.5f, py = v * (float)im->H - 0.5f;
  px = fabsf(px) <= 1e30f ? fminf(fmaxf(px, -1.0f), (float)im->W) : -1.0f;
  py = fabsf(py) <= 1e30f ? fminf(fmaxf(py, -1.0f), (float)im->H) : -1.0f;
  const float fx = floorf(px), fy = floorf(py);
  const float tx = px - fx, ty = py - fy;
  const int x0 = clampi((int)fx, 0, im->W - 1), y0 = clampi((int)fy, 0, im->H - 1);
  const int x1 = clampi((int)fx + 1, 0, im->W - 1), y1 = clampi((int)fy + 1, 0, im->H - 1);
  const px4 a = img_at(im, x0, y0), b = img_at(im, x1, y0);
  const px4 c = img_at(im, x0, y1), d = img_at(im, x1, y1);
  px4 r;
  for (int k = 0; k < 4; k++) {
    const float top = a.v[k] + (b.v[k] - a.v[k]) * tx;
    const float bottom = c.v[k] + (d.v[k] - c.v[k]) * tx;
    r.v[k] = top + (bottom - top) * ty;
  }
  return r;
}

/* smoothstep as the Rhai effects define it */
static float smooth(float a, float b, float x) {
  const float t = clamp01((x - a) / (b - a));
  return t * t * (3.0f - 2.0f * t);
}

static float lum3(float r, float g, float b) { return r * 0.2126f + g * 0.7152f + b * 0.0722f; }

/* sin and cos for the colour effects, by the method the kernels use:
 * Cody-Waite reduction by pi/2 (pi/2 split into three floats: its float
 * rounding 1.57079637050628662109375 and the rounded remainders
 * -4.371138828673793e-8 and -1.7151245100058819e-15, with fused
 * multiply-adds) and the Cephes sinf/cosf minimax polynomials on
 * [-pi/4, pi/4]; beyond |x| = 32768 the reduction in f64 (two-part pi/2)
 * before the same polynomials. Pinned against the mathematical sin/cos
 * to a few ulp by the tests. */
void orc_fx_sincos(float x, float* s, float* c) {
  float r;
  int q;
  if (fabsf(x) <= 32768.0f) {
    const float k = rintf(x * 0.636619772367581343f);
    r = fmaf(-k, 1.57079637050628662109375f, x);
    r = fmaf(-k, -4.37113882867379300e-8f, r);
    r = fmaf(-k, -1.71512451000588190e-15f, r);
    q = (int)k & 3;
  } else if (!(fabsf(x) <= 3.4028235e38f)) {
    *s = *c = NAN;
    return;
  } else {
    /* beyond 32768: the reduction in f64, two-part pi/2 */
    const double xd = (double)x, k = rint(xd * 0.63661977236758134);
    double rd = fma(-k, 1.5707963267948966, xd);
    rd = fma(-k, 6.123233995736766e-17, rd);
    r = (float)rd;
    q = (int)fmod(k, 4.0) & 3;
  }
  const float z = r * r;
  const float sp = fmaf(fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f),
                        z * r, r);
  const float cp = fmaf(fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z,
                             4.166664568298827e-2f),
                        z * z, fmaf(-0.5f, z, 1.0f));
  const float ss = (q & 1) ? cp : sp, cc = (q & 1) ? sp : cp;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

static float fx_sin(float x) {
  float s, c;
  orc_fx_sincos(x, &s, &c);
  return s;
}

static float fx_cos(float x) {
  float s, c;
  orc_fx_sincos(x, &s, &c);
  return c;
}
static float len2(float x, float y) { return sqrtf(x * x + y * y); }

/* ---- density effects ---------------------------------------------------- */

static px4 density_blur(const img_t* im, const float* p, float dm, float ps, px4 c, float u,
                        float v) {
  const float max_radius = p[0] * ps;
  const float density = c.v[3] / fmaxf(dm, 0.000001f);
  const float blur = density < p[1] ? 1.0f : 1.0f - smooth(p[1], p[1] + p[2], density);
  const float radius = max_radius * blur;
  if (radius < 0.5f) return c;
  const float px = 1.0f / (float)im->W, py = 1.0f / (float)im->H;
  const float sigma = radius / 2.0f;
  const float step = fmaxf(1.0f, radius / 4.0f);
  float sum[4] = {0, 0, 0, 0}, weights = 0.0f;
  for (int yi = -4; yi < 5; yi++)
    for (int xi = -4; xi < 5; xi++) {
      const float ox = (float)xi * step, oy = (float)yi * step;
      const float dist = len2(ox, oy);
      if (dist <= radius) {
        const float w = expf(-(dist * dist) / (2.0f * sigma * sigma));
        const px4 s = tex(im, u + ox * px, v + oy * py);
        for (int k = 0; k < 4; k++) sum[k] = sum[k] + s.v[k] * w;
        weights += w;
      }
    }
  if (weights <= 0.0f) return c;
  px4 r;
  for (int k = 0; k < 4; k++) r.v[k] = sum[k] / weights;
  return r;
}

static float gaussian(float x, float sigma) { return expf(-(x * x) / (2.0f * sigma * sigma)); }

static px4 bilateral_blur(const img_t* im, const float* p, float ps, px4 c, float u, float v) {
  const int radius = (int)fminf(fmaxf(p[0] * ps, 1.0f), 10.0f);
  const float sigma_spatial = fmaxf(0.1f, p[1] * ps);
  const float sigma_range = fmaxf(0.01f, p[2]);
  const float px = 1.0f / (float)im->W, py = 1.0f / (float)im->H;
  const float centre = lum3(c.v[0], c.v[1], c.v[2]);
  float sum[4] = {0, 0, 0, 0}, weights = 0.0f;
  for (int y = -radius; y < radius + 1; y++)
    for (int x = -radius; x < radius + 1; x++) {
      const float fx = (float)x, fy = (float)y;
      const px4 s = tex(im, u + fx * px, v + fy * py);
      const float w = gaussian(len2(fx, fy), sigma_spatial) *
                      gaussian(fabsf(centre - lum3(s.v[0], s.v[1], s.v[2])), sigma_range);
      for (int k = 0; k < 4; k++) sum[k] = sum[k] + s.v[k] * w;
      weights += w;
    }
  px4 r;
  for (int k = 0; k < 4; k++) r.v[k] = sum[k] / fmaxf(weights, 0.0001f);
  return r;
}

static px4 sharpen(const img_t* im, const float* p, float ps, px4 c, float u, float v) {
  const float radius = p[1] * ps;
  const float px = 1.0f / (float)im->W, py = 1.0f / (float)im->H;
  const float sigma = radius / 2.0f;
  const int kk = (int)ceilf(radius);
  float sum[4] = {0, 0, 0, 0}, weights = 0.0f;
  for (int y = -kk; y < kk + 1; y++)
    for (int x = -kk; x < kk + 1; x++) {
      const float fx = (float)x, fy = (float)y;
      const float dist = len2(fx, fy);
      if (dist <= radius) {
        const float w = expf(-(dist * dist) / (2.0f * sigma * sigma));
        const px4 s = tex(im, u + fx * px, v + fy * py);
        for (int k = 0; k < 4; k++) sum[k] = sum[k] + s.v[k] * w;
        weights += w;
      }
    }
  if (p[0] < 0.01f) return c;
  px4 r;
  for (int k = 0; k < 3; k++) {
    const float blurred = sum[k] / weights;
    r.v[k] = clamp01(c.v[k] + (c.v[k] - blurred) * p[0]);
  }
  r.v[3] = c.v[3];
  return r;
}

void orc_density_effect(const ffm_effect* fx, const float* in, float* out, int W, int H,
                        float density_mean, float pixel_scale) {
  const img_t im = {in, NULL, W, H};
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const float u = ((float)x + 0.5f) / (float)W, v = ((float)y + 0.5f) / (float)H;
      const px4 c = img_at(&im, x, y);
      px4 o = c;
      switch (fx->type) {
        case FFM_FX_DENSITY_BLUR:
          o = density_blur(&im, fx->params, density_mean, pixel_scale, c, u, v);
          break;
        case FFM_FX_BILATERAL_BLUR: o = bilateral_blur(&im, fx->params, pixel_scale, c, u, v); break;
        case FFM_FX_SHARPEN: o = sharpen(&im, fx->params, pixel_scale, c, u, v); break;
        default: break;
      }
      memcpy(out + 4 * ((size_t)y * W + x), o.v, sizeof o.v);
    }
}

/* ---- the blend library (blend_modes.wgsl) -------------------------------- */

static void rgb_to_hsl(const float c[3], float o[3]) {
  const float cmax = fmaxf(c[0], fmaxf(c[1], c[2])), cmin = fminf(c[0], fminf(c[1], c[2]));
  const float delta = cmax - cmin;
  const float l = (cmax + cmin) * 0.5f;
  if (delta < 0.00001f) {
    o[0] = 0.0f;
    o[1] = 0.0f;
    o[2] = l;
    return;
  }
  /* select(f, t, cond): t when cond */
  const float s = l > 0.5f ? delta / (2.0f - cmax - cmin) : delta / (cmax + cmin);
  float h;
  if (cmax == c[0]) {
    h = (c[1] - c[2]) / delta;
    if (c[1] < c[2]) h += 6.0f;
  } else if (cmax == c[1]) {
    h = (c[2] - c[0]) / delta + 2.0f;
  } else {
    h = (c[0] - c[1]) / delta + 4.0f;
  }
  o[0] = h / 6.0f;
  o[1] = s;
  o[2] = l;
}

static float hue_to_rgb(float p, float q, float t) {
  if (t < 0.0f) t += 1.0f;
  if (t > 1.0f) t -= 1.0f;
  if (t < 1.0f / 6.0f) return p + (q - p) * 6.0f * t;
  if (t < 1.0f / 2.0f) return q;
  if (t < 2.0f / 3.0f) return p + (q - p) * (2.0f / 3.0f - t) * 6.0f;
  return p;
}

static void hsl_to_rgb(const float hsl[3], float o[3]) {
  const float h = hsl[0], s = hsl[1], l = hsl[2];
  if (s < 0.00001f) {
    o[0] = o[1] = o[2] = l;
    return;
  }
  const float q = l < 0.5f ? l * (1.0f + s) : l + s - l * s;
  const float p = 2.0f * l - q;
  o[0] = hue_to_rgb(p, q, h + 1.0f / 3.0f);
  o[1] = hue_to_rgb(p, q, h);
  o[2] = hue_to_rgb(p, q, h - 1.0f / 3.0f);
}

void orc_apply_blend(const float base[3], const float e[3], int mode, float amount, float out[3]) {
  float b[3];
  if (mode >= 9 && mode <= 12) {
    float hb[3], he[3], hm[3];
    rgb_to_hsl(base, hb);
    rgb_to_hsl(e, he);
    if (mode == 9) { hm[0] = he[0]; hm[1] = hb[1]; hm[2] = hb[2]; }       /* hue */
    else if (mode == 10) { hm[0] = hb[0]; hm[1] = he[1]; hm[2] = hb[2]; } /* saturation */
    else if (mode == 11) { hm[0] = he[0]; hm[1] = he[1]; hm[2] = hb[2]; } /* color */
    else { hm[0] = hb[0]; hm[1] = hb[1]; hm[2] = he[2]; }                 /* luminosity */
    hsl_to_rgb(hm, b);
  } else {
    for (int i = 0; i < 3; i++) {
      const float x = base[i], y = e[i];
      float r;
      switch (mode) {
        case 1: r = fminf(x + y, 1.0f); break;                         /* add */
        case 2: r = x * y; break;                                       /* multiply */
        case 3: r = 1.0f - (1.0f - x) * (1.0f - y); break;              /* screen */
        case 4:                                                          /* overlay */
          r = x > 0.5f ? 1.0f - 2.0f * (1.0f - x) * (1.0f - y) : 2.0f * x * y;
          break;
        case 5: {                                                        /* soft light */
          const float d = x > 0.25f ? sqrtf(x) : ((16.0f * x - 12.0f) * x + 4.0f) * x;
          r = y > 0.5f ? x + (2.0f * y - 1.0f) * (d - x) : x - (1.0f - 2.0f * y) * x * (1.0f - x);
          break;
        }
        case 6:                                                          /* hard light */
          r = y > 0.5f ? 1.0f - 2.0f * (1.0f - x) * (1.0f - y) : 2.0f * x * y;
          break;
        case 7: r = y >= 1.0f ? 1.0f : fminf(1.0f, x / (1.0f - y + 0.00001f)); break;
        case 8: r = y <= 0.0f ? 0.0f : 1.0f - fminf(1.0f, (1.0f - x) / (y + 0.00001f)); break;
        default: r = y; break;                                           /* normal */
      }
      b[i] = r;
    }
  }
  /* mix(base, blended, intensity) */
  for (int i = 0; i < 3; i++) out[i] = base[i] * (1.0f - amount) + b[i] * amount;
}

/* ---- colour effects ------------------------------------------------------ */

#define FX_PI 3.14159265359f

/* the effect language's atan2 (src/lang/wgsl.rs uf_atan2): IEEE at (0, 0) */
static float uf_atan2(float y, float x) {
  if (y == 0.0f && x == 0.0f) {
    const float mag = signbit(x) ? 3.14159265358979f : 0.0f;
    return signbit(y) ? -mag : mag;
  }
  return atan2f(y, x);
}

static float fractf(float x) { return x - floorf(x); }

/* hsv_to_rgb as julia.rhai and worley_noise.rhai write it */
static void fx_hsv(float h, float s, float v, float o[3]) {
  const float c = v * s, h6 = h * 6.0f;
  const float x = c * (1.0f - fabsf(h6 - 2.0f * truncf(h6 / 2.0f) - 1.0f));
  const float m = v - c;
  float r, g, b;
  if (h6 < 1.0f) { r = c; g = x; b = 0.0f; }
  else if (h6 < 2.0f) { r = x; g = c; b = 0.0f; }
  else if (h6 < 3.0f) { r = 0.0f; g = c; b = x; }
  else if (h6 < 4.0f) { r = 0.0f; g = x; b = c; }
  else if (h6 < 5.0f) { r = x; g = 0.0f; b = c; }
  else { r = c; g = 0.0f; b = x; }
  o[0] = r + m;
  o[1] = g + m;
  o[2] = b + m;
}

/* the effect language's hash2 (src/lang/wgsl.rs uf_hash2) */
static void fx_hash2(float x, float y, float* hx, float* hy) {
  const float n0 = x * 127.1f + y * 311.7f, n1 = x * 269.5f + y * 183.3f;
  *hx = fractf(fx_sin(n0) * 43758.5453f);
  *hy = fractf(fx_sin(n1) * 43758.5453f);
}

/* domain_warp.rhai's gradient noise */
static float dw_grad(float lx, float ly, float ox, float oy) {
  float gx, gy;
  fx_hash2(lx, ly, &gx, &gy);
  return (gx * 2.0f - 1.0f) * ox + (gy * 2.0f - 1.0f) * oy;
}

static float dw_noise(float x, float y) {
  const float ix = floorf(x), iy = floorf(y);
  const float fx = fractf(x), fy = fractf(y);
  const float ux = fx * fx * (3.0f - 2.0f * fx), uy = fy * fy * (3.0f - 2.0f * fy);
  const float a = dw_grad(ix, iy, fx, fy);
  const float b = dw_grad(ix + 1.0f, iy, fx - 1.0f, fy);
  const float c = dw_grad(ix, iy + 1.0f, fx, fy - 1.0f);
  const float d = dw_grad(ix + 1.0f, iy + 1.0f, fx - 1.0f, fy - 1.0f);
  const float ab = a * (1.0f - ux) + b * ux, cd = c * (1.0f - ux) + d * ux;
  return ab * (1.0f - uy) + cd * uy;
}

/* simplex_noise.rhai's corners and simplex */
static float sx_corner(float lx, float ly, float ox, float oy) {
  const float t = 0.5f - (ox * ox + oy * oy);
  if (!(t >= 0.0f)) return 0.0f;
  float gx, gy;
  fx_hash2(lx, ly, &gx, &gy);
  const float t2 = t * t;
  return t2 * t2 * ((gx * 2.0f - 1.0f) * ox + (gy * 2.0f - 1.0f) * oy);
}

static float sx_noise(float x, float y) {
  const float f2 = 0.36602540378443865f, g2 = 0.21132486540518713f;
  const float s = (x + y) * f2;
  const float ix = floorf(x + s), iy = floorf(y + s);
  const float t = (ix + iy) * g2;
  const float x0 = x - (ix - t), y0 = y - (iy - t);
  const float i1x = x0 > y0 ? 1.0f : 0.0f, i1y = x0 > y0 ? 0.0f : 1.0f;
  const float n0 = sx_corner(ix, iy, x0, y0);
  const float n1 = sx_corner(ix + i1x, iy + i1y, x0 - i1x + g2, y0 - i1y + g2);
  const float n2 = sx_corner(ix + 1.0f, iy + 1.0f, x0 - 1.0f + 2.0f * g2, y0 - 1.0f + 2.0f * g2);
  return 70.0f * (n0 + n1 + n2);
}

static float fx_fbm(int simplex, float x, float y, float octaves) {
  float value = 0.0f, amplitude = 0.5f, px = x, py = y;
  for (int o = 0; (float)o < octaves; o++) {
    value += amplitude * (simplex ? sx_noise(px, py) : dw_noise(px, py));
    const float nx = (0.8f * px + 0.6f * py) * 2.02f;
    const float ny = (-0.6f * px + 0.8f * py) * 2.02f;
    px = nx;
    py = ny;
    amplitude *= 0.5f;
  }
  return value;
}

static uint8_t unorm8(float v) {
  v = (v != v) ? 0.0f : clamp01(v);
  return (uint8_t)floorf(v * 255.0f + 0.5f);
}

/* ---- bloom (assets/effects/bloom.rhai; passes by src/renderer/effect_passes.rs:273-371) ---- */

static px4 p4add(px4 a, px4 b) {
  for (int k = 0; k < 4; k++) a.v[k] = a.v[k] + b.v[k];
  return a;
}
static px4 p4mul(px4 a, float s) {
  for (int k = 0; k < 4; k++) a.v[k] = a.v[k] * s;
  return a;
}

static px4 bloom_bright(px4 c, float threshold, float softness) {
  const float l = fmaxf(c.v[0], fmaxf(c.v[1], c.v[2]));
  const float knee = fmaxf(threshold * softness, 0.00001f);
  const float soft = fminf(fmaxf(l - threshold + knee, 0.0f), 2.0f * knee);
  const float w = fmaxf(soft * soft / (4.0f * knee), l - threshold) / fmaxf(l, 0.00001f);
  return p4mul(c, w);
}

#define ORC_BLOOM_LEVELS 7

/* Buffers "down" and "up": scale 0.5 (rounded half away from zero), up to
 * 7 levels and no more than the longer side's bit length
 * (src/effects/passes.rs:21-33); level l is the base size >> l. */
static void orc_bloom(const float* p, const uint8_t* in, uint8_t* out, int W, int H) {
  const int bw = (int)fmaxf(roundf((float)W * 0.5f), 1.0f), bh = (int)fmaxf(roundf((float)H * 0.5f), 1.0f);
  int bits = 0;
  for (int m = bw > bh ? bw : bh; m; m >>= 1) bits++;
  const int n = bits < ORC_BLOOM_LEVELS ? (bits < 1 ? 1 : bits) : ORC_BLOOM_LEVELS;
  int lw[ORC_BLOOM_LEVELS], lh[ORC_BLOOM_LEVELS];
  float *down[ORC_BLOOM_LEVELS], *up[ORC_BLOOM_LEVELS];
  for (int l = 0; l < n; l++) {
    lw[l] = bw >> l > 0 ? bw >> l : 1;
    lh[l] = bh >> l > 0 ? bh >> l : 1;
    down[l] = (float*)malloc((size_t)lw[l] * lh[l] * 16);
    up[l] = (float*)malloc((size_t)lw[l] * lh[l] * 16);
  }
  const img_t src = {NULL, in, W, H};
  const float th = p[0], so = p[1], inten = p[2], spread = p[3];
  for (int r = 0; r < n; r++) {
    const int w = lw[r], h = lh[r];
    const img_t prev = {r ? down[r - 1] : NULL, NULL, r ? lw[r - 1] : 1, r ? lh[r - 1] : 1};
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        const float u = ((float)x + 0.5f) / (float)w, v = ((float)y + 0.5f) / (float)h;
        const float dx = 0.5f / (float)w, dy = 0.5f / (float)h;
        px4 o;
        if (r == 0) {
          o = p4add(p4add(p4add(bloom_bright(tex(&src, u - dx, v - dy), th, so),
                                bloom_bright(tex(&src, u + dx, v - dy), th, so)),
                          bloom_bright(tex(&src, u - dx, v + dy), th, so)),
                    bloom_bright(tex(&src, u + dx, v + dy), th, so));
          o = p4mul(o, 0.25f);
        } else {
          o = p4add(p4add(p4add(p4add(p4mul(tex(&prev, u, v), 4.0f), tex(&prev, u - dx, v - dy)),
                                tex(&prev, u + dx, v - dy)),
                          tex(&prev, u - dx, v + dy)),
                    tex(&prev, u + dx, v + dy));
          o = p4mul(o, 0.125f);
        }
        memcpy(down[r] + 4 * ((size_t)y * w + x), o.v, 16);
      }
  }
  for (int l = n - 1; l >= 0; l--) {
    const int w = lw[l], h = lh[l];
    const img_t dn = {down[l], NULL, w, h};
    const img_t above = {l + 1 < n ? up[l + 1] : NULL, NULL, l + 1 < n ? lw[l + 1] : 1,
                         l + 1 < n ? lh[l + 1] : 1};
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        const float u = ((float)x + 0.5f) / (float)w, v = ((float)y + 0.5f) / (float)h;
        const px4 here = tex(&dn, u, v);
        px4 o = here;
        if (l != n - 1) {
          const float dx = 1.0f / (float)w, dy = 1.0f / (float)h;
          const px4 side = p4add(p4add(p4add(tex(&above, u - dx, v), tex(&above, u + dx, v)),
                                       tex(&above, u, v - dy)),
                                 tex(&above, u, v + dy));
          px4 tent = p4add(p4mul(tex(&above, u, v), 4.0f), p4mul(side, 2.0f));
          tent = p4add(tent, tex(&above, u - dx, v - dy));
          tent = p4add(tent, tex(&above, u + dx, v - dy));
          tent = p4add(tent, tex(&above, u - dx, v + dy));
          tent = p4add(tent, tex(&above, u + dx, v + dy));
          o = p4add(here, p4mul(tent, 0.0625f * spread));
        }
        memcpy(up[l] + 4 * ((size_t)y * w + x), o.v, 16);
      }
  }
  const img_t glow_img = {up[0], NULL, lw[0], lh[0]};
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const float u = ((float)x + 0.5f) / (float)W, v = ((float)y + 0.5f) / (float)H;
      const px4 c = img_at(&src, x, y), glow = tex(&glow_img, u, v);
      uint8_t* o = out + 4 * ((size_t)y * W + x);
      for (int k = 0; k < 3; k++) o[k] = unorm8(clamp01(c.v[k] + glow.v[k] * 1.0f * inten));
      o[3] = in[4 * ((size_t)y * W + x) + 3];
    }
  for (int l = 0; l < n; l++) {
    free(down[l]);
    free(up[l]);
  }
}

/* ---- distance_field (assets/effects/distance_field.rhai): seed, jump
 * flood over log2_passes (src/effects/passes.rs:110-113) with step
 * 2^(n-1-r), shade. A pixel's coordinates are its centre. ---- */
static px4 texel_u8(const uint8_t* img, int W, int H, int x, int y) {
  const img_t im = {NULL, img, W, H};
  return img_at(&im, clampi(x, 0, W - 1), clampi(y, 0, H - 1));
}

static void orc_distance_field(const float* p, const uint8_t* in, uint8_t* out, int W, int H) {
  float* a = (float*)malloc((size_t)W * H * 8);
  float* b = (float*)malloc((size_t)W * H * 8);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const px4 h = texel_u8(in, W, H, x, y);
      const int seed = lum3(h.v[0], h.v[1], h.v[2]) > p[0];
      a[2 * ((size_t)y * W + x)] = seed ? (float)x + 0.5f : -1.0f;
      a[2 * ((size_t)y * W + x) + 1] = seed ? (float)y + 0.5f : -1.0f;
    }
  int m = (W > H ? W : H);
  if (m < 2) m = 2;
  int n = 0;
  for (unsigned v = (unsigned)(m - 1); v; v >>= 1) n++;
  for (int r = 0; r < n; r++) {
    const int step = 1 << (n - 1 - r);
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) {
        const float px = (float)x + 0.5f, py = (float)y + 0.5f;
        float bx = -1.0f, by = -1.0f, best_d = 1.0e30f;
        for (int dy = -1; dy < 2; dy++)
          for (int dx = -1; dx < 2; dx++) {
            const int qx = x + dx * step, qy = y + dy * step;
            if (qx >= 0 && qy >= 0 && qx < W && qy < H) {
              const float sx = a[2 * ((size_t)qy * W + qx)], sy = a[2 * ((size_t)qy * W + qx) + 1];
              if (sx >= 0.0f) {
                const float d = len2(sx - px, sy - py);
                if (d < best_d) {
                  best_d = d;
                  bx = sx;
                  by = sy;
                }
              }
            }
          }
        b[2 * ((size_t)y * W + x)] = bx;
        b[2 * ((size_t)y * W + x) + 1] = by;
      }
    float* t = a;
    a = b;
    b = t;
  }
  const float spread = p[1], mode = p[2], inten = p[3], bands = p[4];
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const float px = (float)x + 0.5f, py = (float)y + 0.5f;
      const px4 base = texel_u8(in, W, H, x, y);
      const float sx = a[2 * ((size_t)y * W + x)], sy = a[2 * ((size_t)y * W + x) + 1];
      float rgb[3] = {base.v[0], base.v[1], base.v[2]};
      if (sx >= 0.0f) {
        const float range_px = fmaxf(spread, 0.01f) * (float)(W > H ? W : H);
        const float d = len2(sx - px, sy - py);
        const float t = clamp01(d / range_px);
        const px4 sc = texel_u8(in, W, H, (int)sx, (int)sy);
        const float off = d > 0.5f ? 1.0f : 0.0f;
        for (int k = 0; k < 3; k++) {
          if (mode == 0.0f) {
            rgb[k] = base.v[k] + sc.v[k] * expf(-4.0f * t) * off * inten;
          } else if (mode == 1.0f) {
            const float phase = fractf(t * fmaxf(bands, 1.0f));
            const float ring = smooth(0.0f, 0.15f, phase) * (1.0f - smooth(0.35f, 0.5f, phase));
            rgb[k] = base.v[k] + sc.v[k] * ring * (1.0f - t) * inten * off;
          } else if (d > 0.5f) {
            const float mm = clamp01(inten) * (1.0f - t);
            rgb[k] = base.v[k] * (1.0f - mm) + sc.v[k] * mm;
          }
        }
      }
      uint8_t* o = out + 4 * ((size_t)y * W + x);
      for (int k = 0; k < 3; k++) o[k] = unorm8(rgb[k]);
      o[3] = in[4 * ((size_t)y * W + x) + 3];
    }
  free(a);
  free(b);
}

void orc_color_effect(const ffm_effect* fx, const uint8_t* in, uint8_t* out, int W, int H,
                      float pixel_scale) {
  if (fx->type == FFM_FX_DISTANCE_FIELD) {
    orc_distance_field(fx->params, in, out, W, H);
    return;
  }
  if (fx->type == FFM_FX_TRAILS) {
    /* trails.rhai on a still: history zero, so keep() stores the picture
     * (max(c, 0)) and effect() returns c + (trail - c) * intensity = c */
    memcpy(out, in, (size_t)W * H * 4);
    return;
  }
  if (fx->type == FFM_FX_BLOOM) {
    orc_bloom(fx->params, in, out, W, H);
    return;
  }
  const img_t im = {NULL, in, W, H};
  const float* p = fx->params;
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const float u = ((float)x + 0.5f) / (float)W, v = ((float)y + 0.5f) / (float)H;
      const px4 c = img_at(&im, x, y);
      float e[3] = {c.v[0], c.v[1], c.v[2]};
      int mode = 0;
      float amount = 1.0f;
      switch (fx->type) {
        case FFM_FX_VIGNETTE: {
          const float aspect = (float)W / (float)H;
          const float dist = len2((u - 0.5f) * aspect, v - 0.5f);
          const float start = p[1] * 0.5f * fmaxf(1.0f, aspect);
          const float end = start + p[2] * 0.5f * fmaxf(1.0f, aspect);
          const float vig = 1.0f - smooth(start, end, dist);
          for (int k = 0; k < 3; k++) e[k] = c.v[k] * vig;
          amount = p[0];
          mode = (int)p[3];
          break;
        }
        case FFM_FX_FILM_GRAIN: {
          const float ax = u * (float)W + p[1] * 127.1f, ay = v * (float)H + p[1] * 311.7f;
          const float n = ax * 127.1f + ay * 311.7f;
          const float s = fx_sin(n) * 43758.5453f;
          e[0] = e[1] = e[2] = s - floorf(s);
          amount = p[0];
          mode = (int)p[2];
          break;
        }
        case FFM_FX_CHROMATIC_ABERRATION: {
          const float a = p[0] * pixel_scale;
          const float px = 1.0f / (float)W, py = 1.0f / (float)H;
          float ox, oy;
          if (p[1] != 0.0f) {
            const float ex = u - 0.5f, ey = v - 0.5f;
            const float dist = len2(ex, ey) * 2.0f;
            const float vx = ex + 0.0001f, vy = ey + 0.0001f;
            const float vl = len2(vx, vy);
            ox = vx / vl * a * px * dist;
            oy = vy / vl * a * py * dist;
          } else {
            ox = a * px;
            oy = 0.0f;
          }
          e[0] = tex(&im, u + ox, v + oy).v[0];
          e[1] = c.v[1];
          e[2] = tex(&im, u - ox, v - oy).v[2];
          amount = p[2];
          mode = (int)p[3];
          break;
        }
        case FFM_FX_HUE_SHIFT: {
          const float r = c.v[0], g = c.v[1], b = c.v[2];
          const float cmax = fmaxf(r, fmaxf(g, b)), cmin = fminf(r, fminf(g, b));
          const float delta = cmax - cmin;
          float h = 0.0f, s = 0.0f;
          if (delta > 0.0001f) {
            s = delta / cmax;
            if (cmax == r) {
              h = (g - b) / delta;
              if (g < b) h += 6.0f;
            } else if (cmax == g) {
              h = 2.0f + (b - r) / delta;
            } else {
              h = 4.0f + (r - g) / delta;
            }
            h /= 6.0f;
          }
          float hx = h + p[0] / 360.0f;
          hx = hx - floorf(hx);
          const float hh = hx * 6.0f, vv = cmax;
          const float i = floorf(hh), f = hh - i;
          const float pp = vv * (1.0f - s), q = vv * (1.0f - s * f);
          const float t = vv * (1.0f - s * (1.0f - f));
          const float tab[6][3] = {{vv, t, pp}, {q, vv, pp}, {pp, vv, t},
                                   {pp, q, vv}, {t, pp, vv}, {vv, pp, q}};
          const int si = i == 0.0f ? 0 : i == 1.0f ? 1 : i == 2.0f ? 2 : i == 3.0f ? 3 : i == 4.0f ? 4 : 5;
          for (int k = 0; k < 3; k++) e[k] = tab[si][k];
          amount = p[1];
          mode = (int)p[2];
          break;
        }
        case FFM_FX_SOBEL_EDGES: {
          const float tx = 1.0f / (float)W, ty = 1.0f / (float)H;
          const float du[8] = {-tx, 0.0f, tx, -tx, tx, -tx, 0.0f, tx};
          const float dv[8] = {-ty, -ty, -ty, 0.0f, 0.0f, ty, ty, ty};
          float n[8];
          for (int k = 0; k < 8; k++) {
            const px4 s = tex(&im, u + du[k], v + dv[k]);
            n[k] = lum3(s.v[0], s.v[1], s.v[2]);
          }
          /* n0 n1 n2 / n3 . n5 / n6 n7 n8 */
          const float n0 = n[0], n1 = n[1], n2 = n[2], n3 = n[3], n5 = n[4], n6 = n[5];
          const float n7 = n[6], n8 = n[7];
          const float gx = n2 + 2.0f * n5 + n8 - (n0 + 2.0f * n3 + n6);
          const float gy = n0 + 2.0f * n1 + n2 - (n6 + 2.0f * n7 + n8);
          const float edge = smooth(p[1] * 0.5f, p[1], sqrtf(gx * gx + gy * gy));
          const float pc = p[3];
          const float gain = (1.0f + p[2] * 2.0f) * edge;
          for (int k = 0; k < 3; k++) e[k] = (1.0f * (1.0f - pc) + c.v[k] * pc) * gain;
          amount = p[0];
          mode = (int)p[4];
          break;
        }
        case FFM_FX_KALEIDOSCOPE: {
          const float two_pi = 6.28318530718f;
          const float segments = fmaxf(2.0f, p[0]);
          const float rotation = p[1] * FX_PI / 180.0f;
          const float zoom = fmaxf(0.1f, p[2]);
          const int square = p[5] > 0.0f;
          const float edge_offset = p[6] * pixel_scale;
          const float aspect = (float)W / (float)H;
          float x = (u - 0.5f) / zoom, y = (v - 0.5f) / zoom;
          if (square) x = x * aspect;
          const float radius = len2(x, y);
          float angle = uf_atan2(y, x);
          const float segment = two_pi / segments;
          angle = angle - segment * floorf(angle / segment);
          if (angle > segment * 0.5f) angle = segment - angle;
          angle = angle + rotation;
          float rx = fx_cos(angle) * radius;
          const float ry = fx_sin(angle) * radius;
          if (square) rx = rx / aspect;
          const float dx = -fx_cos(rotation), dy = -fx_sin(rotation);
          const float uu = rx + 0.5f + dx * edge_offset * (1.0f / (float)W);
          const float vv = ry + 0.5f + dy * edge_offset * (1.0f / (float)H);
          const px4 s = tex(&im, fminf(fmaxf(uu, 0.001f), 0.999f), fminf(fmaxf(vv, 0.001f), 0.999f));
          for (int q = 0; q < 3; q++) e[q] = s.v[q];
          amount = p[3];
          mode = (int)p[4];
          break;
        }
        case FFM_FX_PLASMA: {
          const float scale = fmaxf(0.5f, p[1]);
          const float time = p[3] * p[2];
          const float direction = p[5] * FX_PI / 180.0f;
          const float x = u * scale, y = v * scale;
          const float dir_phase = x * fx_cos(direction) + y * fx_sin(direction);
          const float t = time + dir_phase;
          float w = fx_sin(x + t);
          w += fx_sin((y + t) * 0.5f);
          w += fx_sin((x + y + t) * 0.5f);
          const float cx = x + 0.5f * fx_sin(t * 0.25f), cy = y + 0.5f * fx_cos(t * 0.5f);
          w += fx_sin(sqrtf(cx * cx + cy * cy + 1.0f) + t);
          w = w * 0.25f;
          e[0] = fx_sin(w * FX_PI) * 0.5f + 0.5f;
          e[1] = fx_sin(w * FX_PI + 2.094f) * 0.5f + 0.5f;
          e[2] = fx_sin(w * FX_PI + 4.189f) * 0.5f + 0.5f;
          amount = p[0];
          mode = (int)p[4];
          break;
        }
        case FFM_FX_TUNNEL: {
          const float aspect = (float)W / (float)H;
          const float x = (u - 0.5f) * aspect, y = v - 0.5f;
          const float angle = uf_atan2(y, x) / FX_PI;
          const float radius = len2(x, y);
          const float wall = fmaxf(radius, 0.001f);
          const float r = wall + p[2] * fx_sin(angle * FX_PI * 4.0f + p[3]) * 0.1f;
          const float au = fractf(angle * 0.5f + 0.5f + p[3] * p[1] * 0.1f);
          const float av = fractf(1.0f / r + p[3] * p[0]);
          const px4 s = tex(&im, au, av);
          const float fade = smooth(0.0f, 0.5f, radius);
          for (int q = 0; q < 3; q++) e[q] = s.v[q] * fade;
          amount = radius < 0.001f ? 0.0f : p[4];
          mode = (int)p[5];
          break;
        }
        case FFM_FX_JULIA: {
          const float max_iter = fmaxf(10.0f, p[3]);
          const float zoom = fmaxf(0.1f, p[4]);
          const float aspect = (float)W / (float)H;
          const float px = (u - 0.5f) * 2.0f * aspect / zoom + p[5];
          const float py = (v - 0.5f) * 2.0f / zoom + p[6];
          const int is_julia = p[0] == 1.0f;
          float zr = is_julia ? px : 0.0f, zi = is_julia ? py : 0.0f;
          const float kr = is_julia ? p[1] : px, ki = is_julia ? p[2] : py;
          int it = 0;
          float x2 = zr * zr, y2 = zi * zi;
          while ((float)it < max_iter && !(x2 + y2 > 4.0f)) {
            const float nr = x2 - y2 + kr;
            const float ni = 2.0f * zr * zi + ki;
            zr = nr;
            zi = ni;
            x2 = zr * zr;
            y2 = zi * zi;
            it += 1;
          }
          const float iter = (float)it;
          const float zmag = zr * zr + zi * zi;
          if (iter >= max_iter) {
            e[0] = e[1] = e[2] = 0.0f;
          } else if (p[7] == 1.0f) {
            const float ln2 = logf(2.0f);
            const float log_zn = logf(zmag) / 2.0f;
            const float nu = logf(log_zn / ln2) / ln2;
            const float t = (iter + 1.0f - nu) / max_iter;
            fx_hsv(fractf(t * 5.0f), 0.8f, 1.0f, e);
          } else if (p[7] == 2.0f) {
            const float f = iter / max_iter;
            for (int q = 0; q < 3; q++) e[q] = c.v[q] * f;
          } else {
            const float t = iter / max_iter;
            fx_hsv(fractf(t * 3.0f), 0.7f, t, e);
          }
          amount = p[8];
          mode = (int)p[9];
          break;
        }
        case FFM_FX_WORLEY_NOISE: {
          const float scale = fmaxf(1.0f, p[1]);
          const float time = p[2];
          const float direction = p[6] * FX_PI / 180.0f;
          const float ax = (u + fx_cos(direction) * time * 0.1f) * scale;
          const float ay = (v + fx_sin(direction) * time * 0.1f) * scale;
          const float cx = floorf(ax), cy = floorf(ay);
          const float fx = fractf(ax), fy = fractf(ay);
          float d1 = 8.0f, d2 = 8.0f;
          for (int j = -1; j < 2; j++)
            for (int i = -1; i < 2; i++) {
              const float nx = (float)i, ny = (float)j;
              float hx, hy;
              fx_hash2(cx + nx, cy + ny, &hx, &hy);
              const float qx = 0.5f + 0.5f * fx_sin(time * 0.5f + 6.2831f * hx);
              const float qy = 0.5f + 0.5f * fx_sin(time * 0.5f + 6.2831f * hy);
              const float dist = len2(nx + qx - fx, ny + qy - fy);
              if (dist < d1) {
                d2 = d1;
                d1 = dist;
              } else if (dist < d2) {
                d2 = dist;
              }
            }
          const float m = p[3];
          const float pattern = m == 1.0f   ? smooth(0.0f, 0.15f, d2 - d1)
                                : m == 2.0f ? 1.0f - smooth(0.0f, 0.5f, d1)
                                : m == 3.0f ? clamp01(d1 * d2 * 2.0f)
                                            : d1;
          if (p[4] == 1.0f) fx_hsv(fractf(d1 + time * 0.1f), 0.8f, pattern, e);
          else if (p[4] == 2.0f) for (int q = 0; q < 3; q++) e[q] = c.v[q] * pattern;
          else e[0] = e[1] = e[2] = pattern;
          amount = p[0];
          mode = (int)p[5];
          break;
        }
        case FFM_FX_DOMAIN_WARP: {
          const float scale = fmaxf(0.5f, p[1]);
          const float octaves = fminf(fmaxf(p[2], 1.0f), 6.0f);
          const float direction = p[5] * FX_PI / 180.0f;
          const float tx = fx_cos(direction) * p[3] * 0.1f, ty = fx_sin(direction) * p[3] * 0.1f;
          const float x = u * scale, y = v * scale;
          const float qx = fx_fbm(0, x + 0.0f + tx, y + 0.0f + ty, octaves);
          const float qy = fx_fbm(0, x + 5.2f + tx * 1.2f, y + 1.3f + ty * 1.2f, octaves);
          const float rx = fx_fbm(0, x + 4.0f * qx + 1.7f, y + 4.0f * qy + 9.2f, octaves);
          const float ry = fx_fbm(0, x + 4.0f * qx + 8.3f, y + 4.0f * qy + 2.8f, octaves);
          const px4 s = tex(&im, clamp01(u + rx * p[0]), clamp01(v + ry * p[0]));
          for (int q = 0; q < 3; q++) e[q] = s.v[q];
          amount = p[0] * 2.0f;
          mode = (int)p[4];
          break;
        }
        case FFM_FX_SIMPLEX_NOISE: {
          const float scale = fmaxf(0.5f, p[1]);
          const float octaves = fminf(fmaxf(p[2], 1.0f), 6.0f);
          const float time = p[3];
          const float direction = p[6] * FX_PI / 180.0f;
          const float nx = u * scale + fx_cos(direction) * time * 0.1f;
          const float ny = v * scale + fx_sin(direction) * time * 0.1f;
          const float noise = fx_fbm(1, nx, ny, octaves);
          const float noise2 = fx_fbm(1, nx + 5.2f, ny + 1.3f, octaves);
          amount = p[0];
          mode = (int)p[5];
          if (p[4] == 1.0f) {
            /* blend -1: the warped read written as is, its alpha too */
            const px4 w = tex(&im, clamp01(u + noise * p[0] * 0.1f), clamp01(v + noise2 * p[0] * 0.1f));
            uint8_t* o = out + 4 * ((size_t)y * W + x);
            for (int q = 0; q < 4; q++) o[q] = unorm8(w.v[q]);
            continue;
          }
          if (p[4] == 2.0f) {
            const float m = smooth(-0.5f, 0.5f, noise);
            for (int q = 0; q < 3; q++) e[q] = c.v[q] * m;
          } else {
            fx_hsv(fractf(noise * 0.5f + 0.5f + time * 0.1f), 0.8f, noise2 * 0.5f + 0.5f, e);
          }
          break;
        }
        default: break;
      }
      const float base[3] = {c.v[0], c.v[1], c.v[2]};
      float r[3];
      orc_apply_blend(base, e, mode, amount, r);
      uint8_t* o = out + 4 * ((size_t)y * W + x);
      o[0] = unorm8(r[0]);
      o[1] = unorm8(r[1]);
      o[2] = unorm8(r[2]);
      o[3] = in[4 * ((size_t)y * W + x) + 3];
    }
}

/* The density chain over acc (W*H*4 floats, in place through scratch). */
void orc_density_chain(const ffm_effects* fx, float* acc, float* scratch, int W, int H,
                       float density_mean) {
  if (!fx) return;
  const size_t n = (size_t)W * H * 4;
  for (int i = 0; i < fx->n_density; i++) {
    if (!fx->density[i].enabled) continue;
    orc_density_effect(&fx->density[i], acc, scratch, W, H, density_mean, fx->pixel_scale);
    memcpy(acc, scratch, n * sizeof(float));
  }
}

/* The colour chain over rgba (W*H*4 bytes, in place through scratch). */
void orc_color_chain(const ffm_effects* fx, uint8_t* rgba, uint8_t* scratch, int W, int H) {
  if (!fx) return;
  const size_t n = (size_t)W * H * 4;
  for (int i = 0; i < fx->n_color; i++) {
    if (!fx->color[i].enabled) continue;
    orc_color_effect(&fx->color[i], rgba, scratch, W, H, fx->pixel_scale);
    memcpy(rgba, scratch, n);
  }
}
