// The render tail's effects (src/renderer/tail.rs): density effects on the
// f32 accumulator before the tone map, colour effects on the RGBA8 picture
// after it. One lane per output pixel; every input read is the effect
// language's texture read, bilinear between texel centres and clamped at
// the edges, by hand in f32 (uf_texture, src/lang/wgsl.rs:184-200), so an
// effect reads the same values on any device. The built-ins follow their
// Rhai definitions (assets/effects/<name>.rhai, whose pixels their WGSL
// predecessors in shaders/effects/ fix); a colour effect's colour goes
// through the blend library (shaders/effects/common/blend_modes.wgsl) over
// its input, the input's alpha kept (src/effects/custom.rs:49-60).
//
// These are per-pixel stencils of at most 21x21 taps read through L1/L2:
// bound by the tap loop (VALU, one exp per tap), not by HBM; each pass
// moves 16 B (density) or 4 B (colour) per pixel in and out.
#include <algorithm>
#include <cmath>

#include "ffm_fastmath.h"
#include "ffm_internal.h"

namespace ffm {

namespace {

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

__device__ __forceinline__ uint8_t to_unorm8(float v) {
  v = (v != v) ? 0.0f : clamp01(v);
  return (uint8_t)floorf(v * 255.0f + 0.5f);
}

// The image an effect reads: f32 RGBA or RGBA8 read as v / 255.
struct ImgF32 {
  const float4* p;
  int W, H;
  __device__ float4 at(int x, int y) const { return p[(size_t)y * W + x]; }
};
struct ImgU8 {
  const uchar4* p;
  int W, H;
  __device__ float4 at(int x, int y) const {
    const uchar4 c = p[(size_t)y * W + x];
    return make_float4((float)c.x / 255.0f, (float)c.y / 255.0f, (float)c.z / 255.0f,
                       (float)c.w / 255.0f);
  }
};

__device__ __forceinline__ float4 f4add(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 f4sub(float4 a, float4 b) {
  return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
}
__device__ __forceinline__ float4 f4mul(float4 a, float s) {
  return make_float4(a.x * s, a.y * s, a.z * s, a.w * s);
}
__device__ __forceinline__ float4 f4div(float4 a, float s) {
  return make_float4(a.x / s, a.y / s, a.z / s, a.w / s);
}

// uf_texture: p = uv * size - 0.5 (a non-finite or huge p reads as -1),
// clamped to [-1, size], then the four texels around it, clamped to the
// image, blended by the fractions.
template <class Img>
__device__ float4 tex(const Img& img, float u, float v) {
  float px = u * (float)img.W - 0.5f, py = v * (float)img.H - 0.5f;
  px = fabsf(px) <= 1e30f ? fminf(fmaxf(px, -1.0f), (float)img.W) : -1.0f;
  py = fabsf(py) <= 1e30f ? fminf(fmaxf(py, -1.0f), (float)img.H) : -1.0f;
  const float fx = floorf(px), fy = floorf(py);
  const float tx = px - fx, ty = py - fy;
  const int x0 = min(max((int)fx, 0), img.W - 1), y0 = min(max((int)fy, 0), img.H - 1);
  const int x1 = min(max((int)fx + 1, 0), img.W - 1), y1 = min(max((int)fy + 1, 0), img.H - 1);
  const float4 a = img.at(x0, y0), b = img.at(x1, y0), c = img.at(x0, y1), d = img.at(x1, y1);
  const float4 top = f4add(a, f4mul(f4sub(b, a), tx));
  const float4 bottom = f4add(c, f4mul(f4sub(d, c), tx));
  return f4add(top, f4mul(f4sub(bottom, top), ty));
}

// WGSL smoothstep as the Rhai effects write it.
__device__ __forceinline__ float smooth(float a, float b, float x) {
  const float t = clamp01((x - a) / (b - a));
  return t * t * (3.0f - 2.0f * t);
}

__device__ __forceinline__ float gaussian(float x, float sigma) {
  return expf(-(x * x) / (2.0f * sigma * sigma));
}

__device__ __forceinline__ float lum3(float r, float g, float b) {
  return r * 0.2126f + g * 0.7152f + b * 0.0722f;
}

__device__ __forceinline__ float length2(float x, float y) { return sqrtf(x * x + y * y); }

// sin and cos of the colour effects: the walker's fixed method
// (ffm_fastmath.h: Cody-Waite reduction and the Cephes polynomials, an f64
// reduction past |x| = 32768), not the platform's (whose last bits the
// reference leaves to the GPU), so the hash-based effects'
// fract(sin(n) * 43758.5453) is reproducible.
__device__ __forceinline__ void fx_sincos(float x, float* s, float* c) { ffm_sincosf(x, s, c); }
__device__ __forceinline__ float fx_sin(float x) {
  float s, c;
  fx_sincos(x, &s, &c);
  return s;
}
__device__ __forceinline__ float fx_cos(float x) {
  float s, c;
  fx_sincos(x, &s, &c);
  return c;
}

// ---------------------------------------------------------- density effects

// density_blur.rhai: a Gaussian of radius up to `radius` where the pixel is
// sparse (its hit count in multiples of the frame's mean below threshold),
// fading out by threshold + falloff; 9x9 taps `step` apart.
__device__ float4 density_blur(const ImgF32& img, const FxConst& k, float4 c, float u, float v) {
  const float max_radius = k.p[0] * k.pixel_scale;
  const float density = c.w / fmaxf(k.density_mean, 0.000001f);
  const float blur =
      density < k.p[1] ? 1.0f : 1.0f - smooth(k.p[1], k.p[1] + k.p[2], density);
  const float radius = max_radius * blur;
  if (radius < 0.5f) return c;
  const float px = 1.0f / (float)img.W, py = 1.0f / (float)img.H;
  const float sigma = radius / 2.0f;
  const float step = fmaxf(1.0f, radius / 4.0f);
  float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float weights = 0.0f;
  for (int yi = -4; yi < 5; yi++) {
    for (int xi = -4; xi < 5; xi++) {
      const float ox = (float)xi * step, oy = (float)yi * step;
      const float dist = length2(ox, oy);
      if (dist <= radius) {
        const float w = expf(-(dist * dist) / (2.0f * sigma * sigma));
        sum = f4add(sum, f4mul(tex(img, u + ox * px, v + oy * py), w));
        weights += w;
      }
    }
  }
  return weights <= 0.0f ? c : f4div(sum, weights);
}

// bilateral_blur.rhai: taps weighed by distance and by luminance
// difference from the centre; all four channels averaged.
__device__ float4 bilateral_blur(const ImgF32& img, const FxConst& k, float4 c, float u, float v) {
  const int radius = (int)fminf(fmaxf(k.p[0] * k.pixel_scale, 1.0f), 10.0f);
  const float sigma_spatial = fmaxf(0.1f, k.p[1] * k.pixel_scale);
  const float sigma_range = fmaxf(0.01f, k.p[2]);
  const float px = 1.0f / (float)img.W, py = 1.0f / (float)img.H;
  const float centre = lum3(c.x, c.y, c.z);
  float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float weights = 0.0f;
  for (int y = -radius; y < radius + 1; y++) {
    for (int x = -radius; x < radius + 1; x++) {
      const float fx = (float)x, fy = (float)y;
      const float4 s = tex(img, u + fx * px, v + fy * py);
      const float w = gaussian(length2(fx, fy), sigma_spatial) *
                      gaussian(fabsf(centre - lum3(s.x, s.y, s.z)), sigma_range);
      sum = f4add(sum, f4mul(s, w));
      weights += w;
    }
  }
  return f4div(sum, fmaxf(weights, 0.0001f));
}

// sharpen.rhai: unsharp mask, c + (c - blur(c)) * amount, colour clamped
// to [0, 1], the hit count kept.
__device__ float4 sharpen(const ImgF32& img, const FxConst& k, float4 c, float u, float v) {
  const float radius = k.p[1] * k.pixel_scale;
  const float px = 1.0f / (float)img.W, py = 1.0f / (float)img.H;
  const float sigma = radius / 2.0f;
  const int kk = (int)ceilf(radius);
  float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float weights = 0.0f;
  for (int y = -kk; y < kk + 1; y++) {
    for (int x = -kk; x < kk + 1; x++) {
      const float fx = (float)x, fy = (float)y;
      const float dist = length2(fx, fy);
      if (dist <= radius) {
        const float w = expf(-(dist * dist) / (2.0f * sigma * sigma));
        sum = f4add(sum, f4mul(tex(img, u + fx * px, v + fy * py), w));
        weights += w;
      }
    }
  }
  if (k.p[0] < 0.01f) return c;
  const float4 blurred = f4div(sum, weights);
  const float4 sharp = f4add(c, f4mul(f4sub(c, blurred), k.p[0]));
  return make_float4(clamp01(sharp.x), clamp01(sharp.y), clamp01(sharp.z), c.w);
}

__global__ __launch_bounds__(256) void density_effect_kernel(const float4* __restrict__ in,
                                                             float4* __restrict__ out, FxConst k) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= k.W || y >= k.H) return;
  const ImgF32 img{in, k.W, k.H};
  const float u = ((float)x + 0.5f) / (float)k.W, v = ((float)y + 0.5f) / (float)k.H;
  const float4 c = img.at(x, y);
  float4 o = c;
  switch (k.type) {
    case FFM_FX_DENSITY_BLUR: o = density_blur(img, k, c, u, v); break;
    case FFM_FX_BILATERAL_BLUR: o = bilateral_blur(img, k, c, u, v); break;
    case FFM_FX_SHARPEN: o = sharpen(img, k, c, u, v); break;
    default: break;
  }
  out[(size_t)y * k.W + x] = o;
}

// ------------------------------------------------------------ blend library

__device__ void rgb_to_hsl(const float c[3], float o[3]) {
  const float cmax = fmaxf(c[0], fmaxf(c[1], c[2])), cmin = fminf(c[0], fminf(c[1], c[2]));
  const float delta = cmax - cmin;
  const float l = (cmax + cmin) * 0.5f;
  if (delta < 0.00001f) {
    o[0] = 0.0f;
    o[1] = 0.0f;
    o[2] = l;
    return;
  }
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
  h /= 6.0f;
  o[0] = h;
  o[1] = s;
  o[2] = l;
}

__device__ float hue_to_rgb(float p, float q, float t) {
  if (t < 0.0f) t += 1.0f;
  if (t > 1.0f) t -= 1.0f;
  if (t < 1.0f / 6.0f) return p + (q - p) * 6.0f * t;
  if (t < 1.0f / 2.0f) return q;
  if (t < 2.0f / 3.0f) return p + (q - p) * (2.0f / 3.0f - t) * 6.0f;
  return p;
}

__device__ void hsl_to_rgb(const float hsl[3], float o[3]) {
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

// apply_blend(base, effect, mode, intensity): mode 0..12 as the library
// numbers them (Normal .. Luminosity; others Normal), then
// mix(base, blended, intensity).
__device__ void apply_blend(const float base[3], const float e[3], int mode, float amount,
                            float out[3]) {
  float b[3];
  if (mode >= 9 && mode <= 12) {
    float hb[3], he[3], hm[3];
    rgb_to_hsl(base, hb);
    rgb_to_hsl(e, he);
    hm[0] = (mode == 9 || mode == 11) ? he[0] : hb[0];
    hm[1] = (mode == 10 || mode == 11) ? he[1] : hb[1];
    hm[2] = mode == 12 ? he[2] : hb[2];
    hsl_to_rgb(hm, b);
  } else {
    for (int i = 0; i < 3; i++) {
      const float x = base[i], y = e[i];
      float r;
      switch (mode) {
        case 1: r = fminf(x + y, 1.0f); break;
        case 2: r = x * y; break;
        case 3: r = 1.0f - (1.0f - x) * (1.0f - y); break;
        case 4: r = x > 0.5f ? 1.0f - 2.0f * (1.0f - x) * (1.0f - y) : 2.0f * x * y; break;
        case 5: {
          const float d = x > 0.25f ? sqrtf(x) : ((16.0f * x - 12.0f) * x + 4.0f) * x;
          r = y > 0.5f ? x + (2.0f * y - 1.0f) * (d - x) : x - (1.0f - 2.0f * y) * x * (1.0f - x);
          break;
        }
        case 6: r = y > 0.5f ? 1.0f - 2.0f * (1.0f - x) * (1.0f - y) : 2.0f * x * y; break;
        case 7: r = y >= 1.0f ? 1.0f : fminf(1.0f, x / (1.0f - y + 0.00001f)); break;
        case 8: r = y <= 0.0f ? 0.0f : 1.0f - fminf(1.0f, (1.0f - x) / (y + 0.00001f)); break;
        default: r = y; break;
      }
      b[i] = r;
    }
  }
  for (int i = 0; i < 3; i++) out[i] = base[i] * (1.0f - amount) + b[i] * amount;
}

// ------------------------------------------------------------ colour effects

// vignette.rhai: the colour darkened by 1 - smoothstep over the
// aspect-corrected distance from the centre.
__device__ void vignette(const FxConst& k, float4 c, float u, float v, float o[3]) {
  const float aspect = (float)k.W / (float)k.H;
  const float dist = length2((u - 0.5f) * aspect, v - 0.5f);
  const float start = k.p[1] * 0.5f * fmaxf(1.0f, aspect);
  const float end = start + k.p[2] * 0.5f * fmaxf(1.0f, aspect);
  const float vig = 1.0f - smooth(start, end, dist);
  o[0] = c.x * vig;
  o[1] = c.y * vig;
  o[2] = c.z * vig;
}

// film_grain.rhai: hash2's first component at the pixel's coordinates
// offset by the seed: fract(sin(dot(p, (127.1, 311.7))) * 43758.5453).
__device__ void film_grain(const FxConst& k, float u, float v, float o[3]) {
  const float ax = u * (float)k.W + k.p[1] * 127.1f, ay = v * (float)k.H + k.p[1] * 311.7f;
  const float n = ax * 127.1f + ay * 311.7f;
  const float s = fx_sin(n) * 43758.5453f;
  const float g = s - floorf(s);
  o[0] = o[1] = o[2] = g;
}

// chromatic_aberration.rhai: red read `offset` one way and blue the other,
// radially (away from the centre, growing to the edges) or sideways.
__device__ void chromatic_aberration(const ImgU8& img, const FxConst& k, float4 c, float u,
                                     float v, float o[3]) {
  const float amount = k.p[0] * k.pixel_scale;
  const float px = 1.0f / (float)k.W, py = 1.0f / (float)k.H;
  float ox, oy;
  if (k.p[1] != 0.0f) {
    const float ex = u - 0.5f, ey = v - 0.5f;
    const float dist = length2(ex, ey) * 2.0f;
    const float vx = ex + 0.0001f, vy = ey + 0.0001f;
    const float vl = length2(vx, vy);
    const float nx = vx / vl, ny = vy / vl;
    ox = nx * amount * px * dist;
    oy = ny * amount * py * dist;
  } else {
    ox = amount * px;
    oy = 0.0f;
  }
  o[0] = tex(img, u + ox, v + oy).x;
  o[1] = c.y;
  o[2] = tex(img, u - ox, v - oy).z;
}

// hue_shift.rhai: HSV with the hue turned by offset degrees.
__device__ void hue_shift(const FxConst& k, float4 c, float o[3]) {
  const float r = c.x, g = c.y, b = c.z;
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
  float hx = h + k.p[0] / 360.0f;
  hx = hx - floorf(hx);
  const float hh = hx * 6.0f, vv = cmax;
  const float i = floorf(hh), f = hh - i;
  const float p = vv * (1.0f - s), q = vv * (1.0f - s * f), t = vv * (1.0f - s * (1.0f - f));
  if (i == 0.0f) {
    o[0] = vv; o[1] = t; o[2] = p;
  } else if (i == 1.0f) {
    o[0] = q; o[1] = vv; o[2] = p;
  } else if (i == 2.0f) {
    o[0] = p; o[1] = vv; o[2] = t;
  } else if (i == 3.0f) {
    o[0] = p; o[1] = q; o[2] = vv;
  } else if (i == 4.0f) {
    o[0] = t; o[1] = p; o[2] = vv;
  } else {
    o[0] = vv; o[1] = p; o[2] = q;
  }
}

// sobel_edges.rhai: Sobel gradient of the luminance, smoothstep'd between
// threshold / 2 and threshold, glowing in a tint between white and the
// pixel's colour.
__device__ void sobel_edges(const ImgU8& img, const FxConst& k, float4 c, float u, float v,
                            float o[3]) {
  const float tx = 1.0f / (float)k.W, ty = 1.0f / (float)k.H;
  auto L = [&](float du, float dv) {
    const float4 s = tex(img, u + du, v + dv);
    return lum3(s.x, s.y, s.z);
  };
  const float n0 = L(-tx, -ty), n1 = L(0.0f, -ty), n2 = L(tx, -ty), n3 = L(-tx, 0.0f);
  const float n5 = L(tx, 0.0f), n6 = L(-tx, ty), n7 = L(0.0f, ty), n8 = L(tx, ty);
  const float gx = n2 + 2.0f * n5 + n8 - (n0 + 2.0f * n3 + n6);
  const float gy = n0 + 2.0f * n1 + n2 - (n6 + 2.0f * n7 + n8);
  const float edge = smooth(k.p[1] * 0.5f, k.p[1], sqrtf(gx * gx + gy * gy));
  const float pc = k.p[3];
  const float gain = (1.0f + k.p[2] * 2.0f) * edge;
  o[0] = (1.0f * (1.0f - pc) + c.x * pc) * gain;
  o[1] = (1.0f * (1.0f - pc) + c.y * pc) * gain;
  o[2] = (1.0f * (1.0f - pc) + c.z * pc) * gain;
}

// The effect language's atan2: IEEE at (0, 0) (+-0 or +-pi by the signs),
// the platform's elsewhere (src/lang/wgsl.rs uf_atan2).
__device__ __forceinline__ float uf_atan2(float y, float x) {
  if (y == 0.0f && x == 0.0f) {
    const float mag = (__float_as_uint(x) & 0x80000000u) ? 3.14159265358979f : 0.0f;
    return (__float_as_uint(y) & 0x80000000u) ? -mag : mag;
  }
  return atan2f(y, x);
}

__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }

constexpr float FX_PI = 3.14159265359f;

// kaleidoscope.rhai: one slice of the image folded into mirrored segments.
__device__ void kaleidoscope(const ImgU8& img, const FxConst& k, float u, float v, float o[3]) {
  const float two_pi = 6.28318530718f;
  const float segments = fmaxf(2.0f, k.p[0]);
  const float rotation = k.p[1] * FX_PI / 180.0f;
  const float zoom = fmaxf(0.1f, k.p[2]);
  const bool square = k.p[5] > 0.0f;
  const float edge_offset = k.p[6] * k.pixel_scale;
  const float aspect = (float)k.W / (float)k.H;
  float x = (u - 0.5f) / zoom, y = (v - 0.5f) / zoom;
  if (square) x = x * aspect;
  const float radius = length2(x, y);
  float angle = uf_atan2(y, x);
  const float segment = two_pi / segments;
  angle = angle - segment * floorf(angle / segment);
  if (angle > segment * 0.5f) angle = segment - angle;
  angle = angle + rotation;
  float rx = fx_cos(angle) * radius;
  const float ry = fx_sin(angle) * radius;
  if (square) rx = rx / aspect;
  const float dx = -fx_cos(rotation), dy = -fx_sin(rotation);
  const float uu = rx + 0.5f + dx * edge_offset * (1.0f / (float)k.W);
  const float vv = ry + 0.5f + dy * edge_offset * (1.0f / (float)k.H);
  const float4 s = tex(img, fminf(fmaxf(uu, 0.001f), 0.999f), fminf(fmaxf(vv, 0.001f), 0.999f));
  o[0] = s.x;
  o[1] = s.y;
  o[2] = s.z;
}

// plasma.rhai: the demoscene plasma, a sum of sines travelling along
// `direction`.
__device__ void plasma(const FxConst& k, float u, float v, float o[3]) {
  const float scale = fmaxf(0.5f, k.p[1]);
  const float time = k.p[3] * k.p[2];
  const float direction = k.p[5] * FX_PI / 180.0f;
  const float x = u * scale, y = v * scale;
  const float dir_phase = x * fx_cos(direction) + y * fx_sin(direction);
  const float t = time + dir_phase;
  float w = fx_sin(x + t);
  w += fx_sin((y + t) * 0.5f);
  w += fx_sin((x + y + t) * 0.5f);
  const float cx = x + 0.5f * fx_sin(t * 0.25f), cy = y + 0.5f * fx_cos(t * 0.5f);
  w += fx_sin(sqrtf(cx * cx + cy * cy + 1.0f) + t);
  w = w * 0.25f;
  o[0] = fx_sin(w * FX_PI) * 0.5f + 0.5f;
  o[1] = fx_sin(w * FX_PI + 2.094f) * 0.5f + 0.5f;
  o[2] = fx_sin(w * FX_PI + 4.189f) * 0.5f + 0.5f;
}

// tunnel.rhai: the image wrapped on the inside of a tunnel; the centre,
// where the tunnel has no wall, keeps the input (amount 0).
__device__ float tunnel(const ImgU8& img, const FxConst& k, float u, float v, float o[3]) {
  const float aspect = (float)k.W / (float)k.H;
  const float x = (u - 0.5f) * aspect, y = v - 0.5f;
  const float angle = uf_atan2(y, x) / FX_PI;
  const float radius = length2(x, y);
  const float wall = fmaxf(radius, 0.001f);
  const float r = wall + k.p[2] * fx_sin(angle * FX_PI * 4.0f + k.p[3]) * 0.1f;
  const float au = fractf(angle * 0.5f + 0.5f + k.p[3] * k.p[1] * 0.1f);
  const float av = fractf(1.0f / r + k.p[3] * k.p[0]);
  const float4 s = tex(img, au, av);
  const float fade = smooth(0.0f, 0.5f, radius);
  o[0] = s.x * fade;
  o[1] = s.y * fade;
  o[2] = s.z * fade;
  return radius < 0.001f ? 0.0f : k.p[4];
}

// hsv_to_rgb as julia.rhai and worley_noise.rhai write it.
__device__ void fx_hsv(float h, float s, float v, float o[3]) {
  const float c = v * s, h6 = h * 6.0f;
  const float x = c * (1.0f - fabsf(h6 - 2.0f * truncf(h6 / 2.0f) - 1.0f));
  const float m = v - c;
  float r, g, b;
  if (h6 < 1.0f) {
    r = c; g = x; b = 0.0f;
  } else if (h6 < 2.0f) {
    r = x; g = c; b = 0.0f;
  } else if (h6 < 3.0f) {
    r = 0.0f; g = c; b = x;
  } else if (h6 < 4.0f) {
    r = 0.0f; g = x; b = c;
  } else if (h6 < 5.0f) {
    r = x; g = 0.0f; b = c;
  } else {
    r = c; g = 0.0f; b = x;
  }
  o[0] = r + m;
  o[1] = g + m;
  o[2] = b + m;
}

// julia.rhai: a Mandelbrot (mode 0) or Julia (mode 1) set, escape-time
// coloured (color_mode 0), smooth (1) or masking the input (2).
__device__ void julia(const FxConst& k, float4 c, float u, float v, float o[3]) {
  const float max_iter = fmaxf(10.0f, k.p[3]);
  const float zoom = fmaxf(0.1f, k.p[4]);
  const float aspect = (float)k.W / (float)k.H;
  const float px = (u - 0.5f) * 2.0f * aspect / zoom + k.p[5];
  const float py = (v - 0.5f) * 2.0f / zoom + k.p[6];
  const bool is_julia = k.p[0] == 1.0f;
  float zr = is_julia ? px : 0.0f, zi = is_julia ? py : 0.0f;
  const float kr = is_julia ? k.p[1] : px, ki = is_julia ? k.p[2] : py;
  int i = 0;
  float x2 = zr * zr, y2 = zi * zi;
  while ((float)i < max_iter && !(x2 + y2 > 4.0f)) {
    const float nr = x2 - y2 + kr;
    const float ni = 2.0f * zr * zi + ki;
    zr = nr;
    zi = ni;
    x2 = zr * zr;
    y2 = zi * zi;
    i += 1;
  }
  const float iter = (float)i;
  const float zmag = zr * zr + zi * zi;
  if (iter >= max_iter) {
    o[0] = o[1] = o[2] = 0.0f;
  } else if (k.p[7] == 1.0f) {
    const float ln2 = logf(2.0f);
    const float log_zn = logf(zmag) / 2.0f;
    const float nu = logf(log_zn / ln2) / ln2;
    const float t = (iter + 1.0f - nu) / max_iter;
    fx_hsv(fractf(t * 5.0f), 0.8f, 1.0f, o);
  } else if (k.p[7] == 2.0f) {
    const float f = iter / max_iter;
    o[0] = c.x * f;
    o[1] = c.y * f;
    o[2] = c.z * f;
  } else {
    const float t = iter / max_iter;
    fx_hsv(fractf(t * 3.0f), 0.7f, t, o);
  }
}

// The effect language's hash2: fract(sin(p . (127.1, 311.7), p . (269.5,
// 183.3)) * 43758.5453) (src/lang/wgsl.rs uf_hash2).
__device__ __forceinline__ void fx_hash2(float x, float y, float& hx, float& hy) {
  const float n0 = x * 127.1f + y * 311.7f, n1 = x * 269.5f + y * 183.3f;
  hx = fractf(fx_sin(n0) * 43758.5453f);
  hy = fractf(fx_sin(n1) * 43758.5453f);
}

// worley_noise.rhai: distances to the nearest and second-nearest of the
// moving points in the 3x3 cells around.
__device__ void worley(const FxConst& k, float4 c, float u, float v, float o[3]) {
  const float scale = fmaxf(1.0f, k.p[1]);
  const float time = k.p[2];
  const float direction = k.p[6] * FX_PI / 180.0f;
  const float ax = (u + fx_cos(direction) * time * 0.1f) * scale;
  const float ay = (v + fx_sin(direction) * time * 0.1f) * scale;
  const float cx = floorf(ax), cy = floorf(ay);
  const float fx = fractf(ax), fy = fractf(ay);
  float d1 = 8.0f, d2 = 8.0f;
  for (int j = -1; j < 2; j++) {
    for (int i = -1; i < 2; i++) {
      const float nx = (float)i, ny = (float)j;
      float hx, hy;
      fx_hash2(cx + nx, cy + ny, hx, hy);
      const float qx = 0.5f + 0.5f * fx_sin(time * 0.5f + 6.2831f * hx);
      const float qy = 0.5f + 0.5f * fx_sin(time * 0.5f + 6.2831f * hy);
      const float dist = length2(nx + qx - fx, ny + qy - fy);
      if (dist < d1) {
        d2 = d1;
        d1 = dist;
      } else if (dist < d2) {
        d2 = dist;
      }
    }
  }
  const float mode = k.p[3];
  const float pattern = mode == 1.0f   ? smooth(0.0f, 0.15f, d2 - d1)
                        : mode == 2.0f ? 1.0f - smooth(0.0f, 0.5f, d1)
                        : mode == 3.0f ? clamp01(d1 * d2 * 2.0f)
                                       : d1;
  if (k.p[4] == 1.0f) {
    fx_hsv(fractf(d1 + time * 0.1f), 0.8f, pattern, o);
  } else if (k.p[4] == 2.0f) {
    o[0] = c.x * pattern;
    o[1] = c.y * pattern;
    o[2] = c.z * pattern;
  } else {
    o[0] = o[1] = o[2] = pattern;
  }
}

// Gradient noise of domain_warp.rhai: per lattice corner a hash2 gradient,
// blended by the smoothstep of the fraction.
__device__ __forceinline__ float dw_grad(float lx, float ly, float ox, float oy) {
  float gx, gy;
  fx_hash2(lx, ly, gx, gy);
  return (gx * 2.0f - 1.0f) * ox + (gy * 2.0f - 1.0f) * oy;
}

__device__ float dw_noise(float x, float y) {
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

// Simplex noise of simplex_noise.rhai: three corners of the skewed
// triangle, each (0.5 - r^2)^4 times its hash2 gradient's dot product.
__device__ __forceinline__ float sx_corner(float lx, float ly, float ox, float oy) {
  const float t = 0.5f - (ox * ox + oy * oy);
  if (!(t >= 0.0f)) return 0.0f;
  float gx, gy;
  fx_hash2(lx, ly, gx, gy);
  const float t2 = t * t;
  return t2 * t2 * ((gx * 2.0f - 1.0f) * ox + (gy * 2.0f - 1.0f) * oy);
}

__device__ float sx_noise(float x, float y) {
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

// fbm of both noise effects: `octaves` octaves, each rotated and scaled by
// 2.02, amplitudes halving from 0.5.
template <bool SIMPLEX>
__device__ float fx_fbm(float x, float y, float octaves) {
  float value = 0.0f, amplitude = 0.5f, px = x, py = y;
  for (int o = 0; (float)o < octaves; o++) {
    value += amplitude * (SIMPLEX ? sx_noise(px, py) : dw_noise(px, py));
    const float nx = (0.8f * px + 0.6f * py) * 2.02f;
    const float ny = (-0.6f * px + 0.8f * py) * 2.02f;
    px = nx;
    py = ny;
    amplitude *= 0.5f;
  }
  return value;
}

// domain_warp.rhai: the image read through noise warped by noise.
__device__ void domain_warp(const ImgU8& img, const FxConst& k, float u, float v, float o[3]) {
  const float scale = fmaxf(0.5f, k.p[1]);
  const float octaves = fminf(fmaxf(k.p[2], 1.0f), 6.0f);
  const float direction = k.p[5] * FX_PI / 180.0f;
  const float tx = fx_cos(direction) * k.p[3] * 0.1f, ty = fx_sin(direction) * k.p[3] * 0.1f;
  const float x = u * scale, y = v * scale;
  const float qx = fx_fbm<false>(x + 0.0f + tx, y + 0.0f + ty, octaves);
  const float qy = fx_fbm<false>(x + 5.2f + tx * 1.2f, y + 1.3f + ty * 1.2f, octaves);
  const float rx = fx_fbm<false>(x + 4.0f * qx + 1.7f, y + 4.0f * qy + 9.2f, octaves);
  const float ry = fx_fbm<false>(x + 4.0f * qx + 8.3f, y + 4.0f * qy + 2.8f, octaves);
  const float4 s = tex(img, clamp01(u + rx * k.p[0]), clamp01(v + ry * k.p[0]));
  o[0] = s.x;
  o[1] = s.y;
  o[2] = s.z;
}

// simplex_noise.rhai: fbm simplex noise as colour (mode 0), a distortion of
// the image written unblended (mode 1, its alpha too) or a mask (mode 2).
// Returns the alpha of an unblended result, or -1 for a blended one.
__device__ float simplex(const ImgU8& img, const FxConst& k, float4 c, float u, float v,
                         float o[3]) {
  const float scale = fmaxf(0.5f, k.p[1]);
  const float octaves = fminf(fmaxf(k.p[2], 1.0f), 6.0f);
  const float time = k.p[3];
  const float direction = k.p[6] * FX_PI / 180.0f;
  const float nx = u * scale + fx_cos(direction) * time * 0.1f;
  const float ny = v * scale + fx_sin(direction) * time * 0.1f;
  const float noise = fx_fbm<true>(nx, ny, octaves);
  const float noise2 = fx_fbm<true>(nx + 5.2f, ny + 1.3f, octaves);
  if (k.p[4] == 1.0f) {
    const float4 w = tex(img, clamp01(u + noise * k.p[0] * 0.1f), clamp01(v + noise2 * k.p[0] * 0.1f));
    o[0] = w.x;
    o[1] = w.y;
    o[2] = w.z;
    return w.w;
  }
  if (k.p[4] == 2.0f) {
    const float m = smooth(-0.5f, 0.5f, noise);
    o[0] = c.x * m;
    o[1] = c.y * m;
    o[2] = c.z * m;
  } else {
    fx_hsv(fractf(noise * 0.5f + 0.5f + time * 0.1f), 0.8f, noise2 * 0.5f + 0.5f, o);
  }
  return -1.0f;
}

__global__ __launch_bounds__(256) void color_effect_kernel(const uchar4* __restrict__ in,
                                                           uchar4* __restrict__ out, FxConst k) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= k.W || y >= k.H) return;
  const ImgU8 img{in, k.W, k.H};
  const float u = ((float)x + 0.5f) / (float)k.W, v = ((float)y + 0.5f) / (float)k.H;
  const float4 c = img.at(x, y);
  float e[3];
  int mode = 0;
  float amount = 1.0f;
  float raw_alpha = -1.0f;  // >= 0: the effect's colour is written as is, unblended
  switch (k.type) {
    case FFM_FX_VIGNETTE:
      vignette(k, c, u, v, e);
      amount = k.p[0];
      mode = (int)k.p[3];
      break;
    case FFM_FX_FILM_GRAIN:
      film_grain(k, u, v, e);
      amount = k.p[0];
      mode = (int)k.p[2];
      break;
    case FFM_FX_CHROMATIC_ABERRATION:
      chromatic_aberration(img, k, c, u, v, e);
      amount = k.p[2];
      mode = (int)k.p[3];
      break;
    case FFM_FX_HUE_SHIFT:
      hue_shift(k, c, e);
      amount = k.p[1];
      mode = (int)k.p[2];
      break;
    case FFM_FX_SOBEL_EDGES:
      sobel_edges(img, k, c, u, v, e);
      amount = k.p[0];
      mode = (int)k.p[4];
      break;
    case FFM_FX_KALEIDOSCOPE:
      kaleidoscope(img, k, u, v, e);
      amount = k.p[3];
      mode = (int)k.p[4];
      break;
    case FFM_FX_PLASMA:
      plasma(k, u, v, e);
      amount = k.p[0];
      mode = (int)k.p[4];
      break;
    case FFM_FX_TUNNEL:
      amount = tunnel(img, k, u, v, e);
      mode = (int)k.p[5];
      break;
    case FFM_FX_JULIA:
      julia(k, c, u, v, e);
      amount = k.p[8];
      mode = (int)k.p[9];
      break;
    case FFM_FX_WORLEY_NOISE:
      worley(k, c, u, v, e);
      amount = k.p[0];
      mode = (int)k.p[5];
      break;
    case FFM_FX_SIMPLEX_NOISE:
      raw_alpha = simplex(img, k, c, u, v, e);
      amount = k.p[0];
      mode = (int)k.p[5];
      break;
    case FFM_FX_DOMAIN_WARP:
      domain_warp(img, k, u, v, e);
      amount = k.p[0] * 2.0f;
      mode = (int)k.p[4];
      break;
    default:
      e[0] = c.x;
      e[1] = c.y;
      e[2] = c.z;
      break;
  }
  if (raw_alpha >= 0.0f) {  // blend -1 (src/effects/custom.rs:49-60): the colour itself
    out[(size_t)y * k.W + x] =
        make_uchar4(to_unorm8(e[0]), to_unorm8(e[1]), to_unorm8(e[2]), to_unorm8(raw_alpha));
    return;
  }
  const float base[3] = {c.x, c.y, c.z};
  float r[3];
  apply_blend(base, e, mode, amount, r);
  out[(size_t)y * k.W + x] = make_uchar4(to_unorm8(r[0]), to_unorm8(r[1]), to_unorm8(r[2]),
                                         in[(size_t)y * k.W + x].w);
}

// ------------------------------------------------------------------ bloom

// bright(): the part of a colour above the threshold, with a soft knee.
__device__ __forceinline__ float4 bloom_bright(float4 c, float threshold, float softness) {
  const float l = fmaxf(c.x, fmaxf(c.y, c.z));
  const float knee = fmaxf(threshold * softness, 0.00001f);
  const float soft = fminf(fmaxf(l - threshold + knee, 0.0f), 2.0f * knee);
  const float w = fmaxf(soft * soft / (4.0f * knee), l - threshold) / fmaxf(l, 0.00001f);
  return f4mul(c, w);
}

// One level of the down pyramid: level 0 averages the bright parts of
// four input reads half a level-0 texel away; higher levels weigh the
// level below's read at the centre 4 and at four diagonals 1, over 8.
__global__ __launch_bounds__(256) void bloom_down_kernel(BloomArgs b, int r) {
  const int w = b.lw[r], h = b.lh[r];
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= w || y >= h) return;
  const float u = ((float)x + 0.5f) / (float)w, v = ((float)y + 0.5f) / (float)h;
  const float dx = 0.5f / (float)w, dy = 0.5f / (float)h;
  float4 o;
  if (r == 0) {
    const ImgU8 in{reinterpret_cast<const uchar4*>(b.in), b.W, b.H};
    o = f4add(f4add(f4add(bloom_bright(tex(in, u - dx, v - dy), b.threshold, b.softness),
                          bloom_bright(tex(in, u + dx, v - dy), b.threshold, b.softness)),
                    bloom_bright(tex(in, u - dx, v + dy), b.threshold, b.softness)),
              bloom_bright(tex(in, u + dx, v + dy), b.threshold, b.softness));
    o = f4mul(o, 0.25f);
  } else {
    const ImgF32 src{reinterpret_cast<const float4*>(b.down[r - 1]), b.lw[r - 1], b.lh[r - 1]};
    o = f4add(f4add(f4add(f4add(f4mul(tex(src, u, v), 4.0f), tex(src, u - dx, v - dy)),
                          tex(src, u + dx, v - dy)),
                    tex(src, u - dx, v + dy)),
              tex(src, u + dx, v + dy));
    o = f4mul(o, 0.125f);
  }
  reinterpret_cast<float4*>(b.down[r])[(size_t)y * w + x] = o;
}

// One level of the up pyramid, from the top: the down level plus, below
// the top, a 3x3 tent of the level above times spread / 16.
__global__ __launch_bounds__(256) void bloom_up_kernel(BloomArgs b, int l) {
  const int w = b.lw[l], h = b.lh[l];
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= w || y >= h) return;
  const float u = ((float)x + 0.5f) / (float)w, v = ((float)y + 0.5f) / (float)h;
  const ImgF32 dn{reinterpret_cast<const float4*>(b.down[l]), w, h};
  const float4 here = tex(dn, u, v);
  float4 o = here;
  if (l != b.levels - 1) {
    const ImgF32 up{reinterpret_cast<const float4*>(b.up[l + 1]), b.lw[l + 1], b.lh[l + 1]};
    const float dx = 1.0f / (float)w, dy = 1.0f / (float)h;
    const float4 side =
        f4add(f4add(f4add(tex(up, u - dx, v), tex(up, u + dx, v)), tex(up, u, v - dy)),
              tex(up, u, v + dy));
    float4 tent = f4add(f4mul(tex(up, u, v), 4.0f), f4mul(side, 2.0f));
    tent = f4add(tent, tex(up, u - dx, v - dy));
    tent = f4add(tent, tex(up, u + dx, v - dy));
    tent = f4add(tent, tex(up, u - dx, v + dy));
    tent = f4add(tent, tex(up, u + dx, v + dy));
    o = f4add(here, f4mul(tent, 0.0625f * b.spread));
  }
  reinterpret_cast<float4*>(b.up[l])[(size_t)y * w + x] = o;
}

// The composite: the picture plus the glow (up level 0) times the
// intensity, clamped; the picture's alpha kept; written unblended.
__global__ __launch_bounds__(256) void bloom_composite_kernel(BloomArgs b) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= b.W || y >= b.H) return;
  const float u = ((float)x + 0.5f) / (float)b.W, v = ((float)y + 0.5f) / (float)b.H;
  const ImgU8 in{reinterpret_cast<const uchar4*>(b.in), b.W, b.H};
  const ImgF32 up{reinterpret_cast<const float4*>(b.up[0]), b.lw[0], b.lh[0]};
  const float4 c = in.at(x, y), glow = tex(up, u, v);
  const uchar4 a = reinterpret_cast<const uchar4*>(b.in)[(size_t)y * b.W + x];
  reinterpret_cast<uchar4*>(b.out)[(size_t)y * b.W + x] =
      make_uchar4(to_unorm8(clamp01(c.x + glow.x * 1.0f * b.intensity)),
                  to_unorm8(clamp01(c.y + glow.y * 1.0f * b.intensity)),
                  to_unorm8(clamp01(c.z + glow.z * 1.0f * b.intensity)), a.w);
}

// --------------------------------------------------------- distance field

// A pixel's coordinates are its centre (x + 0.5, y + 0.5), as the fragment
// position; input_at / buffer_at truncate them back to the texel.
__device__ __forceinline__ float4 texel_u8(const uint8_t* img, int W, int H, int x, int y) {
  x = min(max(x, 0), W - 1);
  y = min(max(y, 0), H - 1);
  const uchar4 c = reinterpret_cast<const uchar4*>(img)[(size_t)y * W + x];
  return make_float4((float)c.x / 255.0f, (float)c.y / 255.0f, (float)c.z / 255.0f,
                     (float)c.w / 255.0f);
}

// seed(): a pixel brighter than the threshold seeds itself, others -1.
__global__ __launch_bounds__(256) void df_seed_kernel(DistArgs d) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= d.W || y >= d.H) return;
  const float4 here = texel_u8(d.in, d.W, d.H, x, y);
  const bool seed = lum3(here.x, here.y, here.z) > d.threshold;
  reinterpret_cast<float2*>(d.a)[(size_t)y * d.W + x] =
      seed ? make_float2((float)x + 0.5f, (float)y + 0.5f) : make_float2(-1.0f, -1.0f);
}

// flood(): the nearest seed among the 3x3 neighbours `step` apart.
__global__ __launch_bounds__(256) void df_flood_kernel(DistArgs d, const float2* __restrict__ src,
                                                       float2* __restrict__ dst, int step) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= d.W || y >= d.H) return;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  float2 best = make_float2(-1.0f, -1.0f);
  float best_d = 1.0e30f;
  for (int dy = -1; dy < 2; dy++) {
    for (int dx = -1; dx < 2; dx++) {
      const int qx = x + dx * step, qy = y + dy * step;
      if (qx >= 0 && qy >= 0 && qx < d.W && qy < d.H) {
        const float2 s = src[(size_t)qy * d.W + qx];
        if (s.x >= 0.0f) {
          const float dist = length2(s.x - px, s.y - py);
          if (dist < best_d) {
            best_d = dist;
            best = s;
          }
        }
      }
    }
  }
  dst[(size_t)y * d.W + x] = best;
}

// effect(): glow (mode 0), contour bands (1) or nearest-seed fill (2)
// over the picture, by the distance to the nearest seed; unreached pixels
// and the alpha as they were; written unblended.
__global__ __launch_bounds__(256) void df_effect_kernel(DistArgs d,
                                                        const float2* __restrict__ coords) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= d.W || y >= d.H) return;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  const float4 base = texel_u8(d.in, d.W, d.H, x, y);
  const float2 s = coords[(size_t)y * d.W + x];
  float r = base.x, g = base.y, bl = base.z;
  if (s.x >= 0.0f) {
    const float range_px = fmaxf(d.spread, 0.01f) * (float)max(d.W, d.H);
    const float dist = length2(s.x - px, s.y - py);
    const float t = clamp01(dist / range_px);
    const float4 sc = texel_u8(d.in, d.W, d.H, (int)s.x, (int)s.y);
    const float off = dist > 0.5f ? 1.0f : 0.0f;
    if (d.mode == 0.0f) {
      const float k = expf(-4.0f * t);
      r = base.x + sc.x * k * off * d.intensity;
      g = base.y + sc.y * k * off * d.intensity;
      bl = base.z + sc.z * k * off * d.intensity;
    } else if (d.mode == 1.0f) {
      const float phase = fractf(t * fmaxf(d.band_count, 1.0f));
      const float ring = smooth(0.0f, 0.15f, phase) * (1.0f - smooth(0.35f, 0.5f, phase));
      r = base.x + sc.x * ring * (1.0f - t) * d.intensity * off;
      g = base.y + sc.y * ring * (1.0f - t) * d.intensity * off;
      bl = base.z + sc.z * ring * (1.0f - t) * d.intensity * off;
    } else if (dist > 0.5f) {
      const float m = clamp01(d.intensity) * (1.0f - t);
      r = base.x * (1.0f - m) + sc.x * m;
      g = base.y * (1.0f - m) + sc.y * m;
      bl = base.z * (1.0f - m) + sc.z * m;
    }
  }
  const uchar4 a = reinterpret_cast<const uchar4*>(d.in)[(size_t)y * d.W + x];
  reinterpret_cast<uchar4*>(d.out)[(size_t)y * d.W + x] =
      make_uchar4(to_unorm8(r), to_unorm8(g), to_unorm8(bl), a.w);
}

}  // namespace

hipError_t launch_density_effect(const FxConst& k, const float* in, float* out, hipStream_t s) {
  dim3 grid((k.W + 15) / 16, (k.H + 15) / 16);
  hipLaunchKernelGGL(density_effect_kernel, grid, dim3(256), 0, s,
                     reinterpret_cast<const float4*>(in), reinterpret_cast<float4*>(out), k);
  return hipGetLastError();
}

// Buffer sizes of bloom's pyramids (BufferDecl::size / levels_at,
// src/effects/passes.rs:21-33: scale 0.5, rounded half away from zero,
// at most 7 levels and no more than the bit length of the longer side).
int bloom_levels(int W, int H, int32_t* lw, int32_t* lh) {
  const int bw = std::max((int)std::round((float)W * 0.5f), 1);
  const int bh = std::max((int)std::round((float)H * 0.5f), 1);
  int bits = 0;
  for (int m = std::max(bw, bh); m; m >>= 1) bits++;
  const int n = std::max(std::min(BLOOM_LEVELS, bits), 1);
  for (int l = 0; l < n; l++) {
    lw[l] = std::max(bw >> l, 1);
    lh[l] = std::max(bh >> l, 1);
  }
  return n;
}

hipError_t launch_bloom(const BloomArgs& b, hipStream_t s) {
  for (int r = 0; r < b.levels; r++) {
    hipLaunchKernelGGL(bloom_down_kernel, dim3((b.lw[r] + 15) / 16, (b.lh[r] + 15) / 16),
                       dim3(256), 0, s, b, r);
    if (hipError_t e = hipGetLastError()) return e;
  }
  for (int l = b.levels - 1; l >= 0; l--) {
    hipLaunchKernelGGL(bloom_up_kernel, dim3((b.lw[l] + 15) / 16, (b.lh[l] + 15) / 16),
                       dim3(256), 0, s, b, l);
    if (hipError_t e = hipGetLastError()) return e;
  }
  hipLaunchKernelGGL(bloom_composite_kernel, dim3((b.W + 15) / 16, (b.H + 15) / 16), dim3(256), 0,
                     s, b);
  return hipGetLastError();
}

// The jump flood's passes: log2_passes (src/effects/passes.rs:110-113),
// step 2^(n-1-r) for pass r.
hipError_t launch_distance_field(const DistArgs& d, hipStream_t s) {
  const dim3 grid((d.W + 15) / 16, (d.H + 15) / 16);
  hipLaunchKernelGGL(df_seed_kernel, grid, dim3(256), 0, s, d);
  if (hipError_t e = hipGetLastError()) return e;
  const uint32_t m = (uint32_t)std::max(std::max(d.W, d.H), 2) - 1u;
  const int n = 32 - __builtin_clz(m);
  float2* cur = reinterpret_cast<float2*>(d.a);
  float2* nxt = reinterpret_cast<float2*>(d.b);
  for (int r = 0; r < n; r++) {
    hipLaunchKernelGGL(df_flood_kernel, grid, dim3(256), 0, s, d, cur, nxt, 1 << (n - 1 - r));
    if (hipError_t e = hipGetLastError()) return e;
    std::swap(cur, nxt);
  }
  hipLaunchKernelGGL(df_effect_kernel, grid, dim3(256), 0, s, d, cur);
  return hipGetLastError();
}

hipError_t launch_color_effect(const FxConst& k, const uint8_t* in, uint8_t* out, hipStream_t s) {
  dim3 grid((k.W + 15) / 16, (k.H + 15) / 16);
  hipLaunchKernelGGL(color_effect_kernel, grid, dim3(256), 0, s,
                     reinterpret_cast<const uchar4*>(in), reinterpret_cast<uchar4*>(out), k);
  return hipGetLastError();
}

}  // namespace ffm
