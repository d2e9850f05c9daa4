// stencils.hip -- the per-pixel passes after the splat, for gfx950:
// accumulate, bilateral histogram blur, flam3-style density estimation and
// the log-density tonemap. All are HBM-bound streaming stencils (no MFMA:
// nothing here is a contraction); each thread owns whole pixels and moves
// them as 16-byte vectors.
#include <hip/hip_runtime.h>

#include "ffm_internal.h"

namespace ffm {

static int grid_for(size_t n, int block, int cap = 16384) {
  size_t g = (n + (size_t)block - 1) / (size_t)block;
  if (g > (size_t)cap) g = (size_t)cap;
  return g < 1 ? 1 : (int)g;
}

// ------------------------------------------------------------ accumulate

// One pixel of accumulate.wgsl: h = this batch's (r, g, b, density) sums
// as f32, hd the density word as an integer (for the exact count).
__device__ __forceinline__ float4 accum_one(float4 prev, float h0, float h1, float h2, float h3,
                                            uint32_t hd, uint2* count, int mode, float blend) {
  const float SCALE = 100.0f;
  if (mode == 2) prev = make_float4(0.f, 0.f, 0.f, 0.f);
  float new_density = h3 / SCALE;
  float total = prev.w + new_density;
  if (count) {
    uint2 c = *count;
    if (prev.w == 0.0f) {
      c = make_uint2(0u, 0u);
    } else if (c.x == 0u && c.y == 0u) {
      float seed = prev.w * SCALE;
      float hi = floorf(seed / 4294967296.0f);
      c = make_uint2((uint32_t)(seed - hi * 4294967296.0f), (uint32_t)hi);
    }
    uint32_t lo = c.x + hd;
    c = make_uint2(lo, c.y + (lo < c.x ? 1u : 0u));
    *count = c;
    total = ((float)c.y * 4294967296.0f + (float)c.x) / SCALE;
    prev.w = total;
  }
  if (new_density <= 0.0f) return prev;
  float dn = fmaxf(h3, 1.0f);
  float n0 = fminf(fmaxf(h0 / dn, 0.0f), 1.0f);
  float n1 = fminf(fmaxf(h1 / dn, 0.0f), 1.0f);
  float n2 = fminf(fmaxf(h2 / dn, 0.0f), 1.0f);
  float cum_alpha = new_density / total;
  float alpha = mode == 1 ? blend : cum_alpha;
  float4 o;
  o.x = prev.x + (n0 - prev.x) * alpha;
  o.y = prev.y + (n1 - prev.y) * alpha;
  o.z = prev.z + (n2 - prev.z) * alpha;
  o.w = total;
  return o;
}

// shaders/accumulate.wgsl:115-230 (no depth buffer). One pixel per thread:
// reads the batch histogram (u32x4 or f32x4) and the previous accumulator,
// writes the new accumulator (may alias prev) and, when `count` is set, the
// exact 64-bit density count.
template <bool F32H>
__global__ __launch_bounds__(256) void accumulate_kernel(const void* __restrict__ hist_v,
                                                         const float4* prev_buf, float4* out_buf,
                                                         uint2* count, size_t npix, int mode,
                                                         float blend) {
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (size_t)gridDim.x * blockDim.x) {
    const float4 prev = mode == 2 ? make_float4(0.f, 0.f, 0.f, 0.f) : prev_buf[p];
    float4 o;
    if (F32H) {
      const float4 h = reinterpret_cast<const float4*>(hist_v)[p];
      o = accum_one(prev, h.x, h.y, h.z, h.w, (uint32_t)h.w, count ? count + p : nullptr, mode, blend);
    } else {
      const uint4 h = reinterpret_cast<const uint4*>(hist_v)[p];
      o = accum_one(prev, (float)h.x, (float)h.y, (float)h.z, (float)h.w, h.w,
                    count ? count + p : nullptr, mode, blend);
    }
    out_buf[p] = o;
  }
}

hipError_t launch_accumulate(const void* hist, int hist_is_f32, const float* prev, float* out,
                             uint32_t* count, int W, int H, int mode, float blend, hipStream_t s) {
  size_t npix = (size_t)W * H;
  int grid = grid_for(npix, 256);
  if (hist_is_f32)
    hipLaunchKernelGGL(accumulate_kernel<true>, dim3(grid), dim3(256), 0, s, hist,
                       reinterpret_cast<const float4*>(prev), reinterpret_cast<float4*>(out),
                       reinterpret_cast<uint2*>(count), npix, mode, blend);
  else
    hipLaunchKernelGGL(accumulate_kernel<false>, dim3(grid), dim3(256), 0, s, hist,
                       reinterpret_cast<const float4*>(prev), reinterpret_cast<float4*>(out),
                       reinterpret_cast<uint2*>(count), npix, mode, blend);
  return hipGetLastError();
}

// ------------------------------------------------------------------ blur

// shaders/histogram_blur.wgsl:39-119, one pass. `center` is the original
// primary buffer, the joint-bilateral reference in both passes. `out` may
// be `center`: a thread reads its own pixel's centre before it writes that
// pixel, and no other thread reads it.
__global__ __launch_bounds__(256) void blur_kernel(const uint4* __restrict__ in, uint4* out,
                                                   const uint4* center, int W, int H,
                                                   float radius_pixels, float density_sigma,
                                                   int direction) {
  const int KMAX = 8;
  const size_t npix = (size_t)W * H;
  float sigma = fmaxf(radius_pixels, 0.0001f);
  int kh = (int)ceilf(3.0f * sigma);
  if (kh > KMAX) kh = KMAX;
  float inv_2s2 = 1.0f / (2.0f * sigma * sigma);
  float sigma_d = fmaxf(density_sigma, 1.0f);
  float inv_2sd2 = 1.0f / (2.0f * sigma_d * sigma_d);
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (size_t)gridDim.x * blockDim.x) {
    int x = (int)(p % (size_t)W), y = (int)(p / (size_t)W);
    float cd = (float)center[p].w;
    float sr = 0.f, sg = 0.f, sb = 0.f, sd = 0.f, wt = 0.f;
    for (int k = -kh; k <= kh; k++) {
      int nx = x, ny = y;
      if (direction == 0)
        nx = min(max(x + k, 0), W - 1);
      else
        ny = min(max(y + k, 0), H - 1);
      uint4 t = in[(size_t)ny * W + nx];
      float tr = (float)t.x, tg = (float)t.y, tb = (float)t.z, td = (float)t.w;
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
    out[p] = make_uint4((uint32_t)(sr * inv_w), (uint32_t)(sg * inv_w), (uint32_t)(sb * inv_w),
                        (uint32_t)(sd * inv_w));
  }
}

// H: hist -> scratch (centre from hist). V: scratch -> hist in place (the
// centre is each pixel's own word of hist, read before it is written).
hipError_t launch_blur(uint32_t* hist, uint32_t* scratch, int W, int H, float sigma,
                       float density_sigma, hipStream_t s) {
  size_t npix = (size_t)W * H;
  int grid = grid_for(npix, 256);
  uint4* h = reinterpret_cast<uint4*>(hist);
  uint4* sc = reinterpret_cast<uint4*>(scratch);
  hipLaunchKernelGGL(blur_kernel, dim3(grid), dim3(256), 0, s, h, sc, h, W, H, sigma,
                     density_sigma, 0);
  hipLaunchKernelGGL(blur_kernel, dim3(grid), dim3(256), 0, s, sc, h, h, W, H, sigma,
                     density_sigma, 1);
  return hipGetLastError();
}

// ---------------------------------------------------- density estimation

// Gather form of the flam3 estimator the host tabulates: pixel q of count
// c uses filter idx(c) with squared half-width h2[idx] and normaliser
// inv_norm[idx]; q reaches p when |p-q|^2 <= h2, with weight
// exp(-4.5 |p-q|^2 / h2) * inv_norm. Each block gathers a 16x16 output
// tile from an LDS tile holding the (16+2R)^2 sources, over the taps of the
// disc |p-q|^2 <= R^2 (no filter reaches further: h2 <= R^2), which is the
// same for every thread.
constexpr int DE_T = 16;
constexpr int DE_RMAX = 16;

__device__ __forceinline__ int de_index(float c, float curve, int n_filt) {
  int idx;
  if (c <= 100.0f)
    idx = (int)ceilf(c) - 1;
  else
    idx = 100 + (int)floorf(powf(c - 100.0f, curve));
  return min(max(idx, 0), n_filt - 1);
}

__global__ __launch_bounds__(256) void de_kernel(const float4* __restrict__ in,
                                                 float4* __restrict__ out, int W, int H,
                                                 const float* __restrict__ filt_h2,
                                                 const float* __restrict__ filt_inv, int n_filt,
                                                 int R, float curve) {
  extern __shared__ float4 de_lds[];
  const int S = DE_T + 2 * R;
  float4* val = de_lds;                                   // S*S: (r c, g c, b c, c)
  float2* flt = reinterpret_cast<float2*>(de_lds + S * S);  // S*S: (4.5/h2, inv_norm) ; h2
  float* fh2 = reinterpret_cast<float*>(flt + S * S);
  const int x0 = blockIdx.x * DE_T - R, y0 = blockIdx.y * DE_T - R;
  for (int i = threadIdx.x; i < S * S; i += blockDim.x) {
    int sx = x0 + i % S, sy = y0 + i / S;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 f = make_float2(0.f, 0.f);
    float h2 = -1.0f;
    if (sx >= 0 && sx < W && sy >= 0 && sy < H) {
      float4 a = in[(size_t)sy * W + sx];
      float c = a.w;
      if (c > 0.0f) {
        v = make_float4(a.x * c, a.y * c, a.z * c, c);
        int idx = de_index(c, curve, n_filt);
        h2 = filt_h2[idx];
        f = make_float2(4.5f / h2, filt_inv[idx]);
      }
    }
    val[i] = v;
    flt[i] = f;
    fh2[i] = h2;
  }
  __syncthreads();
  const int tx = threadIdx.x % DE_T, ty = threadIdx.x / DE_T;
  const int ox = blockIdx.x * DE_T + tx, oy = blockIdx.y * DE_T + ty;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int dy = -R; dy <= R; dy++) {
    const int row = (ty + R + dy) * S + tx + R;
    for (int dx = -R; dx <= R; dx++) {
      if (dx * dx + dy * dy > R * R) continue;
      const int i = row + dx;
      const float d2 = (float)(dx * dx + dy * dy);
      if (d2 <= fh2[i]) {
        float w = expf(-d2 * flt[i].x) * flt[i].y;
        float4 v = val[i];
        acc.x += v.x * w;
        acc.y += v.y * w;
        acc.z += v.z * w;
        acc.w += v.w * w;
      }
    }
  }
  if (ox < W && oy < H) {
    float4 o;
    if (acc.w > 0.0f) {
      o = make_float4(acc.x / acc.w, acc.y / acc.w, acc.z / acc.w, acc.w);
    } else {
      o = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    out[(size_t)oy * W + ox] = o;
  }
}

hipError_t launch_density_estimate(const float* in, float* out, int W, int H, const float* filt_h2,
                                   const float* filt_inv_norm, int n_filt, int max_R, float curve,
                                   hipStream_t s) {
  if (max_R > DE_RMAX) return hipErrorInvalidValue;
  const int S = DE_T + 2 * max_R;
  size_t shm = (size_t)S * S * (sizeof(float4) + sizeof(float2) + sizeof(float));
  dim3 grid((W + DE_T - 1) / DE_T, (H + DE_T - 1) / DE_T);
  hipLaunchKernelGGL(de_kernel, grid, dim3(DE_T * DE_T), shm, s,
                     reinterpret_cast<const float4*>(in), reinterpret_cast<float4*>(out), W, H,
                     filt_h2, filt_inv_norm, n_filt, max_R, curve);
  return hipGetLastError();
}

// --------------------------------------------------------------- tonemap

__device__ __forceinline__ float clampf(float v, float lo, float hi) {
  return fminf(fmaxf(v, lo), hi);
}

__device__ __forceinline__ uint8_t unorm8(float v) {
  v = (v != v) ? 0.0f : clampf(v, 0.0f, 1.0f);
  return (uint8_t)floorf(v * 255.0f + 0.5f);
}

// The 256-texel curve table under a linear, clamp-to-edge sampler at
// texel-space position s = 255c: lut[i] (1 - f) + lut[i + 1] f with
// i = floor(s), f = s - i; s <= 0 (or NaN) reads texel 0, s >= 255 texel 255.
__device__ __forceinline__ float curve_sample(const float* lut, float c) {
  const float s = fminf(fmaxf(c * 255.0f, 0.0f), 255.0f);
  const int i = (int)s;
  const float f = s - (float)i;
  const int j = i < 255 ? i + 1 : 255;
  return lut[i] * (1.0f - f) + lut[j] * f;
}

// The tone map's transcendentals and quotients on the hardware's
// v_log_f32 / v_exp_f32 / v_rcp_f32 (about 1 ulp each): x^g =
// exp2(g log2 x), log10 x = log2 x log10(2), a / b = a rcp(b). The output
// is 8-bit, a level being 1/255 of the range, so these few-ulp differences
// from the library forms move a channel by at most one level (at a rounding
// tie). tm_pow keeps powf's special values on the tone map's inputs: x = 0
// gives 0 for g > 0 and +inf for g < 0, g = 0 gives 1, x < 0 or NaN NaN.
__device__ __forceinline__ float tm_pow(float x, float g) {
  if (g == 0.0f) return 1.0f;
  return __builtin_amdgcn_exp2f(g * __builtin_amdgcn_logf(x));
}

__device__ __forceinline__ float tm_log10(float x) {
  return __builtin_amdgcn_logf(x) * 0.30102999566398120f;
}

__device__ __forceinline__ float tm_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// shaders/tonemap.wgsl:153-211
__device__ void hsv_adjust(float c[3], float hue_shift, float saturation) {
  float r = c[0], g = c[1], b = c[2];
  float mx = fmaxf(fmaxf(r, g), b), mn = fminf(fminf(r, g), b);
  float delta = mx - mn;
  float h = 0.0f, s = 0.0f, v = mx;
  if (delta > 0.00001f) {
    s = tm_div(delta, mx);
    if (r >= mx)
      h = tm_div(g - b, delta);
    else if (g >= mx)
      h = 2.0f + tm_div(b - r, delta);
    else
      h = 4.0f + tm_div(r - g, delta);
    h = h * 60.0f;
    if (h < 0.0f) h = h + 360.0f;
  }
  if (hue_shift != 0.0f) {
    h = h + hue_shift;
    if (h < 0.0f)
      h = h + 360.0f;
    else if (h >= 360.0f)
      h = h - 360.0f;
  }
  if (saturation != 1.0f) s = clampf(s * saturation, 0.0f, 1.0f);
  if (s <= 0.0f) {
    c[0] = c[1] = c[2] = v;
    return;
  }
  float hh = h;
  if (hh >= 360.0f) hh = 0.0f;
  hh = hh * (1.0f / 60.0f);
  uint32_t i = (uint32_t)hh;
  float ff = hh - (float)i;
  float p = v * (1.0f - s);
  float q = v * (1.0f - (s * ff));
  float t = v * (1.0f - (s * (1.0f - ff)));
  switch (i) {
    case 0: c[0] = v; c[1] = t; c[2] = p; break;
    case 1: c[0] = q; c[1] = v; c[2] = p; break;
    case 2: c[0] = p; c[1] = v; c[2] = t; break;
    case 3: c[0] = p; c[1] = q; c[2] = v; break;
    case 4: c[0] = t; c[1] = p; c[2] = v; break;
    default: c[0] = v; c[1] = p; c[2] = q; break;
  }
}

// One pixel of tonemap.wgsl fs_main (shaders/tonemap.wgsl:224-590), with the
// uniform terms precomputed on the host.
__device__ __forceinline__ uchar4 tonemap_one(const float4 a, const TonemapConst& t) {
  const float bucket_count = a.w * 100.0f;
  float color[3];
  float alpha = 0.0f;
  if (t.mode == 0) {
    color[0] = tm_pow(a.x * t.exposure, t.gamma);
    color[1] = tm_pow(a.y * t.exposure, t.gamma);
    color[2] = tm_pow(a.z * t.exposure, t.gamma);
    alpha = clampf(bucket_count * 0.01f * t.density_scale, 0.0f, 1.0f);
  } else if (t.mode == 2) {
    float nd = clampf(bucket_count * 0.01f * t.exposure, 0.0f, 1.0f);
    float gd = tm_pow(nd, t.gamma);
    color[0] = color[1] = color[2] = gd;
    alpha = nd;
  } else {
    float bs = 0.0f;
    if (!(bucket_count < 0.001f)) {
      float l10 = tm_log10(1.0f + t.white * bucket_count * t.k2);
      bs = tm_div(t.k1 * l10, t.white * bucket_count);
    }
    float ls = bs * (1.0f / 67108864.0f);
    float fp0 = ls * (a.x * bucket_count);
    float fp1 = ls * (a.y * bucket_count);
    float fp2 = ls * (a.z * bucket_count);
    float fp3 = ls * bucket_count * t.white;
    if (fp3 > 0.0f) {
      if (fp3 <= t.gamma_threshold) {
        float frac = tm_div(fp3, t.gamma_threshold);
        alpha = (1.0f - frac) * fp3 * t.funcval + frac * tm_pow(fp3, t.gamma);
      } else {
        alpha = tm_pow(fp3, t.gamma);
      }
    }
    ls = fp3 > 0.0f ? tm_div(t.vib * alpha, fp3) : 0.0f;
    if (t.notvib > 0.0f) {
      color[0] = ls * fp0 + t.notvib * tm_pow(fp0, t.gamma);
      color[1] = ls * fp1 + t.notvib * tm_pow(fp1, t.gamma);
      color[2] = ls * fp2 + t.notvib * tm_pow(fp2, t.gamma);
    } else {
      color[0] = ls * fp0;
      color[1] = ls * fp1;
      color[2] = ls * fp2;
    }
  }
  if (t.saturation != 1.0f || t.hue_shift != 0.0f) hsv_adjust(color, t.hue_shift, t.saturation);
  if (t.mode == 1) {
    color[0] = color[0] * t.exposure;
    color[1] = color[1] * t.exposure;
    color[2] = color[2] * t.exposure;
  }
  for (int k = 0; k < 3; k++) color[k] = fmaxf(color[k], 0.0f);
  if (t.highlight == 1) {
    float m = fmaxf(color[0], fmaxf(color[1], color[2]));
    if (m > 1.0f)
      for (int k = 0; k < 3; k++) color[k] = tm_div(color[k], m);
  } else if (t.highlight == 2) {
    float lum = color[0] * 0.2126f + color[1] * 0.7152f + color[2] * 0.0722f;
    if (lum > 1e-6f) {
      float lm = tm_div(lum, 1.0f + lum);
      for (int k = 0; k < 3; k++) color[k] = color[k] * tm_div(lm, lum);
    }
    for (int k = 0; k < 3; k++) color[k] = fminf(color[k], 1.0f);
  } else if (t.highlight == 3) {
    for (int k = 0; k < 3; k++) {
      float c = color[k];
      c = tm_div(c * (2.51f * c + 0.03f), c * (2.43f * c + 0.59f) + 0.14f);
      color[k] = clampf(c, 0.0f, 1.0f);
    }
  } else {
    for (int k = 0; k < 3; k++) color[k] = fminf(color[k], 1.0f);
  }
  // The tone curve (shaders/tonemap.wgsl:479-492): texel i of the table
  // holds the curve at i/255 and is read with a linear-filtering,
  // clamp-to-edge sampler (src/gpu/buffers.rs:1877-1886) at
  // u = (255c + 0.5)/256, whose texel-space position is 255c: the lerp of
  // texels floor(255c) and the next by the fraction, the edge texel outside.
  if (t.use_curve && bucket_count > 0.001f) {
    for (int k = 0; k < 3; k++) color[k] = curve_sample(t.curve_lut, color[k]);
  }
  float leveled = 1.0f;
  if (t.sample_density > 0.0f) {
    float nd = tm_div(a.w, t.sample_density);
    if (t.levels_high <= t.levels_low) {
      leveled = nd > t.levels_low ? 1.0f : 0.0f;
    } else {
      float nrm = tm_div(nd - t.levels_low, t.levels_high - t.levels_low);
      float cl = clampf(nrm, 0.0f, 1.0f);
      leveled = (t.levels_gamma != 1.0f && t.levels_gamma > 0.0f) ? tm_pow(cl, t.levels_gamma) : cl;
    }
  }
  float linear_alpha = clampf(bucket_count * 0.01f * t.density_scale, 0.0f, 1.0f);
  float gamma_alpha = clampf(alpha, 0.0f, 1.0f);
  float st = clampf(tm_div(gamma_alpha - t.alpha_low, t.alpha_high - t.alpha_low), 0.0f, 1.0f);
  float blend_t = st * st * (3.0f - 2.0f * st);
  float base_alpha = gamma_alpha * (1.0f - blend_t) + linear_alpha * blend_t;
  float fa = t.levels_enabled ? fminf(base_alpha, leveled) : base_alpha;
  uchar4 o;
  if (t.transparent) {
    // transparent export (tonemap.wgsl:531-576): the over-black picture
    // C = srgb(colour x alpha); premultiplied keeps C with the opacity,
    // straight alpha takes coverage = max(C) and stores C / coverage
    float ob[3];
    for (int k = 0; k < 3; k++) ob[k] = tm_pow(fmaxf(color[k] * fa, 0.0f), t.inv22);
    if (t.transparent == 2) {
      o.x = unorm8(ob[0]);
      o.y = unorm8(ob[1]);
      o.z = unorm8(ob[2]);
      o.w = unorm8(fa);
      return o;
    }
    float cov = fmaxf(ob[0], fmaxf(ob[1], ob[2]));
    o.x = unorm8(cov > 0.0f ? tm_div(ob[0], cov) : 0.0f);
    o.y = unorm8(cov > 0.0f ? tm_div(ob[1], cov) : 0.0f);
    o.z = unorm8(cov > 0.0f ? tm_div(ob[2], cov) : 0.0f);
    o.w = unorm8(cov);
    return o;
  }
  o.x = unorm8(tm_pow(t.bg_lin[0] * (1.0f - fa) + color[0] * fa, t.inv22));
  o.y = unorm8(tm_pow(t.bg_lin[1] * (1.0f - fa) + color[1] * fa, t.inv22));
  o.z = unorm8(tm_pow(t.bg_lin[2] * (1.0f - fa) + color[2] * fa, t.inv22));
  o.w = 255;
  return o;
}

// fs_main as a compute pass over the accumulator.
__global__ __launch_bounds__(256) void tonemap_kernel(const float4* __restrict__ accum,
                                                      uchar4* __restrict__ rgba, size_t npix,
                                                      TonemapConst t) {
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (size_t)gridDim.x * blockDim.x)
    rgba[p] = tonemap_one(accum[p], t);
}

hipError_t launch_tonemap(const float* accum, uint8_t* rgba, int W, int H, const TonemapConst& tc,
                          hipStream_t s) {
  size_t npix = (size_t)W * H;
  hipLaunchKernelGGL(tonemap_kernel, dim3(grid_for(npix, 256)), dim3(256), 0, s,
                     reinterpret_cast<const float4*>(accum), reinterpret_cast<uchar4*>(rgba), npix,
                     tc);
  return hipGetLastError();
}

// ------------------------------------------------------------- batches

// Accumulate (fresh accumulator, cumulative mode) then tonemap, for many
// flames in one launch: blockIdx.y is the flame, blocks stride its pixels.
__global__ __launch_bounds__(256) void batch_finish_kernel(const uint4* __restrict__ hist,
                                                           uchar4* const* __restrict__ out,
                                                           const BatchPix* __restrict__ fl) {
  // by reference: a copy of the descriptor (its 1 KB curve table indexed by
  // pixel values) would live in scratch memory
  const BatchPix& f = fl[blockIdx.y];
  const uint4* h = hist + f.pix_off;
  uchar4* o = out[blockIdx.y];
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < f.npix;
       p += (size_t)gridDim.x * blockDim.x) {
    const uint4 v = h[p];
    const float4 a = accum_one(make_float4(0.f, 0.f, 0.f, 0.f), (float)v.x, (float)v.y, (float)v.z,
                               (float)v.w, v.w, nullptr, 0, 0.0f);
    o[p] = tonemap_one(a, f.tc);
  }
}

hipError_t launch_batch_finish(const uint32_t* hist, uint8_t* const* out, const BatchPix* fl,
                               int n, hipStream_t s) {
  hipLaunchKernelGGL(batch_finish_kernel, dim3(64, n), dim3(256), 0, s,
                     reinterpret_cast<const uint4*>(hist), reinterpret_cast<uchar4* const*>(out),
                     fl);
  return hipGetLastError();
}

}  // namespace ffm

// ------------------------------------------------------------ supersample

namespace ffm {

// downsample_2x_firefly (src/export/supersample.rs:117-187): each output
// pixel is the 2x2 box of its four RGBA8 samples (row-major order
// (0,0), (1,0), (0,1), (1,1)), with every sample whose luminance proxy
// r + 2g + b exceeds 4 x the quad's median + 64 scaled down to that cap
// first; alpha is averaged unclamped; the mean rounds half away from zero.
// The median of four is the mean of the two middle values: the larger of
// the pairwise minima and the smaller of the pairwise maxima.
__global__ __launch_bounds__(256) void downsample2x_kernel(const uchar4* __restrict__ in,
                                                           uchar4* __restrict__ out, int ow,
                                                           int oh) {
  const int w2 = 2 * ow;
  const size_t n = (size_t)ow * oh;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (size_t)gridDim.x * blockDim.x) {
    const int ox = (int)(p % (size_t)ow), oy = (int)(p / (size_t)ow);
    const size_t r0 = (size_t)(2 * oy) * w2 + 2 * ox, r1 = r0 + w2;
    const uchar4 q[4] = {in[r0], in[r0 + 1], in[r1], in[r1 + 1]};
    float s[4][4], lum[4];
    for (int i = 0; i < 4; i++) {
      s[i][0] = (float)q[i].x;
      s[i][1] = (float)q[i].y;
      s[i][2] = (float)q[i].z;
      s[i][3] = (float)q[i].w;
      lum[i] = s[i][0] + 2.0f * s[i][1] + s[i][2];
    }
    const float lo = fmaxf(fminf(lum[0], lum[1]), fminf(lum[2], lum[3]));
    const float hi = fminf(fmaxf(lum[0], lum[1]), fmaxf(lum[2], lum[3]));
    const float median = 0.5f * (lo + hi);
    const float cap = median * 4.0f + 64.0f;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 4; i++) {
      const float scale = lum[i] > cap ? cap / lum[i] : 1.0f;
      acc[0] += s[i][0] * scale;
      acc[1] += s[i][1] * scale;
      acc[2] += s[i][2] * scale;
      acc[3] += s[i][3];
    }
    uchar4 o;
    o.x = (uint8_t)fminf(fmaxf(roundf(acc[0] * 0.25f), 0.0f), 255.0f);
    o.y = (uint8_t)fminf(fmaxf(roundf(acc[1] * 0.25f), 0.0f), 255.0f);
    o.z = (uint8_t)fminf(fmaxf(roundf(acc[2] * 0.25f), 0.0f), 255.0f);
    o.w = (uint8_t)fminf(fmaxf(roundf(acc[3] * 0.25f), 0.0f), 255.0f);
    out[p] = o;
  }
}

hipError_t launch_downsample2x(const uint8_t* in, uint8_t* out, int ow, int oh, hipStream_t s) {
  const size_t n = (size_t)ow * oh;
  hipLaunchKernelGGL(downsample2x_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     reinterpret_cast<const uchar4*>(in), reinterpret_cast<uchar4*>(out), ow, oh);
  return hipGetLastError();
}

}  // namespace ffm
