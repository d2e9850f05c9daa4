// ffm_device.h -- gfx950 device code of the chaos-game step.
//
// Semantics follow the reference kernel shaders/core/main_template.wgsl
// (assembled 2D reading: tests/shader_dumps/in-app-2d.wgsl:1097-1400).
// Every FP32 expression is written in the reference's operation order and
// the files that include this header are built with -ffp-contract=off, so
// the +,-,*,/ chains round exactly as the reference's source states them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ffm_fastmath.h"
#include "ffm_internal.h"

namespace ffm {

// XCC id of the XCD this wave runs on (gfx940+ hardware register), 0..7:
// kernels that keep one private copy of an image per XCD index it with this.
__device__ __forceinline__ uint32_t xcc_id() {
  uint32_t v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return v & 7u;
}

// ---------------------------------------------------------------- PCG32

// pcg_hash (shaders/core/rng.wgsl:22-27).
__device__ __forceinline__ uint32_t pcg_hash(uint32_t input) {
  uint32_t state = input * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

struct Rng {
  uint32_t state, inc;
};

// rng_init: per-walker stream (shaders/core/rng.wgsl:29-37).
__device__ __forceinline__ Rng rng_init(uint32_t walker, uint32_t seed) {
  Rng r;
  r.state = pcg_hash(walker ^ seed);
  r.inc = (pcg_hash(pcg_hash(walker) + seed) << 1u) | 1u;
  return r;
}

// rng_next: LCG step then RXS-M-XS output of the old state (rng.wgsl:39-45).
__device__ __forceinline__ uint32_t rng_next(Rng& r) {
  uint32_t old = r.state;
  r.state = old * 747796405u + r.inc;
  uint32_t xs = ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
  return (xs >> 22u) ^ xs;
}

// The LCG step alone: a draw whose value is not needed (its output
// permutation skipped) still advances the stream.
__device__ __forceinline__ void rng_skip(Rng& r) { r.state = r.state * 747796405u + r.inc; }

// Top 24 bits times 2^-24 (rng.wgsl:53-55): exact in f32.
__device__ __forceinline__ float rng_nextf(Rng& r) {
  return (float)(rng_next(r) >> 8u) * (1.0f / 16777216.0f);
}

// Pointers into a flame table in global memory, in the constant address
// space: the table does not change while a kernel runs, so the wave-uniform
// reads of it (variation ids, final-chain slot lists) are scalar loads,
// with no VALU address arithmetic and no readfirstlane.
typedef const __attribute__((address_space(4))) int KInt;
typedef const __attribute__((address_space(4))) FlameHeader KHeader;

// ----------------------------------------------------------- variations

constexpr float PI_F = 3.14159265359f;        // the reference's literal
constexpr float TWO_PI_F = 6.28318530718f;
constexpr float PI_ATAN = 3.14159265358979f;  // ff_atan2's literal

// ff_atan2 (shaders/core/utilities.wgsl:604-611).
__device__ __forceinline__ float ff_atan2(float y, float x) {
  if (y == 0.0f && x == 0.0f) {
    float mag = (__float_as_uint(x) & 0x80000000u) ? PI_ATAN : 0.0f;
    return (__float_as_uint(y) & 0x80000000u) ? -mag : mag;
  }
  return ffm_atan2f(y, x);
}

// The variation id is uniform across the wave, so `switch (id)` is a scalar
// branch. Every body depends only on the affine point, which does not
// change over the loop of active variations, so without a barrier the
// compiler hoists ALL bodies' arithmetic (divisions, sqrt, sincos, atan2)
// out of that loop and pays for every one on every transform evaluation.
// The loops over variation slots pass the point and its r^2 = x^2 + y^2
// (computed once for every body that reads it, in the bodies' own order:
// x*x first) through an empty volatile asm at the top of every trip
// (FFM_OPAQUE_POINT): they are then new values each trip, in the same
// registers (no copies), so each body stays behind its branch.
#define FFM_OPAQUE_POINT(px, py, pr2, ptp) \
  asm volatile("" : "+v"(px), "+v"(py), "+v"(pr2), "+v"(ptp))

// One variation body V(x, y). `row` is the transform's table row (waves
// reads its own b, e, d, f), `p` the variation's params. Bodies:
// src/variations/defs/basic.rs:37-205, advanced.rs:32-1700. Quotients use
// the refined hardware reciprocal (ffm_fdiv, ffm_rcp) and square roots the
// 1-ulp hardware form (ffm_sqrt); the reference's WGSL allows 2.5 ulp for
// both, and the per-step bar against the IEEE oracle is 4 ulp.
// The variations beyond the thirteen of the workloads (blob and the RNG
// blurs of the reference's presets), kept out of eval_variation's switch
// unless a flame uses them (the EXT kernels).
__device__ __forceinline__ void eval_ext(int id, const float* __restrict__ p, float x, float y,
                                         float r2, Rng& rng, float& ox, float& oy) {
  switch (id) {
    case FFM_VAR_BLOB: {
      float high = p[0], low = p[1], waves = p[2];
      float r = ffm_sqrt(r2);
      float theta = ff_atan2(x, y);  // Apophysis's atan2(x, y)
      float scale = r * (low + ((high - low) / 2.0f) * (ffm_sinf(waves * theta) + 1.0f));
      float s, c;
      ffm_sincosf(theta, &s, &c);
      ox = scale * c;
      oy = scale * s;
      break;
    }
    case FFM_VAR_NOISE: {
      float theta = rng_nextf(rng) * TWO_PI_F;
      float r = rng_nextf(rng);
      float s, c;
      ffm_sincosf(theta, &s, &c);
      ox = x * r * c;
      oy = y * r * s;
      break;
    }
    case FFM_VAR_BLUR: {
      float theta = rng_nextf(rng) * TWO_PI_F;
      float r = rng_nextf(rng);
      float s, c;
      ffm_sincosf(theta, &s, &c);
      ox = r * c;
      oy = r * s;
      break;
    }
    case FFM_VAR_GAUSSIAN_BLUR: {
      float theta = rng_nextf(rng) * TWO_PI_F;
      float d1 = rng_nextf(rng);
      float d2 = rng_nextf(rng);
      float d3 = rng_nextf(rng);
      float d4 = rng_nextf(rng);
      float r = d1 + d2 + d3 + d4 - 2.0f;
      float s, c;
      ffm_sincosf(theta, &s, &c);
      ox = r * c;
      oy = r * s;
      break;
    }
    default:
      ox = 0.0f;
      oy = 0.0f;
      break;
  }
}

template <bool EXT>
__device__ __forceinline__ void eval_variation(int id, const float* __restrict__ row,
                                               const float* __restrict__ p, float x, float y,
                                               float r2, bool has_tp, float tp, Rng& rng,
                                               float& ox, float& oy) {
  switch (id) {
    case FFM_VAR_LINEAR:
      ox = x;
      oy = y;
      break;
    case FFM_VAR_SINUSOIDAL:
      ox = ffm_sinf(x);
      oy = ffm_sinf(y);
      break;
    case FFM_VAR_SPHERICAL: {
      const float r2e = r2 + 1e-6f;
      ox = ffm_fdiv(x, r2e);
      oy = ffm_fdiv(y, r2e);
      break;
    }
    case FFM_VAR_SWIRL: {
      float s, c;
      ffm_sincosf(r2, &s, &c);
      ox = x * s - y * c;
      oy = x * c + y * s;
      break;
    }
    case FFM_VAR_HORSESHOE: {
      float r = ffm_sqrt_cr(r2) + 1e-6f;  // its rounding reaches both outputs
      float r_inv = ffm_rcp(r);
      ox = (x - y) * (x + y) * r_inv;
      oy = 2.0f * x * y * r_inv;
      break;
    }
    case FFM_VAR_POLAR: {
      float r = ffm_sqrt(r2);
      ox = has_tp ? tp : ffm_fdiv(ffm_atan2f(x, y), PI_F);
      oy = r - 1.0f;
      break;
    }
    case FFM_VAR_DISC: {
      float r = ffm_sqrt_cr(r2);  // feeds the angle pi r
      float theta_pi = has_tp ? tp : ffm_fdiv(ffm_atan2f(x, y), PI_F);
      float pi_r = PI_F * r;
      float s, c;
      ffm_sincosf(pi_r, &s, &c);
      ox = theta_pi * s;
      oy = theta_pi * c;
      break;
    }
    case FFM_VAR_HYPERBOLIC: {
      ox = ffm_fdiv(x, r2 + 1e-6f);
      oy = y;
      break;
    }
    case FFM_VAR_JULIA: {
      float r = ffm_sqrt(r2);
      float theta = ffm_atan2f(y, x);
      float sqrt_r = ffm_sqrt(r);
      float omega = (rng_nextf(rng) < 0.5f) ? PI_F : 0.0f;
      float half_theta = theta / 2.0f + omega;
      float s, c;
      // |theta / 2 + omega| <= 3 pi / 2: always the short reduction
      ffm_sincos_reduced(half_theta, &s, &c);
      ox = sqrt_r * c;
      oy = sqrt_r * s;
      break;
    }
    case FFM_VAR_WAVES: {
      float b = row[XF_B], e = row[XF_E], d = row[XF_D], f = row[XF_F];
      // the row is fixed over the variation loop too: keep e, f opaque here
      asm volatile("" : "+v"(e), "+v"(f));
      ox = x + b * ffm_sinf(ffm_fdiv(y, e * e + 1e-6f));
      oy = y + d * ffm_sinf(ffm_fdiv(x, f * f + 1e-6f));
      break;
    }
    case FFM_VAR_JULIAN: {
      float power = p[0];
      float cpower = p[2];
      float r = powf(r2, cpower);
      float theta = ff_atan2(y, x);
      float trunc_val = floorf(fabsf(power) * rng_nextf(rng));
      float t = ffm_fdiv(theta + TWO_PI_F * trunc_val, power);
      float s, c;
      ffm_sincosf(t, &s, &c);
      ox = r * c;
      oy = r * s;
      break;
    }
    case FFM_VAR_BUBBLE: {
      float r = r2 / 4.0f + 1.0f;
      float scale = ffm_rcp(r);
      ox = scale * x;
      oy = scale * y;
      break;
    }
    case FFM_VAR_CURL: {
      float c1 = p[0], c2 = p[1];
      float re = 1.0f + c1 * x + c2 * (x * x - y * y);
      float im = c1 * y + 2.0f * c2 * x * y;
      float r = ffm_rcp(re * re + im * im);
      ox = (x * re + y * im) * r;
      oy = (y * re - x * im) * r;
      break;
    }
    default:
      if constexpr (EXT) {
        eval_ext(id, p, x, y, r2, rng, ox, oy);
      } else {
        ox = 0.0f;
        oy = 0.0f;
      }
      break;
  }
}

// pre_blur (src/variations/defs/blur.rs PRE_BLUR), a pre-phase variation:
// it moves the affine point that the normal variations then read, by its
// own weight times (six draws summed - 3) at the angle 2 pi x a seventh
// draw. The packed flame puts pre-phase slots first
// (src/shader_builder_v2.rs:2596-2650 runs them before the weighted sum).
constexpr float PRE_BLUR_TWO_PI = 6.28318530717959f;
__device__ __forceinline__ void pre_blur(float w, float& x, float& y, Rng& rng) {
  float g1 = rng_nextf(rng), g2 = rng_nextf(rng), g3 = rng_nextf(rng);
  float g4 = rng_nextf(rng), g5 = rng_nextf(rng), g6 = rng_nextf(rng);
  float rnd_g = w * (g1 + g2 + g3 + g4 + g5 + g6 - 3.0f);
  float rnd_a = rng_nextf(rng) * PRE_BLUR_TWO_PI;
  float s, c;
  ffm_sincosf(rnd_a, &s, &c);
  x = x + rnd_g * c;
  y = y + rnd_g * s;
}

// Affine -> weighted variation sum in the flame's local-index order
// (src/shader_builder_v2.rs:2525-2806) -> post-affine when enabled
// (main_template.wgsl:355-384). `ids` are the flame's variation ids
// (uniform across the wave), `row` the transform's row in LDS.
template <bool EXT>
__device__ __forceinline__ void apply_xform(const float* __restrict__ row, KInt* ids, int nv,
                                            float& x, float& y, Rng& rng, bool has_tp = false) {
  float ax = row[XF_A] * x + row[XF_B] * y + row[XF_E];
  float ay = row[XF_C] * x + row[XF_D] * y + row[XF_F];
  float rx = 0.0f, ry = 0.0f;
  float r2 = ax * ax + ay * ay;
  float tp = has_tp ? ffm_fdiv(ffm_atan2f(ax, ay), PI_F) : 0.0f;
  for (int k = 0; k < nv; k++) {
    FFM_OPAQUE_POINT(ax, ay, r2, tp);
    int id = __builtin_amdgcn_readfirstlane(ids[k]);
    float w = row[XF_W0 + k];
    if (w != 0.0f) {
      if (EXT && id == FFM_VAR_PRE_BLUR) {
        pre_blur(w, ax, ay, rng);
        r2 = ax * ax + ay * ay;
        if (has_tp) tp = ffm_fdiv(ffm_atan2f(ax, ay), PI_F);
        continue;
      }
      float vx, vy;
      eval_variation<EXT>(id, row, row + XF_P0 + k * FFM_MAX_VAR_PARAMS, ax, ay, r2, has_tp, tp, rng, vx,
                          vy);
      rx += w * vx;
      ry += w * vy;
    }
  }
  if (row[XF_POST] != 0.0f) {
    float qx = row[XF_PA] * rx + row[XF_PB] * ry + row[XF_PE];
    float qy = row[XF_PC] * rx + row[XF_PD] * ry + row[XF_PF];
    rx = qx;
    ry = qy;
  }
  x = rx;
  y = ry;
}

// The flame as one lane sees it: LDS rows, plus the header and variation
// ids read with scalar loads from the table in global memory.
struct FlameView {
  const float* rows;         // LDS, (n_xforms + n_finals) * XF_STRIDE
  KInt* ids;                 // the table's variation ids (global), n_vars
  KHeader* hdr;              // the table's header (final-chain slot lists; global), or null
  const float* xcum;         // LDS xaos selection table (n_xforms^2), or null without xaos
  const uint8_t* sel;        // LDS selection buckets (SEL_BITS of the draw), or null
  int sel_byte;              // their byte offset from rows (addressing; sel tells presence)
  int n_xforms, n_finals, n_vars;
  int share_tp;              // FlameHeader::share_tp
};

// Final transform i of the flame: apply_xform's arithmetic with the loop
// over the final's non-zero variation slots only (a zero weight adds
// nothing and draws nothing, so the sum is the same). Every lane applies the
// same final, so each row read is one uniform LDS address.
template <bool EXT>
__device__ __forceinline__ void apply_final(const FlameView& fl, int i, float& x, float& y,
                                            Rng& rng) {
  const float* row = fl.rows + (fl.n_xforms + i) * XF_STRIDE;
  float ax = row[XF_A] * x + row[XF_B] * y + row[XF_E];
  float ay = row[XF_C] * x + row[XF_D] * y + row[XF_F];
  float rx = 0.0f, ry = 0.0f;
  float r2 = ax * ax + ay * ay;
  const bool has_tp = (fl.share_tp >> (1 + i)) & 1;
  float tp = has_tp ? ffm_fdiv(ffm_atan2f(ax, ay), PI_F) : 0.0f;
  const int nact = __builtin_amdgcn_readfirstlane(fl.hdr->fin_nact[i]);
  for (int j = 0; j < nact; j++) {
    FFM_OPAQUE_POINT(ax, ay, r2, tp);
    const int code = __builtin_amdgcn_readfirstlane(fl.hdr->fin_act[i][j]);
    const int k = code & 0xFF, id = code >> 8;
    const float w = row[XF_W0 + k];
    if (EXT && id == FFM_VAR_PRE_BLUR) {
      pre_blur(w, ax, ay, rng);
      r2 = ax * ax + ay * ay;
      if (has_tp) tp = ffm_fdiv(ffm_atan2f(ax, ay), PI_F);
      continue;
    }
    float vx, vy;
    eval_variation<EXT>(id, row, row + XF_P0 + k * FFM_MAX_VAR_PARAMS, ax, ay, r2, has_tp, tp, rng, vx,
                        vy);
    rx += w * vx;
    ry += w * vy;
  }
  if (row[XF_POST] != 0.0f) {
    float qx = row[XF_PA] * rx + row[XF_PB] * ry + row[XF_PE];
    float qy = row[XF_PC] * rx + row[XF_PD] * ry + row[XF_PF];
    rx = qx;
    ry = qy;
  }
  x = rx;
  y = ry;
}

// Saturating f32 -> i32 as the reference's i32(floor(.)) behaves on GPUs:
// out-of-range values clamp to INT_MIN / INT_MAX and NaN gives 0, which is
// exactly v_cvt_i32_f32's definition (one instruction; a C cast would be
// undefined out of range).
__device__ __forceinline__ int32_t sat_i32(float f) {
  int32_t r;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(f));
  return r;
}

// world_to_pixel (shaders/core/utilities.wgsl:417-445).
__device__ __forceinline__ void world_to_pixel(const ViewConst& v, float x, float y, int32_t& px,
                                               int32_t& py) {
  float tx = x - v.pan_x, ty = y - v.pan_y;
  float rx = tx * v.cos_r - ty * v.sin_r;
  float ry = tx * v.sin_r + ty * v.cos_r;
  rx = rx * v.zoom;
  ry = ry * v.zoom;
  px = sat_i32(floorf(v.cx + rx * v.scale));
  py = sat_i32(floorf(v.cy + ry * v.scale));
}

// Nearest texel of the clamp-to-edge palette at u = colour index.
// u < 0 and NaN give texel 0 (fmaxf returns the non-NaN operand), u at or
// past size - 1 gives size - 1; in between, floor(u).
__device__ __forceinline__ int palette_texel(float ci, int size) {
  float u = fminf(fmaxf(ci * (float)size, 0.0f), (float)(size - 1));
  return (int)floorf(u);
}

enum StepFlags { ST_PLOT = 1, ST_RESPAWN = 2, ST_BURN = 4 };


// The first half of an iteration (main_template.wgsl:235-249): the
// transform pick and the opacity draw. select_transform_const
// (utilities.wgsl:35-55): first i with u * total <= cum_i, else the last;
// cum is the sequential f32 sum. The opacity draw is always consumed; a draw
// is at most 1 - 2^-24, so at opacity >= 1 it always passes and only
// advances the stream.
__device__ __forceinline__ int select_const(const FlameView& fl, float u) {
  const int n = fl.n_xforms;
  float target = u * fl.rows[(n - 1) * XF_STRIDE + XF_CUM];
  int xi = n - 1;
  for (int i = n - 2; i >= 0; i--)
    if (target <= fl.rows[i * XF_STRIDE + XF_CUM]) xi = i;
  return xi;
}

// The same pick from the 24-bit draw d (u = d 2^-24) through the selection
// buckets: one LDS byte; in a bucket that holds a boundary the pick is the
// first transform from the bucket's lowest pick on whose cumulative weight
// reaches the target (every earlier one falls short for the whole bucket).
__device__ __forceinline__ int select_bucket(const FlameView& fl, uint32_t d, float u) {
  const uint32_t e =
      reinterpret_cast<const uint8_t*>(fl.rows)[fl.sel_byte + (int)(d >> (24 - SEL_BITS))];
  int xi = (int)(e & 0x7Fu);
  if (e > 0x7Fu) {
    const int n = fl.n_xforms;
    const float target = u * fl.rows[(n - 1) * XF_STRIDE + XF_CUM];
    while (xi < n - 1 && !(target <= fl.rows[__umul24(xi, XF_STRIDE) + XF_CUM])) xi++;
  }
  return xi;
}

// With xaos the pick is select_transform_xaos (utilities.wgsl:107-140): the
// same rule over row `prev` of the table of running sums of
// weight[j] * xaos[prev][j], and the pick becomes the next `prev`.
template <bool EXT>
__device__ __forceinline__ int choose_xform(const FlameView& fl, Rng& rng, uint32_t& should_plot,
                                            int& prev) {
  const uint32_t d = rng_next(rng) >> 8u;
  const float u = (float)d * (1.0f / 16777216.0f);  // rng_nextf's value
  int xi;
  // flames without usable buckets take the EXT kernels
  if (!EXT || fl.sel) {
    xi = select_bucket(fl, d, u);
  } else if (EXT && fl.xcum) {
    const int n = fl.n_xforms;
    const float* xr = fl.xcum + prev * n;
    float target = u * xr[n - 1];
    xi = n - 1;
    for (int i = n - 2; i >= 0; i--)
      if (target <= xr[i]) xi = i;
    prev = xi;
  } else {
    xi = select_const(fl, u);
  }
  const float opacity = fl.rows[__umul24(xi, XF_STRIDE) + XF_OPACITY];
  should_plot = 1u;
  if (opacity >= 1.0f) rng_skip(rng);
  else should_plot = rng_nextf(rng) < opacity ? 1u : 0u;
  // held in a vector register over the step (as a lane mask it would live
  // in a scalar pair the whole step and spill)
  asm volatile("" : "+v"(should_plot));
  return xi;
}

// The rest of the iteration (main_template.wgsl:355-1383), 2D, palette
// colour mode, direct histogram, for transform xi. Walker state is
// (x, y, ci, fuse). Returns ST_* flags; on ST_PLOT, (px, py) is the
// in-frame pixel and `texel` the palette texel of the sample.
template <bool EXT>
__device__ __forceinline__ int step_xform(const FlameView& fl, const ViewConst& v, int pal_size,
                                          uint32_t burn_in, Rng& rng, float& x, float& y,
                                          float& ci, uint32_t& fuse, int xi,
                                          uint32_t should_plot, int32_t& px, int32_t& py,
                                          int& texel) {
  const int n = fl.n_xforms;
  const float* row = fl.rows + __umul24(xi, XF_STRIDE);
  float nx = x, ny = y;
  const int nv = fl.n_vars;
  if (!EXT) __builtin_assume(nv > 0);  // flames without variations take the EXT kernels
  apply_xform<EXT>(row, fl.ids, nv, nx, ny, rng, fl.share_tp & 1);
  // Bad-value respawn + re-fuse (main_template.wgsl:397-479).
  if (!(fabsf(nx) <= 1e32f) || !(fabsf(ny) <= 1e32f)) {
    x = rng_nextf(rng) * 2.0f - 1.0f;
    y = rng_nextf(rng) * 2.0f - 1.0f;
    fuse = burn_in;
    px = py = -1;
    return ST_RESPAWN;
  }
  x = nx;
  y = ny;
  // Colour-index blend, symmetry semantics (main_template.wgsl:496-503).
  ci = ci * row[XF_C1] + row[XF_C2];
  if (fuse > 0u) {
    fuse -= 1u;
    px = py = -1;
    return ST_BURN;
  }
  // Final chain: reshapes the plot only (main_template.wgsl:663-717), the
  // finals this transform attaches, in order.
  float fx = nx, fy = ny;
  // without EXT every transform applies the same chain (the host picks the
  // EXT kernels for chains that differ by transform), empty without finals
  if (!EXT || fl.n_finals > 0) {
    if (!EXT || fl.hdr->fin_uniform) {
      const int nc = __builtin_amdgcn_readfirstlane(fl.hdr->fin_chain_n);
      for (int j = 0; j < nc; j++)
        apply_final<EXT>(fl, __builtin_amdgcn_readfirstlane(fl.hdr->fin_chain[j]), fx, fy, rng);
    } else {
      const int nc = (int)row[XF_NFIN];
      for (int j = 0; j < nc; j++)
        apply_xform<EXT>(fl.rows + (n + (int)row[XF_FIN0 + j]) * XF_STRIDE, fl.ids, fl.n_vars, fx,
                         fy, rng);
    }
  }
  world_to_pixel(v, fx, fy, px, py);
  // in frame: 0 <= px < width and 0 <= py < height, as unsigned compares
  if ((uint32_t)px < (uint32_t)v.width && (uint32_t)py < (uint32_t)v.height && should_plot) {
    texel = palette_texel(ci, pal_size);
    return ST_PLOT;
  }
  return 0;
}

// One whole iteration of main_template.wgsl's loop body (:193-1383).
template <bool EXT>
__device__ __forceinline__ int chaos_step(const FlameView& fl, const ViewConst& v, int pal_size,
                                          uint32_t burn_in, Rng& rng, float& x, float& y,
                                          float& ci, uint32_t& fuse, int& prev, int& xi_out,
                                          int32_t& px, int32_t& py, int& texel) {
  uint32_t should_plot;
  const int xi = choose_xform<EXT>(fl, rng, should_plot, prev);
  xi_out = xi;
  return step_xform<EXT>(fl, v, pal_size, burn_in, rng, x, y, ci, fuse, xi, should_plot, px, py,
                    texel);
}

// Walker start or resume (main_template.wgsl:25-70,125-181). A start is
// the point (two draws), colour index 0 and a full fuse; with xaos one more
// draw picks the previous transform, unconditioned (select_transform_const),
// always. A persistent orbit (`orbit` not null) resumes its point, colour,
// fuse and, with xaos, previous transform (packed as fuse | prev << 24)
// unless this dispatch starts every orbit (`fresh`) or refreshes this walker
// ((walker + seed) % refresh_period == 0); an orbit that starts with xaos
// takes a first pick stratified over the walkers instead: the Kronecker
// point (walker * 2654435769 + seed) >> 8, over 2^24.
template <bool EXT>
__device__ __forceinline__ void walker_begin(const FlameView& fl, const IterArgs& a,
                                             uint32_t walker, const float4* orbit, Rng& rng,
                                             float& x, float& y, float& ci, uint32_t& fuse,
                                             int& prev) {
  rng = rng_init(walker, a.seed);
  x = rng_nextf(rng) * 2.0f - 1.0f;
  y = rng_nextf(rng) * 2.0f - 1.0f;
  ci = 0.0f;
  fuse = a.burn_in;
  prev = 0;
  if (EXT && fl.xcum) prev = select_const(fl, rng_nextf(rng));
  if (!orbit) return;
  const bool restart =
      a.fresh || (a.refresh_period != 0u && (walker + a.seed) % a.refresh_period == 0u);
  if (!restart) {
    const float4 o = *orbit;
    x = o.x;
    y = o.y;
    ci = o.z;
    const uint32_t w = __float_as_uint(o.w);
    fuse = EXT && fl.xcum ? (w & 0xFFFFFFu) : w;
    if (EXT && fl.xcum) prev = min((int)(w >> 24), fl.n_xforms - 1);  // a foreign word stays in the table
  } else if (EXT && fl.xcum) {
    const uint32_t k = walker * 2654435769u + a.seed;
    prev = select_const(fl, (float)(k >> 8) * (1.0f / 16777216.0f));
  }
}

// The orbit word that walker_begin reads back.
template <bool EXT>
__device__ __forceinline__ float4 walker_orbit(const FlameView& fl, float x, float y, float ci,
                                               uint32_t fuse, int prev) {
  const uint32_t w = EXT && fl.xcum ? (fuse | (uint32_t)prev << 24) : fuse;
  return make_float4(x, y, ci, __uint_as_float(w));
}

struct LdsFlame {
  FlameView fv;
  const uint32_t* pal_u32;
  const float* pal_f32;
};

// Copy the flame table and palette into LDS; returns the LDS views.
template <bool F32PAL>
__device__ __forceinline__ LdsFlame load_table(const float* table, const uint32_t* pal_u32,
                                               const float* pal_f32, int pal_size, uint32_t* lds) {
  const FlameHeader* hdr = reinterpret_cast<const FlameHeader*>(table);
  const int nx = hdr->n_xforms, nf = hdr->n_finals, nv = hdr->n_vars;
  const int rows_f = hdr->table_rows * XF_STRIDE;
  const int hdr_w = (int)(sizeof(FlameHeader) / 4);
  const uint32_t* src = reinterpret_cast<const uint32_t*>(table);
  for (int i = threadIdx.x; i < hdr_w + rows_f; i += blockDim.x) lds[i] = src[i];
  uint32_t* pal = lds + hdr_w + rows_f;
  for (int i = threadIdx.x; i < pal_size; i += blockDim.x) pal[i] = pal_u32[i];
  float* palf = reinterpret_cast<float*>(pal + pal_size);
  if (F32PAL)
    for (int i = threadIdx.x; i < 3 * pal_size; i += blockDim.x) palf[i] = pal_f32[i];
  __syncthreads();
  LdsFlame l;
  l.fv.rows = reinterpret_cast<const float*>(lds + hdr_w);
  l.fv.ids = (KInt*)table + 4;
  l.fv.hdr = (KHeader*)table;
  l.fv.xcum = __builtin_amdgcn_readfirstlane(hdr->xaos)
                  ? l.fv.rows + __builtin_amdgcn_readfirstlane(hdr->xaos_off)
                  : nullptr;
  const int sel_off = __builtin_amdgcn_readfirstlane(hdr->sel_off);
  l.fv.sel = sel_off ? reinterpret_cast<const uint8_t*>(l.fv.rows + sel_off) : nullptr;
  l.fv.sel_byte = sel_off * 4;
  l.fv.n_xforms = __builtin_amdgcn_readfirstlane(nx);
  l.fv.n_finals = __builtin_amdgcn_readfirstlane(nf);
  l.fv.n_vars = __builtin_amdgcn_readfirstlane(nv);
  l.fv.share_tp = __builtin_amdgcn_readfirstlane(hdr->share_tp);
  l.pal_u32 = pal;
  l.pal_f32 = palf;
  return l;
}

template <bool F32PAL>
__device__ __forceinline__ LdsFlame load_lds(const IterArgs& a, uint32_t* lds) {
  return load_table<F32PAL>(a.flame, a.pal_u32, a.pal_f32, a.pal_size, lds);
}

}  // namespace ffm
