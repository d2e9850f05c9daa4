// ffm_internal.h -- layouts shared by the host runtime and the gfx950 kernels.
//
// The flame table is what FlameRenderer::load_config uploads
// (src/renderer/compute_kernel.rs:2748; GpuTransform src/gpu/buffers.rs:94-160)
// re-laid for LDS: one row of XF_STRIDE floats per transform (odd stride, so
// lanes reading different transforms hit different LDS banks), followed by
// the per-flame uniform header.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ffm.h"

namespace ffm {

// Row layout of one transform in the table (floats).
enum XfField : int {
  XF_A = 0, XF_B, XF_C, XF_D, XF_E, XF_F,
  XF_PA, XF_PB, XF_PC, XF_PD, XF_PE, XF_PF,
  XF_CUM,      // sequential f32 cumulative selection weight (normal xforms)
  XF_C1,       // (1 + s) / 2          (main_template.wgsl:496-503)
  XF_C2,       // colour * (1 - s) / 2
  XF_OPACITY,
  XF_POST,     // 1.0 when the post-affine applies
  XF_W0,       // FFM_MAX_VARS variation weights
  XF_P0 = XF_W0 + FFM_MAX_VARS,  // FFM_MAX_VARS x FFM_MAX_VAR_PARAMS params
  XF_NFIN = XF_P0 + FFM_MAX_VARS * FFM_MAX_VAR_PARAMS,  // finals this transform applies
  XF_FIN0,                                              // their indices (FFM_MAX_FINALS)
  XF_STRIDE = XF_FIN0 + FFM_MAX_FINALS + 1  // 103: odd (one pad word)
};
static_assert(XF_STRIDE % 2 == 1, "odd row stride");

// Per-flame header that precedes the rows (ints stored as int32).
struct FlameHeader {
  int32_t n_xforms;
  int32_t n_finals;
  int32_t n_vars;
  // Bit 0: some transforms use polar and some use disc (or one both), so the
  // transform pass computes their common atan2(x, y) / pi once for every
  // lane; bit 1 + i: final i uses both.
  int32_t share_tp;
  int32_t var_id[FFM_MAX_VARS];
  // Each final transform's variation slots with a non-zero weight, in
  // local-index order, as slot | variation id << 8 (one scalar load a slot):
  // every lane applies the same final, so its loop runs over these only.
  int32_t fin_nact[FFM_MAX_FINALS];
  int32_t fin_act[FFM_MAX_FINALS][FFM_MAX_VARS];
  // 1 when every normal transform applies the same final chain (then
  // fin_chain lists it and the chain is uniform across a wave); 0 when the
  // chain depends on the transform (its row's XF_NFIN / XF_FIN0 list).
  int32_t fin_uniform;
  int32_t fin_chain_n;
  int32_t fin_chain[FFM_MAX_FINALS];
  // Rows the table holds after the header: the transforms' and finals'
  // rows, then, with xaos, ceil(n_xforms^2 / XF_STRIDE) rows holding the
  // xaos selection table at float xaos_off from the first row: for each
  // previous transform i, the sequential f32 sums of weight[j] * xaos[i][j]
  // (select_transform_xaos, shaders/core/utilities.wgsl:107-140).
  int32_t table_rows;
  int32_t xaos;
  int32_t xaos_off;
  // Without xaos, SEL_ROWS rows after the transforms' hold the pick of each
  // of 2^SEL_BITS buckets of the 24-bit selection draw, one byte each, at
  // float sel_off from the first row (0: the buckets are not usable, as
  // when a negative weight makes the cumulative sums decrease).
  int32_t sel_off;
};

// Rows of the xaos selection table of a flame with n transforms.
__host__ __device__ inline int xaos_rows(int n) { return (n * n + XF_STRIDE - 1) / XF_STRIDE; }

// Selection buckets: the top SEL_BITS bits of the 24-bit draw d index a
// byte holding the pick when it is the same for every d of the bucket (the
// pick never decreases with d while the cumulative weights never decrease),
// else the pick at the bucket's lowest d with bit 7 set (the pick is then
// the first transform from there whose cumulative weight reaches the target).
constexpr int SEL_BITS = 10;
constexpr int SEL_ROWS = ((1 << SEL_BITS) / 4 + XF_STRIDE - 1) / XF_STRIDE;
static_assert(FFM_MAX_XFORMS <= 128, "a bucket's pick fits in 7 bits");

// A flame table in device memory: header then n_rows rows.
__host__ __device__ inline size_t flame_table_floats(int n_rows) {
  return sizeof(FlameHeader) / 4 + (size_t)n_rows * XF_STRIDE;
}

// View constants with the uniform hoists of world_to_pixel done on the host
// (cos/sin of the rotation; shaders/core/utilities.wgsl:417-445).
struct ViewConst {
  float pan_x, pan_y, cos_r, sin_r, zoom, scale, cx, cy;
  int32_t width, height;
};

// Arguments of one iterate dispatch.
struct IterArgs {
  const float* flame;       // FlameHeader + rows
  int32_t n_rows;           // table rows after the header (sizes the LDS copy)
  const uint32_t* pal_u32;  // packed splat colour per texel: r | g<<8 | b<<16
  const float* pal_f32;     // linear colour per texel x3 (f32 splat mode)
  int32_t pal_size;
  ViewConst view;
  uint64_t walkers;
  uint32_t walker0;         // first walker id (batched launches)
  uint32_t ipt, burn_in, seed;
  float4* orbits;           // x, y, colour index, fuse bits (with xaos: fuse | previous
                            // transform << 24); may be null
  int32_t fresh;
  uint32_t refresh_period;  // walkers with (walker + seed) % period == 0 restart (0: none)
  int32_t ext;              // the flame has xaos or a variation beyond the workloads' thirteen
  void* hist;               // u32 or f32 words
  uint32_t hist_copy_words; // XCD mode: words per private copy
  unsigned long long* plotted;
  unsigned long long* inframe;
};

// Dynamic LDS a kernel may request: gfx950 has 160 KB per workgroup; 8 KB
// stay for the kernels' static __shared__ variables.
constexpr int LDS_DYNAMIC_MAX = 152 * 1024;

// Screen tiles of the LDS splat: 2^tile_wl x 2^tile_hl pixels (TileArgs),
// at most TILE_PX_MAX (a record's tile-local pixel has 13 bits). One flame
// renders with 128x64 tiles = 128 KB of u32x4 in LDS; the batched gallery
// tunes between those and 64x64 (64 KB: two accumulating workgroups a CU).
constexpr int TILE_W = 128, TILE_H = 64, TILE_PX = TILE_W * TILE_H;
constexpr int TILE_WL = 7, TILE_HL = 6;
constexpr int TILE_PX_MAX = 8192;
constexpr int TILE_MAX = 2048;  // bins per image the LDS counting sort holds

// Records are sorted into bins: a bin is 2^group_shift consecutive tiles of
// one tile row (one tile per bin while the image has at most TILE_MAX
// tiles; larger images group tiles so the bins stay within TILE_MAX).
// A record is tile-local pixel | texel << 13 | tile-in-bin << 25 (texels
// below 4096, at most 128 tiles per bin).
constexpr int REC_TEXEL_SHIFT = 13, REC_SUB_SHIFT = 25;
constexpr uint32_t REC_LOCAL_MASK = 0x1FFFu, REC_TEXEL_MASK = 0xFFFu;

struct TileArgs {
  uint32_t* tile_fill;   // n_tiles: records requested per bin
  uint64_t* region;      // n_tiles: first slot of the bin's region in `records` (up to 2^33)
  uint32_t* cap;         // n_tiles: slots in the region
  uint32_t* item_start;  // n_tiles + 1: work list of tile_accum
  uint32_t* records;     // local | texel << 13 | tile-in-bin << 25
  uint32_t* sorted;      // bins of several tiles: records re-ordered by tile, same offsets
  uint32_t* hist;        // the u32x4 (or f32x4) histogram
  const float* pal_f32;  // f32 splat: linear colour x 100 per texel, x3
  int32_t tiles_x;       // tiles per tile row
  int32_t n_tiles;       // bins (the length of the per-bin arrays)
  int32_t bins_x;        // bins per tile row: ceil(tiles_x / 2^group_shift)
  int32_t group_shift;   // log2 of the tiles per bin
  int32_t width, height;
  int32_t f32;           // 1: f32 histogram (float4 accumulation)
  int32_t flush_every;   // iterations between staging-fill checks (1..8)
  int32_t stage;         // staged records per walker workgroup (bin_shape)
  int32_t walk_block;    // lanes per walker workgroup (bin_shape): 512 or 1024
  int32_t walk_per_cu;   // walker workgroups per CU (bin_shape): 3 or 2 of 512, 1 of 1024
  int32_t bins_lds;      // bin slots of the walkers' LDS counting sort: >= n_tiles, a multiple of 64
  int32_t tile_wl;       // log2 of the tile width (pixels)
  int32_t tile_hl;       // log2 of the tile height
};

// The walker launch for a flame's LDS needs: lanes per workgroup and
// staged records per workgroup.
void bin_shape(int n_rows, int pal_size, int nb, int* block, int* stage, int* per_cu,
               int max_per_cu);
size_t bin_lds_bytes(int n_rows, int pal_size, int stage, int nb);
size_t accum_lds_bytes(int pal_size, bool f32, int tile_px);
hipError_t launch_bin_iterate(const IterArgs& a, const TileArgs& t, bool count_only, int grid,
                              hipStream_t s);
hipError_t launch_tile_capacity(const TileArgs& t, float scale, uint32_t margin,
                                uint64_t buffer_records, hipStream_t s);
hipError_t launch_tile_accum(const TileArgs& t, const uint32_t* pal, int pal_size, int grid,
                             hipStream_t s);

// Kernel launchers (defined in the .hip files).
hipError_t launch_iterate(const IterArgs& a, int splat_mode, int grid, int block,
                          hipStream_t s);
hipError_t launch_trace(const IterArgs& a, uint32_t walker, uint32_t steps, uint32_t* out,
                        hipStream_t s);
hipError_t launch_fold_xcd(uint32_t* hist, const uint32_t* copies, size_t words, int ncopies,
                           hipStream_t s);
hipError_t launch_clear_u32(void* p, size_t words, hipStream_t s);
// The per-device scratch of ffm_iterate_noalloc (noalloc.hip).
hipError_t noalloc_scratch_init(int device);
void* noalloc_scratch(int device);
hipError_t launch_accumulate(const void* hist, int hist_is_f32, const float* prev, float* out,
                             uint32_t* count, int W, int H, int mode, float blend,
                             hipStream_t s);
hipError_t launch_blur(uint32_t* hist, uint32_t* scratch, int W, int H, float sigma,
                       float density_sigma, hipStream_t s);
hipError_t launch_density_estimate(const float* in, float* out, int W, int H,
                                   const float* filt_h2, const float* filt_inv_norm, int n_filt,
                                   int max_R, float curve, hipStream_t s);

struct TonemapConst {
  int32_t mode, highlight, levels_enabled;
  float k1, k2, gamma, funcval, gamma_threshold, vib, notvib, white, exposure;
  float density_scale, saturation, hue_shift, alpha_low, alpha_high;
  float levels_low, levels_high, levels_gamma, sample_density;
  float bg_lin[3];
  float inv22;
  int32_t transparent;  // transparent_mode: 0 opaque, 1 straight, 2 premultiplied
  int32_t use_curve;    // map clipped colours through curve_lut
  float curve_lut[256]; // the tone curve at i/255, each rounded to f16
};
hipError_t launch_downsample2x(const uint8_t* in, uint8_t* out, int ow, int oh, hipStream_t s);
hipError_t launch_tonemap(const float* accum, uint8_t* rgba, int W, int H,
                          const TonemapConst& tc, hipStream_t s);

// ---- the tail's effects (effects.hip)
struct FxConst {
  int32_t type;                       // ffm_effect_type
  float p[FFM_MAX_EFFECT_PARAMS];     // the effect's parameters
  int32_t W, H;
  float density_mean, pixel_scale;
};
hipError_t launch_density_effect(const FxConst& k, const float* in, float* out, hipStream_t s);
hipError_t launch_color_effect(const FxConst& k, const uint8_t* in, uint8_t* out, hipStream_t s);

// bloom (assets/effects/bloom.rhai): a down pyramid of the bright parts, an
// up pyramid of tent-filtered sums, and the composite. Level l of each is
// (lw[l], lh[l]) float4 texels.
constexpr int BLOOM_LEVELS = 7;
struct BloomArgs {
  const uint8_t* in;  // the picture (RGBA8), W x H
  uint8_t* out;
  int32_t W, H, levels;
  int32_t lw[BLOOM_LEVELS], lh[BLOOM_LEVELS];
  float* down[BLOOM_LEVELS];
  float* up[BLOOM_LEVELS];
  float threshold, softness, intensity, spread;
};
// Pyramid level sizes for a W x H frame; returns the level count.
int bloom_levels(int W, int H, int32_t* lw, int32_t* lh);
hipError_t launch_bloom(const BloomArgs& b, hipStream_t s);

// distance_field (assets/effects/distance_field.rhai): seed, jump flood
// (log2 passes, ping-pong a / b of nearest-seed coordinates), shade.
struct DistArgs {
  const uint8_t* in;
  uint8_t* out;
  float* a;  // W x H float2
  float* b;
  int32_t W, H;
  float threshold, spread, mode, intensity, band_count;
};
hipError_t launch_distance_field(const DistArgs& d, hipStream_t s);

// ---- batched gallery: many flames in one persistent launch

struct BatchFlame {
  uint32_t table_off;  // floats into the concatenated flame tables
  uint32_t pal_off;    // texels into the concatenated palettes
  int32_t n_rows, pal_size;
  ViewConst view;
  uint32_t ipt, burn_in, seed;
  uint32_t tile_base;  // first of this flame's tiles in the global tile arrays
  int32_t tiles_x, n_tiles;
  uint64_t hist_off;   // pixels into the concatenated histograms
};

struct BatchItem {
  uint32_t flame, w0, w1, pad;  // walkers [w0, w1) of one flame
};

struct BatchArgs {
  const BatchFlame* flames;
  int32_t n_flames;
  const BatchItem* items;
  int32_t n_items;
  const float* tables;
  const uint32_t* pal_u32;
  int32_t max_rows, max_pal;
  uint32_t count_div;  // pre-pass walks 1/count_div of every item
  uint32_t count_iter_div;  // for burn_in + 1/count_iter_div of the plotted iterations
  int32_t ext;         // some flame uses a variation beyond the workloads' thirteen
  uint32_t* next_item; // work-stealing counter of this launch (zeroed before it)
  TileArgs g;          // global tile arrays; g.hist = base of the histograms
};

struct BatchPix {
  uint64_t pix_off;  // pixels into the concatenated histograms / images
  uint64_t npix;
  TonemapConst tc;
};

hipError_t launch_bin_batch(const BatchArgs& b, bool count_only, int grid, hipStream_t s);
hipError_t launch_tile_capacity_n(const TileArgs& t, float scale, uint32_t margin,
                                  uint64_t buffer_records, hipStream_t s);
hipError_t launch_tile_accum_batch(const BatchArgs& b, int grid, hipStream_t s);
hipError_t launch_batch_finish(const uint32_t* hist, uint8_t* const* out, const BatchPix* fl,
                               int n, hipStream_t s);

}  // namespace ffm
