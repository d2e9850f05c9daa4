/*
 * ffm.h -- C ABI of the MI355X fractal-flame chaos-game engine (libffm.so).
 *
 * The path this library replaces is the reference's flame render:
 * iterate -> splat -> (blur | density estimate) -> accumulate -> tonemap.
 * Every entry point below names the reference interface it stands in for.
 * Reference files are cited relative to the calibas/fflame-rust tree.
 *
 * Conventions (mirroring the reference's render API and pyfae's C ABI):
 *  - Every call returns FFM_OK (0) or a negative ffm_status; the message of
 *    the last failure on a context is ffm_last_error(ctx)
 *    (src/renderer/render.rs:312-330 RenderError; wasm/python/src/lib.rs:80-101
 *    returns status + error text the same way).
 *  - Descriptors are caller-owned and copied on set.
 *  - Device buffers passed to the granular calls are plain device pointers
 *    (hipMalloc'd or a torch tensor's data_ptr()); `stream` is a hipStream_t
 *    passed as void*, NULL meaning the context's own stream. That stream is
 *    a blocking stream: its work is ordered after and before work on the
 *    legacy default stream (torch's default stream), so buffers filled there
 *    are ready when a NULL-stream call runs.
 *  - One context per host thread; calls on one context are serialised.
 *  - Plain C: no torch or HIP types appear in any signature.
 */
#ifndef FFM_H
#define FFM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFM_ABI_VERSION 10

/* Limits. The reference caps a flame at 128 transforms
 * (src/gpu/buffers.rs:28 MAX_TRANSFORMS) and a palette texture at 256..4096
 * texels (src/config/defaults.rs:30, src/gpu/buffers.rs:1511). */
#define FFM_MAX_XFORMS 128
#define FFM_MAX_FINALS 4
#define FFM_MAX_VARS 16       /* distinct variations active in one flame */
#define FFM_MAX_VAR_PARAMS 4  /* user params + derived slots per variation */
#define FFM_MAX_XAOS 32       /* transforms of a flame with xaos weights */
#define FFM_MIN_PALETTE 256
#define FFM_MAX_PALETTE 4096

/* Status codes: RenderError::{OutOfMemory, TooLarge, Cancelled,
 * PixelReadFailed, NoPaletteFound} (src/renderer/render.rs:312-330) plus
 * argument and device errors. */
typedef enum {
  FFM_OK = 0,
  FFM_ERR_INVALID_ARGUMENT = -1,
  FFM_ERR_OUT_OF_MEMORY = -2,
  FFM_ERR_TOO_LARGE = -3,
  FFM_ERR_CANCELLED = -4,
  FFM_ERR_PIXEL_READ_FAILED = -5,
  FFM_ERR_NO_PALETTE = -6,
  FFM_ERR_DEVICE = -7,
  FFM_ERR_NO_FLAME = -8,
  FFM_ERR_PARSE = -9,       /* a config that is not valid JSON / XML, or misses a required field */
  FFM_ERR_UNSUPPORTED = -10 /* a config feature this path does not render */
} ffm_status;

/* Variation ids are the reference registry indices, which are append-only
 * "for preset compatibility" (src/variations/defs/mod.rs:442-501). */
typedef enum {
  FFM_VAR_LINEAR = 0,
  FFM_VAR_SINUSOIDAL = 1,
  FFM_VAR_SPHERICAL = 2,
  FFM_VAR_SWIRL = 3,
  FFM_VAR_HORSESHOE = 4,
  FFM_VAR_POLAR = 5,
  FFM_VAR_DISC = 8,
  FFM_VAR_HYPERBOLIC = 10,
  FFM_VAR_JULIA = 13,
  FFM_VAR_WAVES = 15,
  FFM_VAR_JULIAN = 24, /* params: [0]=power, [1]=dist; [2]=cpower derived */
  FFM_VAR_BLOB = 25,   /* params: [0]=high, [1]=low, [2]=waves */
  FFM_VAR_BUBBLE = 27,
  FFM_VAR_NOISE = 29,
  FFM_VAR_BLUR = 30,
  FFM_VAR_GAUSSIAN_BLUR = 31,
  FFM_VAR_CURL = 47,   /* params: [0]=c1, [1]=c2 */
  FFM_VAR_PRE_BLUR = 54 /* pre phase: moves the affine point before the others */
} ffm_variation_id;

/* One transform. Field meaning follows the reference Transform / GPU
 * Transform (src/scene/transforms.rs:337-482; shaders/core/header.wgsl:5-65):
 * x' = a*x + b*y + e, y' = c*x + d*y + f (shaders/core/affine.wgsl:4-19).
 * `var_weight[k]` is the weight of the flame's k-th active variation
 * (ffm_flame_desc.var_id[k]); 0 leaves it out, as the reference's
 * `if (xform.variations[k] != 0.0)` gate does
 * (src/shader_builder_v2.rs:2742-2758). */
typedef struct {
  float a, b, c, d, e, f;
  float post_a, post_b, post_c, post_d, post_e, post_f;
  int32_t post_enabled; /* reference: post_enabled > 0.5 */
  float weight;         /* selection weight (ignored for finals) */
  float color;          /* palette position of this transform */
  float color_speed;    /* Apophysis "symmetry" s (shaders/core/main_template.wgsl:496-503) */
  float opacity;
  float var_weight[FFM_MAX_VARS];
  float var_param[FFM_MAX_VARS][FFM_MAX_VAR_PARAMS];
  /* The final transforms this normal transform applies, in order, as
   * indices into ffm_flame_desc.finals, read when the flame's
   * per_transform_finals is set (Transform::final_attachments; the
   * reference's AttachmentList, shaders/core/header.wgsl:173-178 and
   * src/gpu/buffers.rs:596-650). */
  int32_t n_final_attach;
  int32_t final_attach[FFM_MAX_FINALS];
} ffm_xform;

/* A flame: normal transforms, the final chain, and the flame's active
 * variation list in emission (local index) order
 * (src/scene/transforms.rs:30-85 compute_local_index_map). */
typedef struct {
  int32_t n_xforms;
  int32_t n_finals;
  int32_t n_vars;
  int32_t var_id[FFM_MAX_VARS];
  ffm_xform xforms[FFM_MAX_XFORMS];
  ffm_xform finals[FFM_MAX_FINALS];
  /* 0: every normal transform applies every final, in order -- a single
   * legacy final (migrate_legacy_final, src/scene/transforms.rs:2257-2266)
   * and .flame imports; 1: each applies its own final_attach list. */
  int32_t per_transform_finals;
  /* Xaos (Flame::xaos, src/scene/transforms.rs:2040-2044): when has_xaos,
   * the transform after transform i is drawn with weights
   * weight[j] * xaos[i][j] (select_transform_xaos,
   * shaders/core/utilities.wgsl:107-140), for flames of at most
   * FFM_MAX_XAOS transforms. */
  int32_t has_xaos;
  float xaos[FFM_MAX_XAOS][FFM_MAX_XAOS];
} ffm_flame_desc;

/* The view: FractalConfig zoom/pan/rotation as world_to_pixel reads them
 * (shaders/core/utilities.wgsl:417-445). */
typedef struct {
  int32_t width, height;
  float zoom, pan_x, pan_y, rotation;
} ffm_view;

/* Histogram / splat modes. The reference splats 4 x u32 per in-frame
 * sample with global atomicAdd (shaders/core/main_template.wgsl:1253-1320). */
typedef enum {
  FFM_SPLAT_U32_GLOBAL = 0, /* 4 x u32 device-scope atomics, as the reference */
  FFM_SPLAT_U32_XCD = 1,    /* u32, one private copy per XCD, folded after */
  FFM_SPLAT_U32_LDS = 2,    /* u32, LDS-privatised screen tiles flushed with integer atomics */
  FFM_SPLAT_F32_GLOBAL = 3, /* 4 x f32 atomics ("float4 accumulation") */
  FFM_SPLAT_F32_LDS = 4     /* f32 through the LDS screen tiles (palette <= 2048 texels) */
} ffm_splat_mode;

#define FFM_MAX_CURVE_POINTS 16

/* Tonemap parameters: TonemapParams (shaders/tonemap.wgsl:9-63) as the
 * host fills it (src/renderer/compute_kernel.rs:3813-3918), log branch and
 * the modes around it. */
typedef struct {
  int32_t tonemap_mode;   /* 0 linear, 1 logarithmic (default), 2 density */
  int32_t highlight_mode; /* 0 clip, 1 maxnorm, 2 reinhard, 3 aces */
  float exposure, gamma, gamma_threshold, brightness, vibrancy, white_level;
  float density_scale, saturation, hue_shift;
  float alpha_blend_low, alpha_blend_high;
  int32_t levels_enabled;
  float levels_low, levels_high, levels_gamma;
  float background[3]; /* sRGB */
  /* transparent_mode of TonemapParams (shaders/tonemap.wgsl:531-576):
   * 0 opaque over the background; 1 transparent, straight alpha split so a
   * flatten over black (rgb x a) gives the over-black picture; 2
   * transparent, premultiplied (rgb = the over-black colour, a = opacity) */
  int32_t transparent_mode;
  /* The tone curve (ToneCurve, src/scene/tonemap.rs:142-300; applied at
   * shaders/tonemap.wgsl:479-492): with use_curve set, every pixel whose
   * bucket count exceeds 0.001 maps each clipped channel c through a
   * 256-texel f16 table of the curve at i/255, read by linear filtering at
   * texel-space position 255c (the reference's linear clamp-to-edge curve
   * sampler, src/gpu/buffers.rs:1877-1886). The curve is the monotone
   * cubic Hermite through curve_n
   * points (2..FFM_MAX_CURVE_POINTS, x non-decreasing, finite);
   * curve_n = 0 with use_curve set is the linear curve (0,0)-(1,1). */
  int32_t use_curve;
  int32_t curve_n;
  float curve_x[FFM_MAX_CURVE_POINTS], curve_y[FFM_MAX_CURVE_POINTS];
} ffm_tonemap_params;

/* A render job: RenderJob (src/renderer/render.rs:111-138) plus the knobs
 * the loop reads. The loop dispatches `walkers` walkers of `ipt` iterations
 * until `target_samples` samples have been plotted (0: one dispatch),
 * accumulating every `batch_size` dispatches (render.rs:350-351,624-710). */
typedef struct {
  ffm_view view;
  uint64_t walkers;
  uint32_t ipt, burn_in;
  uint32_t seed_base, seed_stride; /* seed_k = base + k*stride (compute_kernel.rs:3541-3554) */
  uint64_t target_samples;
  uint32_t batch_size;
  int32_t splat_mode;      /* ffm_splat_mode */
  int32_t persistent_orbits; /* 1: walkers resume across dispatches */
  float filter_radius;     /* histogram_blur sigma in px (0: off) */
  float filter_blur_edges; /* bilateral edge slider, 0..1 (compute_kernel.rs:1668-1700) */
  float de_max_radius, de_min_radius, de_curve; /* flam3 density estimation (0 max: off) */
  ffm_tonemap_params tonemap;
} ffm_job;

typedef struct {
  uint64_t total_samples;   /* plotted samples (RenderOutput::total_iterations) */
  uint64_t dispatches;
  double render_time_ms;
  double iterate_ms;        /* GPU time of the dispatch phase (walkers + splat) */
  double walker_kernel_ms;  /* summed durations of the walker-kernel launches */
  uint64_t walker_launches; /* number of walker-kernel launches (pre-pass excluded) */
  double tile_accum_ms;     /* summed durations of the tile-accumulate launches */
} ffm_render_stats;

typedef struct ffm_ctx ffm_ctx;

/* ---- context ---------------------------------------------------------- */

/* FlameRenderer::for_job (src/renderer/compute_kernel.rs:784): one renderer
 * per device, owning its stream and device buffers. */
int ffm_ctx_create(int device, ffm_ctx** out);
int ffm_ctx_destroy(ffm_ctx* ctx);
/* pyfae's error text (wasm/python/src/lib.rs:80-101 status 1 + message). */
const char* ffm_last_error(const ffm_ctx* ctx);
int ffm_abi_version(void);

/* FlameRenderer::load_config (src/renderer/compute_kernel.rs:2748): packs
 * the transforms (GpuTransform, src/gpu/buffers.rs:94-160) and derives
 * variation init values (init_julian, src/variations/defs/advanced.rs:467). */
int ffm_flame_set(ffm_ctx* ctx, const ffm_flame_desc* flame);

/* Palette upload: stops (position, r, g, b) as sRGB floats, sampled into a
 * `table_size`-texel table with linear interpolation
 * (src/scene/palette.rs:265-360), stored as f16 (src/gpu/buffers.rs:3215-3221)
 * and read nearest (src/gpu/buffers.rs:1866-1875). */
int ffm_palette_set(ffm_ctx* ctx, const float* stops_prgb, int n_stops, int table_size);

/* ---- whole render ------------------------------------------------------ */

/* render / render_with (src/renderer/render.rs:371,483-800) ending in the
 * tonemap and pixel readback (src/renderer/tail.rs:112-114): RGBA8 sRGB
 * pixels into `rgba_out` (width*height*4 bytes, host memory). */
int ffm_render(ffm_ctx* ctx, const ffm_job* job, uint8_t* rgba_out, ffm_render_stats* stats);

/* The same render, leaving the accumulator (RGBA32F, row-major; after the
 * density estimator and the context's density effects: what the tone map
 * would read) in `accum_out` (host, width*height*4 floats) instead of
 * tonemapping. */
int ffm_render_accum(ffm_ctx* ctx, const ffm_job* job, float* accum_out, ffm_render_stats* stats);

/* The same render with the RGBA8 pixels left in device memory
 * (`rgba_dev`, width*height*4 bytes): no host readback, so a caller that
 * keeps working on the GPU (torch, a video encoder) pays no PCIe copy. */
int ffm_render_to_device(ffm_ctx* ctx, const ffm_job* job, void* rgba_dev,
                         ffm_render_stats* stats);

/* RenderProgress::is_cancelled (src/renderer/render.rs:293-302): polled
 * between dispatches; the render returns FFM_ERR_CANCELLED. */
int ffm_cancel(ffm_ctx* ctx);

/* ---- granular kernels on device buffers --------------------------------- */

/* One compute dispatch (compute_pass, src/renderer/compute_kernel.rs:1243;
 * kernel shaders/core/main_template.wgsl:20-1428): `walkers` walkers run
 * `ipt` iterations from rng_init(walker, seed). `orbits` (16 B per walker:
 * x, y, colour index, fuse) holds resumable walker state
 * (main_template.wgsl:125-181,1384-1416); NULL starts every walker afresh.
 * `fresh` != 0 (re)starts walkers whose state is in `orbits`.
 * `hist` is width*height*4 words (u32, or f32 for FFM_SPLAT_F32_GLOBAL).
 * `plotted` (device u64, may be NULL) gets the number of samples that
 * passed the fuse (in frame or not); `inframe` (device u64, may be NULL)
 * the number splatted. */
int ffm_iterate(ffm_ctx* ctx, const ffm_view* view, uint64_t walkers, uint32_t ipt,
                uint32_t burn_in, uint32_t seed, void* orbits, int fresh, void* hist,
                int splat_mode, uint64_t* plotted, uint64_t* inframe, void* stream);

/* Per-step trace of one walker for parity tests: for each of `steps`
 * iterations writes 8 words {rng state after the step, xform index,
 * bits(x), bits(y), bits(colour index), pixel x, pixel y, flags}
 * (flags: bit0 plotted in frame, bit1 respawned, bit2 burning in).
 * `out` is device memory of steps*8 u32. */
int ffm_trace(ffm_ctx* ctx, const ffm_view* view, uint32_t walker, uint32_t seed,
              uint32_t burn_in, uint32_t steps, uint32_t* out, void* stream);

/* Separable bilateral Gaussian on the batch histogram
 * (shaders/histogram_blur.wgsl:39-119, host compute_kernel.rs:1668-1727):
 * H pass hist -> scratch, V pass scratch -> hist in place; scratch holds
 * width x height x 4 u32 words. */
int ffm_blur(ffm_ctx* ctx, uint32_t* hist, uint32_t* scratch, int width, int height,
             float sigma_px, float density_sigma, void* stream);

/* flam3-style density estimation on the accumulator (no reference
 * counterpart: src/flame_xml.rs:1488-1490 keeps the estimator attributes
 * inert). Each pixel's colour sums and count spread over a Gaussian disc
 * whose radius shrinks with the pixel's count as max_radius/(count)^curve,
 * floored at min_radius (the Draves-Reckase estimator, flam3's filter
 * table). In/out are RGBA32F accumulators (mean colour, count). */
int ffm_density_estimate(ffm_ctx* ctx, const float* accum_in, float* accum_out, int width,
                         int height, float max_radius, float min_radius, float curve,
                         void* stream);

/* accumulate.wgsl (shaders/accumulate.wgsl:115-230): batch histogram into
 * the RGBA32F running accumulator. mode 0 cumulative mean, 1 fixed EMA at
 * blend_factor, 2 overwrite. `count` (u32x2 per pixel) non-NULL enables the
 * exact 64-bit density count. `hist_is_f32` reads an f32 histogram. */
int ffm_accumulate(ffm_ctx* ctx, const void* hist, int hist_is_f32, const float* prev,
                   float* out, uint32_t* count, int width, int height, int mode,
                   float blend_factor, void* stream);

/* tonemap.wgsl fs_main (shaders/tonemap.wgsl:224-590) as a compute
 * stencil: accumulator -> RGBA8 sRGB. `samples_in_buffer` feeds
 * sample_density (compute_kernel.rs:1796-1841); area per update_tonemap
 * (compute_kernel.rs:3813-3918). */
int ffm_tonemap(ffm_ctx* ctx, const float* accum, uint8_t* rgba, int width, int height,
                float zoom, double samples_in_buffer, const ffm_tonemap_params* tp,
                void* stream);

/* ffm_iterate without allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that forbid hipMalloc). The same dispatch as
 * ffm_iterate for splat modes U32_GLOBAL, U32_LDS, F32_GLOBAL and F32_LDS
 * (U32_XCD's per-XCD histograms do not fit; refused), with the LDS tile
 * splat's bin tables and staged records in a device-resident scratch of
 * FFM_NOALLOC_SCRATCH_BYTES per GPU instead of context buffers: walkers run
 * in chunks whose records fit it, the bin plan is made afresh each call, no
 * timing events are recorded, and nothing is allocated or freed. Refused
 * with FFM_ERR_TOO_LARGE when one walker's ipt + 1 records exceed three
 * quarters of the record space (the whole scratch for single-tile bins,
 * half of it for images past 2048 tiles). Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU race, so issue them one at a
 * time per device. */
#define FFM_NOALLOC_SCRATCH_BYTES (256u << 20)
int ffm_iterate_noalloc(ffm_ctx* ctx, const ffm_view* view, uint64_t walkers, uint32_t ipt,
                        uint32_t burn_in, uint32_t seed, void* orbits, int fresh, void* hist,
                        int splat_mode, uint64_t* plotted, uint64_t* inframe, void* stream);

/* downsample_2x_firefly (src/export/supersample.rs:117-187): the RGBA8
 * picture `in` of (2 out_w) x (2 out_h) boxed 2x2 down into `out` (out_w x
 * out_h), each quad's samples brighter (r + 2g + b) than 4x the quad's
 * median + 64 first scaled down to that, alpha averaged unclamped. Device
 * pointers. */
int ffm_downsample2x(ffm_ctx* ctx, const uint8_t* in, uint8_t* out, int out_w, int out_h,
                     void* stream);

/* Fold the per-XCD private histograms of FFM_SPLAT_U32_XCD into `hist`. */
int ffm_fold_xcd(ffm_ctx* ctx, uint32_t* hist, int width, int height, void* stream);

/* ---- batched gallery ------------------------------------------------------ */

/* Many independent flames in one persistent launch (no reference
 * counterpart: the gallery renders in series, wasm/render/src/lib.rs:45-53).
 * Flame i renders jobs[i] (its own view, walkers, ipt, seed) with palette
 * i (stops_prgb[i], n_stops[i]); its RGBA8 pixels go to rgba_out[i]. */
int ffm_batch_render(ffm_ctx* ctx, int n, const ffm_flame_desc* flames,
                     const float* const* stops_prgb, const int* n_stops,
                     const ffm_job* jobs, uint8_t* const* rgba_out, ffm_render_stats* stats);

/* The same flames launched one after another (the reference's behaviour),
 * for the batched/serial comparison. */
int ffm_batch_render_serial(ffm_ctx* ctx, int n, const ffm_flame_desc* flames,
                            const float* const* stops_prgb, const int* n_stops,
                            const ffm_job* jobs, uint8_t* const* rgba_out,
                            ffm_render_stats* stats);

/* ffm_batch_render and ffm_batch_render_serial with the pictures left in
 * device memory: rgba_dev[i] is a device buffer of flame i's width x height
 * RGBA8 pixels, as ffm_render_to_device gives one render's (no reference
 * counterpart: the reference's gallery reads every picture back). */
int ffm_batch_render_to_device(ffm_ctx* ctx, int n, const ffm_flame_desc* flames,
                               const float* const* stops_prgb, const int* n_stops,
                               const ffm_job* jobs, uint8_t* const* rgba_dev,
                               ffm_render_stats* stats);
int ffm_batch_render_serial_to_device(ffm_ctx* ctx, int n, const ffm_flame_desc* flames,
                                      const float* const* stops_prgb, const int* n_stops,
                                      const ffm_job* jobs, uint8_t* const* rgba_dev,
                                      ffm_render_stats* stats);

/* ---- differentiable soft splat ---------------------------------------------- */

/* Bilinear soft splat of the chaos game into an f32 density image, and its
 * adjoint (no reference counterpart: splat_size is hard-wired to 1.0 and
 * read by nothing, src/renderer/compute_kernel.rs:1354). The walk is the
 * chaos game's (selection, opacity draw, transform, respawn, fuse); each
 * plotted sample adds density_scale x its bilinear weights to the four
 * pixels around its continuous position (world_to_pixel without the floor,
 * minus half a pixel). Colour does not enter the density.
 *
 * The differentiable parameters of a flame with n transforms and v active
 * variations are theta = [n x (a, b, c, d, e, f)] then [n x v weights]
 * (the transforms' pre-affine coefficients and variation weights);
 * everything else (selection weights, opacity, post-affine, variation
 * params) comes from the flame descriptor. Several independent fits run in
 * one launch: theta, density and grad arrays hold the fits back to back in
 * the order given to ffm_softsplat_set. Soft-splat flames have no final
 * transforms, at most FFM_SOFT_MAX_XFORMS transforms and
 * FFM_SOFT_MAX_VARS variations, and no waves variation (its body reads the
 * affine coefficients themselves). */
#define FFM_SOFT_MAX_XFORMS 8
#define FFM_SOFT_MAX_VARS 4
#define FFM_SOFT_MAX_WINDOW 16

typedef struct ffm_soft_job {
  ffm_view view;        /* density image size and camera */
  uint64_t walkers;     /* 1 .. 2^32 */
  uint32_t ipt;         /* iterations per walker */
  uint32_t burn_in;     /* fuse iterations before a walker plots */
  uint32_t seed;        /* rng_init seed of every walker */
  uint32_t window;      /* adjoint replay window K, 1 .. FFM_SOFT_MAX_WINDOW */
  uint32_t truncate;    /* 0: exact gradient (dL/dpoint carried across windows);
                           1: truncated, dL/dpoint restarts at zero each window */
  float density_scale;  /* density = density_scale x sum of bilinear weights */
  uint32_t walk_f64;    /* 0: the walk in f32, the chaos game's own arithmetic;
                           1: in f64 (same transform picks and draws), for
                           gradients that f32 rounding of long walks would blur */
} ffm_soft_job;

/* Number of differentiable parameters of `flame` (n * (6 + v)), or a
 * negative status when the flame cannot be fitted. */
int ffm_softsplat_param_count(const ffm_flame_desc* flame);

/* theta of `flame` as the descriptor holds it (host floats). */
int ffm_softsplat_pack(const ffm_flame_desc* flame, float* theta_out);

/* Structure of n fits (copied): flames[i] is fitted on jobs[i]. */
int ffm_softsplat_set(ffm_ctx* ctx, int n, const ffm_flame_desc* flames,
                      const ffm_soft_job* jobs);

/* Total floats of theta (and grad_theta) and of density (and grad_density)
 * for the fits set last. */
int ffm_softsplat_sizes(const ffm_ctx* ctx, uint64_t* theta_floats, uint64_t* density_floats);

/* density (device f32) := the soft splat of every fit at theta (device f32). */
int ffm_softsplat_forward(ffm_ctx* ctx, const float* theta, float* density, void* stream);

/* grad_theta (device f32, overwritten) := sum over plotted samples of
 * grad_density (device f32, dL/d density) pulled back through the splat
 * and the walk. Each walker's trajectory is recomputed by RNG replay in
 * windows of K iterations from 16-byte checkpoints; nothing per iteration
 * is stored between calls. Sums are reduced in f64 in a fixed order, so a
 * call is deterministic. */
int ffm_softsplat_adjoint(ffm_ctx* ctx, const float* theta, const float* grad_density,
                          float* grad_theta, void* stream);

/* ---- configs: .fflame / .fae JSON and .flame XML -------------------------- */

/* The reference's two on-disk forms, read into what this path renders:
 * FractalConfig JSON (src/config/fractal_config.rs:178, migrated by version
 * as FractalConfig::from_json does, :1746-2050; Transform and Flame serde,
 * src/scene/transforms.rs:1307-1560,2567-2690) and Apophysis / JWildfire
 * XML (parse_flame_xml, src/flame_xml.rs:41-560,600-1067). A config that
 * asks for something this path does not render (a variation outside
 * ffm_variation, linked transforms, subflames, post symmetry, a tone
 * curve beyond FFM_MAX_CURVE_POINTS or with unsorted x, user-written
 * effects, escape-time or simulation modes, a 3D
 * camera) is refused with FFM_ERR_UNSUPPORTED and a message naming it,
 * never rendered approximately. */
#define FFM_MAX_STOPS 1024
#define FFM_SCENE_NAME 128

/* PaletteTransform (src/scene/palette.rs:511-600): the palette lookup table
 * is squeezed, log-redistributed, rotated and reversed, in that order
 * (transform_lookup_table, palette.rs:703-800). All zero but
 * squeeze_factor = 1 is the identity. */
typedef struct {
  int32_t squeeze_mode;  /* 0 linear, 1 geometric */
  float squeeze_factor;  /* linear: repeat count; geometric: octave ratio r */
  float log_strength;    /* exponential redistribution; 0 = off */
  float rotation;        /* cyclic shift, fraction of the table (-1..1) */
  int32_t reverse;
} ffm_palette_xf;

/* ---- the tail: density and colour effects ----------------------------
 * The stages after the chaos game (src/renderer/tail.rs:1-162): density
 * effects on the accumulator (f32 RGBA: mean colour, hit count) before the
 * tone map, colour effects on the tone-mapped RGBA8 picture after it, each
 * effect reading the previous one's output (EffectChainRunner::
 * run_density_effects / run_color_effects, src/renderer/effect_chain.rs:
 * 737-830,874-962). The built-in effects below follow their Rhai
 * definitions (assets/effects/<name>.rhai); a colour effect's colour is
 * blended over its input by the blend library
 * (shaders/effects/common/blend_modes.wgsl apply_blend; the input's alpha
 * kept, src/effects/custom.rs:49-60) and stored as RGBA8. */
#define FFM_MAX_EFFECTS 16
#define FFM_MAX_EFFECT_PARAMS 12
enum ffm_effect_type {
  FFM_FX_DENSITY_BLUR = 1,           /* radius 3, threshold 0.25, falloff 0.5 */
  FFM_FX_BILATERAL_BLUR = 2,         /* radius 3, sigma_spatial 3, sigma_range 0.1 */
  FFM_FX_SHARPEN = 3,                /* amount 1, radius 1 */
  FFM_FX_VIGNETTE = 16,              /* intensity 0.5, radius 0.5, softness 0.5, blend_mode 0 */
  FFM_FX_FILM_GRAIN = 17,            /* intensity 0.1, seed 0, blend_mode 4 */
  FFM_FX_CHROMATIC_ABERRATION = 18,  /* amount 2, radial 1, intensity 1, blend_mode 0 */
  FFM_FX_HUE_SHIFT = 19,             /* offset 0 (degrees), intensity 1, blend_mode 0 */
  FFM_FX_SOBEL_EDGES = 20,           /* intensity 1, threshold 0.1, glow 0.5, preserve_color 0.5,
                                        blend_mode 1 */
  FFM_FX_KALEIDOSCOPE = 21,          /* segments 6, rotation 0 (degrees), zoom 1, intensity 1,
                                        blend_mode 0, square_mode 1, edge_offset 0 */
  FFM_FX_PLASMA = 22,                /* intensity 0.3, scale 3, speed 1, time 0, blend_mode 1,
                                        direction 225 (degrees) */
  FFM_FX_TUNNEL = 23,                /* speed 1, rotation_speed 0.5, distortion 0, time 0,
                                        intensity 1, blend_mode 0 */
  FFM_FX_JULIA = 24,                 /* mode 1 (0 Mandelbrot, 1 Julia), julia_cx -0.7, julia_cy 0.27,
                                        max_iter 100 (at most 5000), zoom 1, center_x 0, center_y 0,
                                        color_mode 1, intensity 1, blend_mode 0 */
  FFM_FX_WORLEY_NOISE = 25,          /* intensity 0.5, scale 8, time 0, mode 0, color_mode 1,
                                        blend_mode 4, direction 225 (degrees) */
  FFM_FX_SIMPLEX_NOISE = 26,         /* intensity 0.5, scale 5, octaves 4, time 0, mode 0
                                        (colour, distort, mask), blend_mode 4, direction 225 */
  FFM_FX_DOMAIN_WARP = 27,           /* intensity 0.1, scale 3, octaves 4, time 0, blend_mode 0,
                                        direction 225 (degrees) */
  FFM_FX_BLOOM = 28,                 /* threshold 0.45, softness 0.3, intensity 1.5, spread 0.7
                                        (tint white); a 7-level half-size pyramid */
  FFM_FX_DISTANCE_FIELD = 29,        /* threshold 0.15, spread 0.25, mode 0 (glow, contours, fill),
                                        intensity 1, band_count 12; a jump flood */
  FFM_FX_TRAILS = 30                 /* persistence 0.85, intensity 1: a temporal effect; a still
                                        render has no history, so the picture passes unchanged */
};
/* An EffectInstance (src/effects/mod.rs:319-336): params in the effect's
 * declared order (the slot order of get_param), as listed above. */
typedef struct {
  int32_t type;     /* ffm_effect_type */
  int32_t enabled;  /* 0: skipped */
  float params[FFM_MAX_EFFECT_PARAMS];
} ffm_effect;
/* A config's two chains (FractalConfig density_effects / color_effects,
 * src/config/fractal_config.rs:543-548). */
typedef struct {
  int32_t n_density, n_color;
  ffm_effect density[FFM_MAX_EFFECTS];
  ffm_effect color[FFM_MAX_EFFECTS];
  float pixel_scale;  /* render pixels per logical point (set_pixel_scale; 1 for exports) */
} ffm_effects;

/* The registry's defaults for `type` (the Rhai PARAMS values). */
int ffm_effect_defaults(int32_t type, ffm_effect* out);

/* Effects every later ffm_render / ffm_render_to_device / ffm_render_accum
 * of this context applies (NULL: none). ffm_render_accum returns the
 * accumulator after the density chain: what the tone map reads. The
 * batched gallery and the soft splat apply none. */
int ffm_effects_set(ffm_ctx* ctx, const ffm_effects* fx);

/* One density effect on device buffers (f32 RGBA, width*height*4 floats
 * each, in != out); density_mean is the frame's mean density
 * (FlameRenderer::mean_density, src/renderer/compute_kernel.rs:1786-1792). */
int ffm_density_effect(ffm_ctx* ctx, const ffm_effect* fx, const float* in, float* out, int width,
                       int height, float density_mean, float pixel_scale, void* stream);

/* One colour effect on device buffers (RGBA8, in != out). */
int ffm_color_effect(ffm_ctx* ctx, const ffm_effect* fx, const uint8_t* in, uint8_t* out,
                     int width, int height, float pixel_scale, void* stream);

/* One config as this path reads it. */
typedef struct {
  ffm_flame_desc flame;
  int32_t n_stops;
  float stops[FFM_MAX_STOPS][4]; /* position, r, g, b (sRGB 0..1), in file order */
  int32_t palette_size;          /* lookup-table texels, 256..4096 */
  ffm_palette_xf palette_xf;
  float zoom, rotation;
  double pan_x, pan_y;
  float filter_radius, filter_blur_edges;
  uint64_t max_iterations;       /* the config's sample budget */
  int32_t reset_orbits;          /* restart every orbit each dispatch */
  int32_t deterministic_rng;     /* fixed seed schedule (else seeded from the clock) */
  int32_t render_3d;             /* render mode 3d with a flat camera (every .flame import) */
  uint32_t image_width, image_height; /* the saved size (`image_size`, XML `size`) */
  ffm_tonemap_params tonemap;
  char name[FFM_SCENE_NAME];     /* the flame's name, truncated, NUL-terminated */
  ffm_effects effects;           /* the config's enabled density and colour effects */
} ffm_scene;

/* Number of configs in `text` (a JSON object, a JSON array of configs --
 * FractalConfig::from_json_multi, fractal_config.rs:1944-1973 -- or an XML
 * document's <flame> / <jwf-flame> elements). Errors go to err[err_cap]. */
int ffm_scene_count(const char* text, size_t len, int* count, char* err, size_t err_cap);

/* Parse config `index` of `text` into `out`. */
int ffm_scene_parse(const char* text, size_t len, int index, ffm_scene* out, char* err,
                    size_t err_cap);

/* ffm_palette_set with a lookup-table transform applied (NULL: none). */
int ffm_palette_set_xf(ffm_ctx* ctx, const float* stops_prgb, int n_stops, int table_size,
                       const ffm_palette_xf* xf);

/* ffm_scene_job / ffm_render_config flags. A scene's 3D render mode (a
 * flat camera) renders through the 2D chaos game only with
 * FFM_RENDER_3D_AS_2D. FFM_RENDER_TRANSPARENT and FFM_RENDER_PREMULTIPLIED
 * are pyfae render's transparent= / premultiplied= (transparent_mode 1 or
 * 2, src/renderer/compute_kernel.rs:3771-3784).
 * FFM_RENDER_SUPERSAMPLE_2X is the export's 2x antialiasing
 * (supersample_plan / scale_config_for_supersample,
 * src/export/supersample.rs:54-100): the job renders at twice the width and
 * height with 4x the sample budget and twice the histogram filter radius,
 * and ffm_render_config boxes its final picture down to the requested size
 * (ffm_downsample2x). By default a job dispatches the reference's 8192
 * walkers (128 workgroups of 64, src/renderer/render.rs:350-351): the same
 * walker streams, seeds and orbits as the reference's render, whose picture
 * depends on that count wherever the walk mixes slowly (the reference's own
 * visual config, an identity map, draws exactly its 8192 start points).
 * FFM_RENDER_FILL_GPU dispatches enough walkers to fill the GPU instead
 * (about half the budget in one dispatch, 8192 to 2^20): for a flame whose
 * walk mixes within the burn-in, the same picture within the Monte-Carlo
 * noise in a few dispatches. */
#define FFM_RENDER_3D_AS_2D 1u
#define FFM_RENDER_TRANSPARENT 2u
#define FFM_RENDER_PREMULTIPLIED 4u
#define FFM_RENDER_SUPERSAMPLE_2X 8u
#define FFM_RENDER_FILL_GPU 16u

/* The job the reference's export builds for a config (RenderJob::new,
 * src/renderer/render.rs:111-138, with pyfae's render arguments,
 * python/pyfae/__init__.py:201-221): `iterations` 0 takes the config's
 * max_iterations, `ipt` 0 the CLI's 1024. */
int ffm_scene_job(const ffm_scene* scene, int width, int height, uint64_t iterations,
                  uint32_t ipt, uint32_t flags, ffm_job* out, char* err, size_t err_cap);

/* pyfae's Engine.render(config, width, height, iterations=,
 * iterations_per_thread=) (python/pyfae/__init__.py:201-221; the engine's
 * "render" request): parse config 0 of `text`, set its flame and palette,
 * render with its effects (not the context's),
 * render, RGBA8 into rgba_out (host, width*height*4). */
int ffm_render_config(ffm_ctx* ctx, const char* text, size_t len, int width, int height,
                      uint64_t iterations, uint32_t ipt, uint32_t flags, uint8_t* rgba_out,
                      ffm_render_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* FFM_H */
