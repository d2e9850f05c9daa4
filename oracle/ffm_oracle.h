/*
 * ffm_oracle.h -- CPU restatement of the reference's flame path, used ONLY
 * as the checker: by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg. The product (libffm.so) never links, loads or calls it.
 *
 * Every function restates the reference's algorithm from its source text
 * (file:line cited at each definition); FP32 arithmetic is evaluated in the
 * order the reference's WGSL writes it, with FMA contraction off
 * (compiled -ffp-contract=off), so integer paths and +,-,*,/ chains are
 * exact restatements and transcendental calls use the host libm.
 */
#ifndef FFM_ORACLE_H
#define FFM_ORACLE_H

#include <stdint.h>

#include "../include/ffm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t state, inc;
} orc_rng;

typedef struct {
  float x, y, color_index;
  uint32_t fuse;
  uint32_t prev; /* with xaos: the transform applied last */
} orc_walker;

/* --- RNG (shaders/core/rng.wgsl:22-55) --- */
uint32_t orc_pcg_hash(uint32_t input);
void orc_rng_init(orc_rng* r, uint32_t thread_id, uint32_t seed);
uint32_t orc_rng_next(orc_rng* r);
float orc_rng_nextf(orc_rng* r);
/* Bulk: n consecutive words of walker `tid`'s stream (tests). */
void orc_rng_stream(uint32_t tid, uint32_t seed, uint32_t n, uint32_t* out);

/* --- transform selection (shaders/core/utilities.wgsl:35-55) --- */
int orc_select(const ffm_flame_desc* f, float rand_val);
/* select_transform_xaos (shaders/core/utilities.wgsl:107-140) */
int orc_select_xaos(const ffm_flame_desc* f, float rand_val, int prev);
/* Flame::xaos_walk_splits_orbits (src/scene/xaos.rs:64-124) as the orbit
 * refresh period (src/renderer/compute_kernel.rs:46-60): 4 or 0. */
uint32_t orc_xaos_refresh_period(const ffm_flame_desc* f);

/* --- one variation / one transform --- */
void orc_variation(int var_id, const float* params, const ffm_xform* xf, float x, float y,
                   orc_rng* rng, float* ox, float* oy);
void orc_apply_xform(const ffm_flame_desc* f, const ffm_xform* xf, float* x, float* y,
                     orc_rng* rng);
/* Fill derived variation slots (init_julian) in place. */
void orc_flame_prepare(ffm_flame_desc* f);

/* --- palette (src/scene/palette.rs:265-360; src/gpu/buffers.rs:3215-3221) --- */
void orc_palette_sample(const float* stops_prgb, int n_stops, float t, float out_rgb[3]);
uint16_t orc_f32_to_f16(float v);
float orc_f16_to_f32(uint16_t h);
/* Linear-light RGB per texel (f16 storage, then srgb_to_linear). */
void orc_palette_linear(const float* stops_prgb, int n_stops, int size, float* lin_rgb);
int orc_palette_texel(float color_index, int size);
/* The palette lookup-table transform (src/scene/palette.rs:703-800) as the
 * sampled texel each entry shows; NULL is the identity. */
void orc_palette_index_map(int size, const ffm_palette_xf* xf, int32_t* src);
void orc_palette_linear_xf(const float* stops_prgb, int n_stops, int size,
                           const ffm_palette_xf* xf, float* lin_rgb);

/* --- view (shaders/core/utilities.wgsl:417-445) --- */
void orc_world_to_pixel(const ffm_view* v, float x, float y, int32_t* px, int32_t* py);

/* --- the chaos game (shaders/core/main_template.wgsl) --- */
/* One dispatch of walkers [walker0, walker0+n). hist: W*H*4 u32 (NULL: no
 * splat). orbits: per-walker resumable state indexed by walker-walker0
 * (NULL: start fresh, state dropped); fresh!=0 restarts stored walkers.
 * Returns samples past the fuse; *inframe (may be NULL) the splatted ones. */
uint64_t orc_iterate(const ffm_flame_desc* f, const float* pal_lin, int pal_size,
                     const ffm_view* v, uint32_t walker0, uint32_t n, uint32_t ipt,
                     uint32_t burn_in, uint32_t seed, orc_walker* orbits, int fresh,
                     uint32_t refresh_period, uint32_t* hist, uint64_t* inframe);
/* Same, with several host threads each splatting into a private histogram
 * that is summed at the end (the CPU baseline). */
uint64_t orc_iterate_mt(const ffm_flame_desc* f, const float* pal_lin, int pal_size,
                        const ffm_view* v, uint32_t n, uint32_t ipt, uint32_t burn_in,
                        uint32_t seed, uint32_t* hist, int threads);
/* Per-step records of one walker, 8 words a step (see ffm_trace). */
void orc_trace(const ffm_flame_desc* f, const float* pal_lin, int pal_size, const ffm_view* v,
               uint32_t walker, uint32_t seed, uint32_t burn_in, uint32_t steps, uint32_t* out);

int orc_step_from(const ffm_flame_desc* f, const float* pal_lin, int pal_size, const ffm_view* v,
                  uint32_t burn_in, orc_rng* rng, orc_walker* w, uint32_t* out8);

/* --- host sample accounting: OrbitFuses (src/renderer/compute_kernel.rs:73-102) --- */
typedef struct {
  uint32_t* left;
  uint32_t len;
} orc_orbit_fuses;
uint64_t orc_orbit_fuses_dispatch(orc_orbit_fuses* c, uint32_t threads, uint32_t ipt,
                                  uint32_t burn_in, uint32_t period, uint32_t seed);
void orc_orbit_fuses_free(orc_orbit_fuses* c);

/* --- stencils --- */
void orc_blur_pass(const uint32_t* in, uint32_t* out, const uint32_t* center_src, int W, int H,
                   float sigma, float density_sigma, int direction);
void orc_blur(uint32_t* hist, uint32_t* scratch, int W, int H, float sigma, float density_sigma);
float orc_blur_density_sigma(double samples, int W, int H, float blur_edges);
void orc_accumulate(const void* hist, int hist_is_f32, const float* prev, float* out,
                    uint32_t* count, int W, int H, int mode, float blend);
void orc_density_estimate(const float* in, float* out, int W, int H, float max_r, float min_r,
                          float curve);
float orc_round_f16(float v);
float orc_curve_eval(const float* px, const float* py, int n, float x);
void orc_curve_lut(const ffm_tonemap_params* tp, float* lut);
float orc_curve_sample(const float* lut, float c);
float orc_step_term_scale(void);
void orc_downsample2x(const uint8_t* in, uint8_t* out, int ow, int oh);
void orc_use_libm(int on);
float orc_fm_atan2(float y, float x);
void orc_tonemap(const float* accum, uint8_t* rgba, int W, int H, float zoom,
                 double samples_in_buffer, const ffm_tonemap_params* tp);
void orc_tonemap_defaults(ffm_tonemap_params* tp);

/* --- whole render (src/renderer/render.rs:483-800) --- */
int orc_render(const ffm_flame_desc* f, const float* stops_prgb, int n_stops, int pal_size,
               const ffm_palette_xf* pal_xf, const ffm_job* job, uint64_t exact_after,
               float* accum_out, uint8_t* rgba_out, uint64_t* total_out, int threads,
               const ffm_effects* fx);

/* --- the render tail's effects (ffm_effects.c; src/renderer/tail.rs) --- */
void orc_density_effect(const ffm_effect* fx, const float* in, float* out, int W, int H,
                        float density_mean, float pixel_scale);
void orc_color_effect(const ffm_effect* fx, const uint8_t* in, uint8_t* out, int W, int H,
                      float pixel_scale);
void orc_apply_blend(const float base[3], const float e[3], int mode, float amount, float out[3]);
void orc_fx_sincos(float x, float* s, float* c);
void orc_density_chain(const ffm_effects* fx, float* acc, float* scratch, int W, int H,
                       float density_mean);
void orc_color_chain(const ffm_effects* fx, uint8_t* rgba, uint8_t* scratch, int W, int H);

/* --- soft splat, FP64 (ffm_soft.c; no reference counterpart) --- */
/* Parameters of a fittable flame, n * (6 + v), or -1. */
int orc_soft_param_count(const ffm_flame_desc* f);
/* density (W*H doubles) := the soft splat at theta (double, the ABI's layout). */
int orc_soft_forward(const ffm_flame_desc* f, const double* theta, const ffm_soft_job* job,
                     double* density, int threads);
/* grad_theta := dL/dtheta for dL/d density = grad_density; exact when
 * job->truncate == 0, else dL/dpoint restarts at zero at every window start. */
int orc_soft_adjoint(const ffm_flame_desc* f, const double* theta, const ffm_soft_job* job,
                     const double* grad_density, double* grad_theta, int threads);
/* One variation's value V[2] and Jacobian J[4] = [dVx/dx, dVx/dy, dVy/dx, dVy/dy]. */
int orc_soft_variation(int id, const float* params, double x, double y, double* V, double* J);

/* --- CPU microbenchmarks of benches/flame_bench.rs (flame_bench.c) --- */
enum {
  ORC_MB_AFFINE_2X3 = 0,
  ORC_MB_AFFINE_FULL = 1,
  ORC_MB_R_SQUARED = 2,
  ORC_MB_R = 3,
  ORC_MB_THETA = 4,
  ORC_MB_PHI = 5,
  ORC_MB_PALETTE_POSITION = 6,
  ORC_MB_SINGLE_TRANSFORM_ITER = 7
};
/* ns per iteration of benchmark `which`. params: a b c d e f, x y (the
 * point), speed, colour, then g (affine_full) or linear weight + has-linear
 * flag (single_transform_iter). iters = 0 calibrates to min_seconds.
 * state_out (3 floats, may be NULL): the last values computed. */
double orc_microbench(int which, const float* params, uint64_t iters, double min_seconds,
                      float* state_out);

#ifdef __cplusplus
}
#endif
#endif
