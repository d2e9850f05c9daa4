/* CPU microbenchmarks of the reference's Criterion suite, restated in C.
 *
 * TEST / MEASUREMENT INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times
 * these on the host beside the GPU run, to tie the host to the per-op
 * figures the reference publishes (benchmark_results/last_run.json:5-75).
 * Nothing of the product links or calls this file.
 *
 * Each benchmark follows benches/flame_bench.rs:
 *   affine_transform/affine_2x3_matrix, affine_full   (flame_bench.rs:21-56)
 *   point_helpers/r_squared, r, theta, phi           (flame_bench.rs:58-82;
 *       Point helpers src/scene/transforms.rs:1983-2006, f32)
 *   color_blending/palette_position/{0,0.5,1}         (flame_bench.rs:84-105)
 *   preset_iteration/single_transform_iter/<preset>   (flame_bench.rs:107-149)
 * The Rust benches whose cost is the host's own data structures (HashMap
 * lookups, the variation registry, Transform construction and cloning,
 * flame_bench.rs:151-237) are not on the chaos-game path and are not
 * restated.
 *
 * Criterion passes each result through black_box, which forces the value
 * to memory; sink() does the same. Inputs go through source() once per
 * iteration so the compiler cannot fold the loop body to a constant.
 * Arithmetic is f32 without contraction (-ffp-contract=off), as the
 * reference's Point and Transform fields are f32 (transforms.rs:1975-1978);
 * subnormals are kept (no flush-to-zero), as Rust keeps them.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdint.h>
#include <time.h>

#include "ffm_oracle.h"

static inline void sink(const void* p) { __asm__ volatile("" : : "r"(p) : "memory"); }
#define SINK(v)            \
  do {                     \
    __typeof__(v) s_ = (v); \
    sink(&s_);             \
  } while (0)
#define SOURCE(v) __asm__ volatile("" : "+m"(v))

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* One timed run of `iters` iterations of benchmark `which` with parameters
 * p[] (see orc_microbench). Returns seconds; *state_out receives the last
 * values so a test can check what was computed. */
static double run(int which, const float* p, uint64_t iters, float* state_out) {
  float a = p[0], b = p[1], c = p[2], d = p[3], e = p[4], f = p[5];
  float px = p[6], py = p[7], s = p[8], col = p[9];
  float ox = 0.f, oy = 0.f, oc = 0.f;
  const double t0 = now_s();
  switch (which) {
    case ORC_MB_AFFINE_2X3:
    case ORC_MB_AFFINE_FULL:
      for (uint64_t i = 0; i < iters; i++) {
        SOURCE(px);
        SOURCE(py);
        if (which == ORC_MB_AFFINE_FULL) SINK(p[10]);  /* the transform's z coefficient g */
        ox = a * px + b * py + e;
        oy = c * px + d * py + f;
        SINK(ox);
        SINK(oy);
      }
      break;
    case ORC_MB_R_SQUARED:
      for (uint64_t i = 0; i < iters; i++) {
        SOURCE(px);
        SOURCE(py);
        ox = px * px + py * py;
        SINK(ox);
      }
      break;
    case ORC_MB_R:
      for (uint64_t i = 0; i < iters; i++) {
        SOURCE(px);
        SOURCE(py);
        ox = sqrtf(px * px + py * py);
        SINK(ox);
      }
      break;
    case ORC_MB_THETA:
      for (uint64_t i = 0; i < iters; i++) {
        SOURCE(px);
        SOURCE(py);
        ox = atan2f(py, px);
        SINK(ox);
      }
      break;
    case ORC_MB_PHI:
      for (uint64_t i = 0; i < iters; i++) {
        SOURCE(px);
        SOURCE(py);
        ox = atan2f(px, py);
        SINK(ox);
      }
      break;
    case ORC_MB_PALETTE_POSITION: {
      /* current_pos * (1 - speed) + transform_color * speed, with
       * current_pos = px and transform_color = py (flame_bench.rs:92-99) */
      for (uint64_t i = 0; i < iters; i++) {
        SOURCE(px);
        SOURCE(py);
        SOURCE(s);
        ox = px * (1.0f - s) + py * s;
        SINK(ox);
      }
      break;
    }
    case ORC_MB_SINGLE_TRANSFORM_ITER: {
      /* The preset's first transform iterated from (0.5, 0.5), colour 0.5:
       * affine, then the linear weight p[10] when the transform has one
       * (p[11] != 0) or the affine point as is, then the colour blend
       * color_pos * (1 - s) + color * s (flame_bench.rs:121-143). The
       * state carries across iterations, so a contracting transform walks
       * into subnormals as the reference's loop does. */
      const float lw = p[10];
      const int has_linear = p[11] != 0.0f;
      float x = 0.5f, y = 0.5f, cp = 0.5f;
      for (uint64_t i = 0; i < iters; i++) {
        const float ax = a * x + b * y + e;
        const float ay = c * x + d * y + f;
        if (has_linear) {
          x = ax * lw;
          y = ay * lw;
        } else {
          x = ax;
          y = ay;
        }
        cp = cp * (1.0f - s) + col * s;
        SINK(x);
        SINK(y);
        SINK(cp);
      }
      ox = x;
      oy = y;
      oc = cp;
      break;
    }
    default:
      return -1.0;
  }
  const double dt = now_s() - t0;
  if (state_out) {
    state_out[0] = ox;
    state_out[1] = oy;
    state_out[2] = oc;
  }
  return dt;
}

double orc_microbench(int which, const float* params, uint64_t iters, double min_seconds,
                      float* state_out) {
  if (iters == 0) {
    /* calibrate: double the count until one run takes min_seconds */
    iters = 1u << 16;
    for (;;) {
      const double dt = run(which, params, iters, state_out);
      if (dt < 0) return dt;
      if (dt >= min_seconds || iters >= (1ull << 40)) return 1e9 * dt / (double)iters;
      iters *= 2;
    }
  }
  const double dt = run(which, params, iters, state_out);
  return dt < 0 ? dt : 1e9 * dt / (double)iters;
}
