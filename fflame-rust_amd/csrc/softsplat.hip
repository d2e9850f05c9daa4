// softsplat.hip -- differentiable soft splat: forward density and its adjoint.
//
// The walk is the chaos game's (ffm_device.h: selection, opacity draw,
// transform, bad-value respawn, fuse; shaders/core/main_template.wgsl
// :193-514), with the transforms' affine coefficients and variation weights
// read from a device parameter vector theta. Each plotted sample adds
// density_scale x its bilinear weights to the four pixels around its
// continuous pixel position (world_to_pixel, utilities.wgsl:417-445, without
// the floor, minus half a pixel). The reference has no such mode
// (splat_size is hard-wired to 1.0 and read by nothing,
// src/renderer/compute_kernel.rs:1354).
//
// Walk precision. T = float walks with the chaos game's own arithmetic
// (apply_xform). T = double walks in f64 with the same discrete choices
// (transform picks, opacity, julia's omega, julian's branch all come from the
// same f32 draws and comparisons): a walk's pathwise gradient multiplies
// Jacobians along the trajectory, and on attractors with expanding pieces
// (spherical near the origin) f32 rounding of the positions alone moves the
// gradient by a few 1e-3, so accurate gradients of long walks need the f64
// walk.
//
// Adjoint. For dL/d density = G, dL/dtheta sums, over plotted samples, the
// gradient of the bilinear gather of G at the sample, pulled back through
// the steps that produced the sample. Each walker replays its trajectory
// from its RNG stream in windows of K iterations: a window's per-step
// records (input point, RNG state before the step, flags) live in LDS, and
// the reverse sweep over them re-draws each step's random numbers from the
// recorded state. In exact mode the walk first stores a checkpoint
// (x, y, RNG state, fuse) at every window start, then windows are swept
// last to first with dL/dpoint carried across window boundaries, which is
// the full reverse-mode gradient of the walk; in truncated mode windows are
// swept in walk order and dL/dpoint restarts at zero in each one.
//
// Per-lane parameter gradients accumulate in a lane-private LDS column
// (no atomics: each lane owns its slots); a block's columns are summed in
// f64 into per-block partials, and a finish kernel sums the partials of
// each parameter in a fixed order, so the gradient is deterministic.
#include "ffm_device.h"
#include "ffm_soft.h"

namespace ffm {

// One block size for both kernels, so every fit has one block numbering.
constexpr int SOFT_BLOCK = 128;
constexpr int ACC_STRIDE = SOFT_BLOCK + 1;  // lane-private columns, odd stride

// Scalar math of a walk precision.
template <typename T> struct Fm;
template <> struct Fm<float> {
  static constexpr float PI = PI_F, TWO_PI = TWO_PI_F, EPS = 1e-6f;
  __device__ static void sincos(float x, float* s, float* c) { ffm_sincosf(x, s, c); }
  __device__ static float atan2(float y, float x) { return ffm_atan2f(y, x); }
  __device__ static float atan2o(float y, float x) { return ff_atan2(y, x); }
  __device__ static float sqrt(float x) { return sqrtf(x); }
  __device__ static float pow(float x, float y) { return powf(x, y); }
};
template <> struct Fm<double> {
  static constexpr double PI = 3.14159265358979323846, TWO_PI = 2.0 * PI, EPS = 1e-6;
  __device__ static void sincos(double x, double* s, double* c) { ::sincos(x, s, c); }
  __device__ static double atan2(double y, double x) { return ::atan2(y, x); }
  // IEEE atan2 already gives ff_atan2's signed-zero results at the origin
  __device__ static double atan2o(double y, double x) { return ::atan2(y, x); }
  __device__ static double sqrt(double x) { return ::sqrt(x); }
  __device__ static double pow(double x, double y) { return ::pow(x, y); }
};

// Fit of a flat block index (blocks of all fits back to back).
__device__ __forceinline__ int soft_fit_of(const SoftFit* fits, int n, uint32_t b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (fits[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The fit's table (FlameHeader + rows) in LDS with theta's affine
// coefficients and variation weights written over the descriptor's.
__device__ __forceinline__ FlameView soft_load(const SoftArgs& a, const SoftFit& F, float* lds) {
  const int nx = F.n_xforms, nv = F.n_vars;
  const int hdr_w = (int)(sizeof(FlameHeader) / 4);
  const int words = hdr_w + nx * XF_STRIDE;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(a.tables + F.table_off);
  uint32_t* dst = reinterpret_cast<uint32_t*>(lds);
  for (int i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
  const float* th = a.theta + F.theta_off;
  float* rows = lds + hdr_w;
  for (int i = threadIdx.x; i < nx * 6; i += blockDim.x)
    rows[(i / 6) * XF_STRIDE + XF_A + i % 6] = th[i];
  for (int i = threadIdx.x; i < nx * nv; i += blockDim.x)
    rows[(i / nv) * XF_STRIDE + XF_W0 + i % nv] = th[nx * 6 + i];
  __syncthreads();
  FlameView fv;
  fv.rows = rows;
  fv.ids = (KInt*)(a.tables + F.table_off) + 4;
  fv.hdr = nullptr;  // no final transforms in a soft-splat flame
  fv.sel = nullptr;  // the LDS copy holds the transforms' rows only
  fv.sel_byte = 0;
  fv.xcum = nullptr;  // nor xaos
  fv.n_xforms = nx;
  fv.n_finals = 0;
  fv.n_vars = nv;
  fv.share_tp = 0;
  return fv;
}

// ------------------------------------------------------- variations

// V(x, y) and, when JAC, J = [dVx/dx, dVx/dy, dVy/dx, dVy/dy] of one
// variation in precision T. `draw`: whether the variation consumes its
// random draw (the forward draws only for a non-zero weight). Bodies as
// eval_variation (src/variations/defs/basic.rs, advanced.rs); the f64 forms
// are the oracle's.
#define FFM_NO_SPECULATE_XY() asm volatile("" : "+v"(x), "+v"(y))
template <typename T, bool JAC>
__device__ __forceinline__ void var_jac(int id, const float* __restrict__ p, T x, T y, Rng& rng,
                                        bool draw, T V[2], T J[4]) {
  using M = Fm<T>;
  const T one = (T)1, two = (T)2, zero = (T)0;
  switch (id) {
    case FFM_VAR_LINEAR:
      V[0] = x; V[1] = y;
      if (JAC) { J[0] = one; J[1] = zero; J[2] = zero; J[3] = one; }
      break;
    case FFM_VAR_SINUSOIDAL: {
      FFM_NO_SPECULATE_XY();
      T sx, cx, sy, cy;
      M::sincos(x, &sx, &cx);
      M::sincos(y, &sy, &cy);
      V[0] = sx; V[1] = sy;
      if (JAC) { J[0] = cx; J[1] = zero; J[2] = zero; J[3] = cy; }
      break;
    }
    case FFM_VAR_SPHERICAL:
    case FFM_VAR_HYPERBOLIC: {
      FFM_NO_SPECULATE_XY();
      T s = x * x + y * y + M::EPS;
      V[0] = x / s;
      V[1] = id == FFM_VAR_SPHERICAL ? y / s : y;
      if (JAC) {
        T inv = one / s, inv2 = inv * inv;
        J[0] = inv - two * x * x * inv2;
        J[1] = -two * x * y * inv2;
        J[2] = id == FFM_VAR_SPHERICAL ? J[1] : zero;
        J[3] = id == FFM_VAR_SPHERICAL ? inv - two * y * y * inv2 : one;
      }
      break;
    }
    case FFM_VAR_SWIRL: {
      FFM_NO_SPECULATE_XY();
      T r2 = x * x + y * y, S, C;
      M::sincos(r2, &S, &C);
      V[0] = x * S - y * C;
      V[1] = x * C + y * S;
      if (JAC) {
        J[0] = S + two * x * V[1]; J[1] = -C + two * y * V[1];
        J[2] = C - two * x * V[0]; J[3] = S - two * y * V[0];
      }
      break;
    }
    case FFM_VAR_HORSESHOE: {
      FFM_NO_SPECULATE_XY();
      T rho = M::sqrt(x * x + y * y), r = rho + M::EPS, ir = one / r;
      V[0] = (x - y) * (x + y) * ir;
      V[1] = two * x * y * ir;
      if (JAC) {
        T irho = rho > zero ? one / rho : zero;
        J[0] = (two * x - V[0] * x * irho) * ir; J[1] = (-two * y - V[0] * y * irho) * ir;
        J[2] = (two * y - V[1] * x * irho) * ir; J[3] = (two * x - V[1] * y * irho) * ir;
      }
      break;
    }
    case FFM_VAR_POLAR:
    case FFM_VAR_DISC: {
      FFM_NO_SPECULATE_XY();
      T r2 = x * x + y * y, rho = M::sqrt(r2);
      T a = M::atan2(x, y) / M::PI;
      T sb = zero, cb = zero;
      if (id == FFM_VAR_POLAR) {
        V[0] = a; V[1] = rho - one;
      } else {
        M::sincos(M::PI * rho, &sb, &cb);
        V[0] = a * sb; V[1] = a * cb;
      }
      if (JAC) {
        T ipr2 = r2 > zero ? one / (M::PI * r2) : zero;
        T irho = rho > zero ? one / rho : zero;
        T dax = y * ipr2, day = -x * ipr2;
        if (id == FFM_VAR_POLAR) {
          J[0] = dax; J[1] = day; J[2] = x * irho; J[3] = y * irho;
        } else {
          T dbx = M::PI * x * irho, dby = M::PI * y * irho;
          J[0] = dax * sb + a * cb * dbx; J[1] = day * sb + a * cb * dby;
          J[2] = dax * cb - a * sb * dbx; J[3] = day * cb - a * sb * dby;
        }
      }
      break;
    }
    case FFM_VAR_JULIA: {
      FFM_NO_SPECULATE_XY();
      T r2 = x * x + y * y, rho = M::sqrt(r2);
      T omega = (draw && rng_nextf(rng) < 0.5f) ? M::PI : zero;
      T s, c;
      M::sincos(M::atan2(y, x) / two + omega, &s, &c);
      T sr = M::sqrt(rho);
      V[0] = sr * c; V[1] = sr * s;
      if (JAC) {
        T i2 = r2 > zero ? (T)0.5 / r2 : zero;  // f' = V / (2 z)
        T re = (V[0] * x + V[1] * y) * i2, im = (V[1] * x - V[0] * y) * i2;
        J[0] = re; J[1] = -im; J[2] = im; J[3] = re;
      }
      break;
    }
    case FFM_VAR_JULIAN: {
      FFM_NO_SPECULATE_XY();
      const float power = p[0];
      // f32: the prepared cpower slot; f64: dist / max(|power|, 1e-30) / 2 in f64
      T cp = sizeof(T) == 4 ? (T)p[2]
                            : (T)p[1] / (T)fmax(fabs((double)power), 1e-30) / (T)2;
      T r2 = x * x + y * y;
      T r = M::pow(r2, cp);
      float tr = draw ? floorf(fabsf(power) * rng_nextf(rng)) : 0.0f;
      T t = (M::atan2o(y, x) + M::TWO_PI * (T)tr) / (T)power;
      T st, ct;
      M::sincos(t, &st, &ct);
      V[0] = r * ct; V[1] = r * st;
      if (JAC) {
        T ir2 = r2 > zero ? one / r2 : zero;
        T drx = two * cp * r * x * ir2, dry = two * cp * r * y * ir2;
        T dtx = -y * ir2 / (T)power, dty = x * ir2 / (T)power;
        J[0] = drx * ct - r * st * dtx; J[1] = dry * ct - r * st * dty;
        J[2] = drx * st + r * ct * dtx; J[3] = dry * st + r * ct * dty;
      }
      break;
    }
    case FFM_VAR_BUBBLE: {
      FFM_NO_SPECULATE_XY();
      T s = (x * x + y * y) / (T)4 + one, is = one / s;
      V[0] = x * is; V[1] = y * is;
      if (JAC) {
        T h = (T)0.5 * is * is;
        J[0] = is - x * x * h; J[1] = -x * y * h; J[2] = J[1]; J[3] = is - y * y * h;
      }
      break;
    }
    case FFM_VAR_CURL: {
      FFM_NO_SPECULATE_XY();
      T c1 = p[0], c2 = p[1];
      T re = one + c1 * x + c2 * (x * x - y * y), im = c1 * y + two * c2 * x * y;
      T d = re * re + im * im;
      V[0] = (x * re + y * im) / d;
      V[1] = (y * re - x * im) / d;
      if (JAC) {
        // f = z / w, f' = (1 - c2 z^2) / w^2 = (1 - c2 z^2) conj(w)^2 / |w|^4
        T id2 = one / d;
        T nr = one - c2 * (x * x - y * y), ni = -two * c2 * x * y;
        T wr = re * re - im * im, wi = -two * re * im;
        T fr = (nr * wr - ni * wi) * id2 * id2, fi = (nr * wi + ni * wr) * id2 * id2;
        J[0] = fr; J[1] = -fi; J[2] = fi; J[3] = fr;
      }
      break;
    }
    default:
      V[0] = V[1] = zero;
      if (JAC) J[0] = J[1] = J[2] = J[3] = zero;
      break;
  }
}

// The transform in f64: affine -> weight-gated variation sum in local-index
// order -> post-affine (the f64 form of apply_xform).
__device__ __forceinline__ void apply_xform_f64(const float* __restrict__ row, KInt* ids,
                                                int nv, double& x, double& y, Rng& rng) {
  const double qx = (double)row[XF_A] * x + (double)row[XF_B] * y + (double)row[XF_E];
  const double qy = (double)row[XF_C] * x + (double)row[XF_D] * y + (double)row[XF_F];
  double rx = 0.0, ry = 0.0;
  for (int k = 0; k < nv; k++) {
    const int id = __builtin_amdgcn_readfirstlane(ids[k]);
    const double w = row[XF_W0 + k];
    if (w != 0.0) {
      double V[2], J[4];
      var_jac<double, false>(id, row + XF_P0 + k * FFM_MAX_VAR_PARAMS, qx, qy, rng, true, V, J);
      rx += w * V[0];
      ry += w * V[1];
    }
  }
  if (row[XF_POST] != 0.0f) {
    const double tx = (double)row[XF_PA] * rx + (double)row[XF_PB] * ry + (double)row[XF_PE];
    const double ty = (double)row[XF_PC] * rx + (double)row[XF_PD] * ry + (double)row[XF_PF];
    rx = tx;
    ry = ty;
  }
  x = rx;
  y = ry;
}

// One step of the walk: chaos_step without colour, final chain or binning.
template <typename T>
__device__ __forceinline__ int soft_step(const FlameView& fl, uint32_t burn_in, Rng& rng, T& x,
                                         T& y, uint32_t& fuse, int& xi_out) {
  float u = rng_nextf(rng);
  const int n = fl.n_xforms;
  float target = u * fl.rows[(n - 1) * XF_STRIDE + XF_CUM];
  int xi = n - 1;
  for (int i = n - 2; i >= 0; i--)
    if (target <= fl.rows[i * XF_STRIDE + XF_CUM]) xi = i;
  xi_out = xi;
  const float* row = fl.rows + xi * XF_STRIDE;
  bool should_plot = rng_nextf(rng) < row[XF_OPACITY];
  T nx = x, ny = y;
  if constexpr (sizeof(T) == 4) apply_xform<false>(row, fl.ids, fl.n_vars, nx, ny, rng);
  else apply_xform_f64(row, fl.ids, fl.n_vars, nx, ny, rng);
  if (!(fabs(nx) <= (T)1e32) || !(fabs(ny) <= (T)1e32)) {
    x = (T)(rng_nextf(rng) * 2.0f - 1.0f);
    y = (T)(rng_nextf(rng) * 2.0f - 1.0f);
    fuse = burn_in;
    return ST_RESPAWN;
  }
  x = nx;
  y = ny;
  if (fuse > 0u) {
    fuse -= 1u;
    return 0;
  }
  return should_plot ? ST_PLOT : 0;
}

// Continuous pixel position of a world point; false when no tap is in frame.
template <typename T>
__device__ __forceinline__ bool soft_pixel(const ViewConst& v, T x, T y, T& sx, T& sy) {
  T tx = x - (T)v.pan_x, ty = y - (T)v.pan_y;
  if constexpr (sizeof(T) == 4) {
    float rx = (tx * v.cos_r - ty * v.sin_r) * v.zoom;
    float ry = (tx * v.sin_r + ty * v.cos_r) * v.zoom;
    sx = (v.cx - 0.5f) + rx * v.scale;
    sy = (v.cy - 0.5f) + ry * v.scale;
  } else {
    const double k = (double)v.zoom * (double)v.scale;
    sx = ((double)v.cx - 0.5) + (tx * (double)v.cos_r - ty * (double)v.sin_r) * k;
    sy = ((double)v.cy - 0.5) + (tx * (double)v.sin_r + ty * (double)v.cos_r) * k;
  }
  return sx > (T)-1 && sy > (T)-1 && sx < (T)v.width && sy < (T)v.height;
}

template <typename T>
__device__ __forceinline__ void walker_init(uint32_t tid, uint32_t seed, uint32_t burn_in,
                                            Rng& rng, T& x, T& y, uint32_t& fuse) {
  rng = rng_init(tid, seed);
  x = (T)(rng_nextf(rng) * 2.0f - 1.0f);
  y = (T)(rng_nextf(rng) * 2.0f - 1.0f);
  fuse = burn_in;
}

// ------------------------------------------------------------- forward

template <typename T>
__global__ __launch_bounds__(SOFT_BLOCK) void soft_forward_kernel(SoftArgs a) {
  extern __shared__ float lds[];
  const int fi = soft_fit_of(a.fits, a.n_fits, blockIdx.x);
  const SoftFit F = a.fits[fi];
  if (F.f64 != (int)(sizeof(T) == 8)) return;  // the other precision's launch owns it
  const FlameView fv = soft_load(a, F, lds);
  const uint64_t tid = (uint64_t)(blockIdx.x - F.block0) * blockDim.x + threadIdx.x;
  if (tid >= F.walkers) return;
  Rng rng;
  T x, y;
  uint32_t fuse;
  walker_init((uint32_t)tid, F.seed, F.burn_in, rng, x, y, fuse);
  // this XCD's private copy: its atomics resolve in the XCD's own L2
  float* dens = a.dens_xcd + (size_t)xcc_id() * a.total_pixels + F.pix_off;
  const int W = F.view.width, H = F.view.height;
  for (uint32_t it = 0; it < F.ipt; it++) {
    int xi;
    if (!(soft_step(fv, F.burn_in, rng, x, y, fuse, xi) & ST_PLOT)) continue;
    T sx, sy;
    if (!soft_pixel(F.view, x, y, sx, sy)) continue;
    T x0 = floor(sx), y0 = floor(sy);
    T ux = sx - x0, uy = sy - y0;
    int ix = (int)x0, iy = (int)y0;
    T s = (T)F.density_scale, o = (T)1;
    float w[4] = {(float)((o - ux) * (o - uy) * s), (float)(ux * (o - uy) * s),
                  (float)((o - ux) * uy * s), (float)(ux * uy * s)};
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int X = ix + (c & 1), Y = iy + (c >> 1);
      if (X >= 0 && X < W && Y >= 0 && Y < H)
        __hip_atomic_fetch_add(dens + (size_t)Y * W + X, w[c], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
}

// density = the sum of the XCD copies, in copy order.
__global__ __launch_bounds__(256) void soft_fold_kernel(SoftArgs a) {
  const uint64_t n = a.total_pixels;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < SOFT_XCD_COPIES; c++) s += a.dens_xcd[(size_t)c * n + i];
    a.density[i] = s;
  }
}

// ------------------------------------------------------------- adjoint

// dL/d(world point) of one plotted sample: the bilinear gather of G at the
// sample, differentiated (pixels out of frame hold 0).
template <typename T>
__device__ __forceinline__ void splat_grad(const ViewConst& v, const float* __restrict__ G,
                                           T kscale, T x, T y, T& lx, T& ly) {
  T sx, sy;
  if (!soft_pixel(v, x, y, sx, sy)) return;
  T x0 = floor(sx), y0 = floor(sy);
  T ux = sx - x0, uy = sy - y0, o = (T)1;
  int ix = (int)x0, iy = (int)y0;
  const int W = v.width, H = v.height;
  T g[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    int X = ix + (c & 1), Y = iy + (c >> 1);
    g[c] = (X >= 0 && X < W && Y >= 0 && Y < H) ? (T)G[(size_t)Y * W + X] : (T)0;
  }
  T gsx = (o - uy) * (g[1] - g[0]) + uy * (g[3] - g[2]);
  T gsy = (o - ux) * (g[2] - g[0]) + ux * (g[3] - g[1]);
  // s = c + k R (p - pan): dL/dp = k R^T dL/ds
  const T cr = v.cos_r, sr = v.sin_r;
  lx += kscale * (cr * gsx + sr * gsy);
  ly += kscale * (-sr * gsx + cr * gsy);
}

// Pull (lx, ly) = dL/d(output point) back through the step taken with
// transform `xi` from (x, y), RNG state `rs` before the step. Adds the
// parameter gradient to this lane's accumulator column; leaves dL/d(input
// point) in (lx, ly).
template <typename T>
__device__ __forceinline__ void step_back(const FlameView& fl, T x, T y, Rng rs, int xi, T& lx,
                                          T& ly, T* __restrict__ acc) {
  (void)rng_nextf(rs);  // selection
  (void)rng_nextf(rs);  // opacity
  const float* row = fl.rows + xi * XF_STRIDE;
  const T a = row[XF_A], b = row[XF_B], c = row[XF_C], d = row[XF_D];
  const T qx = a * x + b * y + (T)row[XF_E];
  const T qy = c * x + d * y + (T)row[XF_F];
  T vx = lx, vy = ly;
  if (row[XF_POST] != 0.0f) {
    vx = (T)row[XF_PA] * lx + (T)row[XF_PC] * ly;
    vy = (T)row[XF_PB] * lx + (T)row[XF_PD] * ly;
  }
  T J0 = 0, J1 = 0, J2 = 0, J3 = 0;
  const int nx = fl.n_xforms, nv = fl.n_vars;
  T* gw = acc + (6 * nx + nv * xi) * ACC_STRIDE;
  for (int k = 0; k < nv; k++) {
    const int id = __builtin_amdgcn_readfirstlane(fl.ids[k]);
    const T w = row[XF_W0 + k];
    T V[2], J[4];
    var_jac<T, true>(id, row + XF_P0 + k * FFM_MAX_VAR_PARAMS, qx, qy, rs, w != (T)0, V, J);
    T gv = vx * V[0] + vy * V[1];
    if (isfinite(gv)) gw[k * ACC_STRIDE] += gv;
    if (w != (T)0) {
      J0 += w * J[0]; J1 += w * J[1]; J2 += w * J[2]; J3 += w * J[3];
    }
  }
  const T mx = J0 * vx + J2 * vy;  // mu = Js^T v
  const T my = J1 * vx + J3 * vy;
  T* ga = acc + 6 * xi * ACC_STRIDE;
  ga[0 * ACC_STRIDE] += mx * x;
  ga[1 * ACC_STRIDE] += mx * y;
  ga[2 * ACC_STRIDE] += my * x;
  ga[3 * ACC_STRIDE] += my * y;
  ga[4 * ACC_STRIDE] += mx;
  ga[5 * ACC_STRIDE] += my;
  lx = a * mx + c * my;  // A^T mu
  ly = b * mx + d * my;
}

// A window's records in LDS, (K + 1) rows of SOFT_BLOCK lanes: the step's
// input point, and (RNG state before the step, flags | transform << 8).
template <typename T> struct Rec {
  T* xy;  // 2 per slot
  uint2* sf;
  __device__ void put(int j, T x, T y, uint32_t s, uint32_t f) {
    xy[2 * j * SOFT_BLOCK] = x;
    xy[2 * j * SOFT_BLOCK + 1] = y;
    sf[j * SOFT_BLOCK] = make_uint2(s, f);
  }
};

// Replay `len` steps from the state, recording each step's input.
template <typename T>
__device__ __forceinline__ void replay(const FlameView& fl, uint32_t burn_in, Rng& rng, T& x, T& y,
                                       uint32_t& fuse, int len, Rec<T> rec) {
  for (int j = 0; j < len; j++) {
    const uint32_t st = rng.state;
    const T px = x, py = y;
    int xi;
    int f = soft_step(fl, burn_in, rng, x, y, fuse, xi);
    rec.put(j, px, py, st, (uint32_t)(f | (xi << 8)));
  }
  rec.put(len, x, y, 0u, 0u);
}

// Reverse sweep over a replayed window.
template <typename T>
__device__ __forceinline__ void sweep(const FlameView& fl, const SoftFit& F, const float* G,
                                      uint32_t inc, int len, Rec<T> rec, T& lx, T& ly, T* acc) {
  const T kscale = (T)F.density_scale * (T)F.view.zoom * (T)F.view.scale;
  T ox = rec.xy[2 * len * SOFT_BLOCK], oy = rec.xy[2 * len * SOFT_BLOCK + 1];
  for (int j = len - 1; j >= 0; j--) {
    const T x = rec.xy[2 * j * SOFT_BLOCK], y = rec.xy[2 * j * SOFT_BLOCK + 1];
    const uint2 sf = rec.sf[j * SOFT_BLOCK];
    if (sf.y & ST_PLOT) splat_grad(F.view, G, kscale, ox, oy, lx, ly);
    if (sf.y & ST_RESPAWN) {
      lx = ly = (T)0;
    } else {
      Rng rs;
      rs.state = sf.x;
      rs.inc = inc;
      step_back(fl, x, y, rs, (int)(sf.y >> 8), lx, ly, acc);
    }
    ox = x;
    oy = y;
  }
}

// Byte offsets of the adjoint's LDS regions for one fit.
struct AdjLayout {
  size_t rec_xy, rec_sf, acc, total;
};
__host__ __device__ inline AdjLayout adj_layout(int n_xforms, int n_params, int window,
                                                int tsize) {
  const size_t hdr_w = sizeof(FlameHeader) / 4;
  AdjLayout l;
  l.rec_xy = ((hdr_w + (size_t)n_xforms * XF_STRIDE) * 4 + 15) & ~(size_t)15;
  const size_t slots = (size_t)(window + 1) * SOFT_BLOCK;
  l.rec_sf = l.rec_xy + slots * 2 * tsize;
  l.acc = l.rec_sf + slots * 8;
  l.total = l.acc + (size_t)n_params * ACC_STRIDE * tsize;
  return l;
}

template <typename T>
__global__ __launch_bounds__(SOFT_BLOCK) void soft_adjoint_kernel(SoftArgs a) {
  extern __shared__ float lds[];
  const int fi = soft_fit_of(a.fits, a.n_fits, blockIdx.x);
  const SoftFit F = a.fits[fi];
  if (F.f64 != (int)(sizeof(T) == 8)) return;  // the other precision's launch owns it
  const FlameView fv = soft_load(a, F, lds);
  const int K = (int)F.window;
  const AdjLayout L = adj_layout(F.n_xforms, F.n_params, K, (int)sizeof(T));
  char* base = reinterpret_cast<char*>(lds);
  Rec<T> rec;
  rec.xy = reinterpret_cast<T*>(base + L.rec_xy) + 2 * threadIdx.x;
  rec.sf = reinterpret_cast<uint2*>(base + L.rec_sf) + threadIdx.x;
  T* acc_base = reinterpret_cast<T*>(base + L.acc);
  T* acc = acc_base + threadIdx.x;
  for (int q = 0; q < F.n_params; q++) acc[q * ACC_STRIDE] = (T)0;

  const uint32_t blk = blockIdx.x - F.block0;
  const uint64_t tid = (uint64_t)blk * SOFT_BLOCK + threadIdx.x;
  if (tid < F.walkers) {
    const float* G = a.grad_density + F.pix_off;
    Rng rng;
    T x, y;
    uint32_t fuse;
    walker_init((uint32_t)tid, F.seed, F.burn_in, rng, x, y, fuse);
    const uint32_t inc = rng.inc;
    const int nw = (int)((F.ipt + (uint32_t)K - 1u) / (uint32_t)K);
    T lx = 0, ly = 0;
    if (F.truncate) {
      for (int w = 0; w < nw; w++) {
        const int len = min(K, (int)F.ipt - w * K);
        replay(fv, F.burn_in, rng, x, y, fuse, len, rec);
        lx = ly = (T)0;
        sweep(fv, F, G, inc, len, rec, lx, ly, acc);
      }
    } else {
      // checkpoints: nw x walkers points (2 T each), then nw x walkers (state, fuse)
      T* ck_xy = reinterpret_cast<T*>(a.ckpt + F.ckpt_off);
      uint2* ck_sf = reinterpret_cast<uint2*>(a.ckpt + F.ckpt_off +
                                              (uint64_t)nw * F.walkers * 2 * sizeof(T));
      // walk once, keeping each window's starting state
      for (int w = 0; w < nw; w++) {
        const uint64_t slot = (uint64_t)w * F.walkers + tid;
        ck_xy[2 * slot] = x;
        ck_xy[2 * slot + 1] = y;
        ck_sf[slot] = make_uint2(rng.state, fuse);
        const int len = min(K, (int)F.ipt - w * K);
        for (int j = 0; j < len; j++) {
          int xi;
          (void)soft_step(fv, F.burn_in, rng, x, y, fuse, xi);
        }
      }
      for (int w = nw - 1; w >= 0; w--) {
        const uint64_t slot = (uint64_t)w * F.walkers + tid;
        x = ck_xy[2 * slot];
        y = ck_xy[2 * slot + 1];
        const uint2 sf = ck_sf[slot];
        rng.state = sf.x;
        fuse = sf.y;
        const int len = min(K, (int)F.ipt - w * K);
        replay(fv, F.burn_in, rng, x, y, fuse, len, rec);
        sweep(fv, F, G, inc, len, rec, lx, ly, acc);
      }
    }
  }
  __syncthreads();
  // Block partials: column sums in f64, fixed order.
  for (int q = threadIdx.x; q < F.n_params; q += SOFT_BLOCK) {
    const T* col = acc_base + q * ACC_STRIDE;
    double s = 0.0;
    for (int t = 0; t < SOFT_BLOCK; t++) s += (double)col[t];
    a.partials[F.part_off + (uint64_t)q * F.n_blocks + blk] = s;
  }
}

// grad_theta[q] of fit blockIdx.y = the sum of its blocks' partials, in a
// fixed order (strided per-thread sums, then a tree).
__global__ __launch_bounds__(256) void soft_finish_kernel(SoftArgs a) {
  const SoftFit F = a.fits[blockIdx.y];
  const int q = blockIdx.x;
  if (q >= F.n_params) return;
  const double* p = a.partials + F.part_off + (uint64_t)q * F.n_blocks;
  double s = 0.0;
  for (uint32_t b = threadIdx.x; b < F.n_blocks; b += blockDim.x) s += p[b];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.grad_theta[F.theta_off + q] = (float)red[0];
}

// ------------------------------------------------------------- host side

size_t soft_adjoint_lds_bytes(int n_xforms, int n_params, int window, bool f64) {
  return adj_layout(n_xforms, n_params, window, f64 ? 8 : 4).total;
}

size_t soft_ckpt_bytes(uint32_t ipt, uint32_t window, uint64_t walkers, bool f64) {
  const uint64_t nw = (ipt + window - 1) / window;
  return (size_t)(nw * walkers * (2 * (f64 ? 8 : 4) + 8));
}

size_t soft_forward_lds_bytes(int n_xforms) {
  return 4 * (sizeof(FlameHeader) / 4 + (size_t)n_xforms * XF_STRIDE);
}

int soft_block() { return SOFT_BLOCK; }

hipError_t launch_soft_forward(const SoftArgs& a, uint32_t blocks, size_t lds, bool any_f32,
                               bool any_f64, hipStream_t s) {
  hipError_t e = launch_clear_u32(a.dens_xcd, (size_t)SOFT_XCD_COPIES * a.total_pixels, s);
  if (e != hipSuccess) return e;
  if (any_f32)
    hipLaunchKernelGGL(soft_forward_kernel<float>, dim3(blocks), dim3(SOFT_BLOCK), lds, s, a);
  if (any_f64)
    hipLaunchKernelGGL(soft_forward_kernel<double>, dim3(blocks), dim3(SOFT_BLOCK), lds, s, a);
  const uint64_t fold_blocks = std::min<uint64_t>((a.total_pixels + 255) / 256, 4096);
  hipLaunchKernelGGL(soft_fold_kernel, dim3((uint32_t)fold_blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_soft_adjoint(const SoftArgs& a, uint32_t blocks, size_t lds32, size_t lds64,
                               hipStream_t s) {
  hipError_t e;
  if (lds32) {
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(soft_adjoint_kernel<float>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, LDS_DYNAMIC_MAX);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(soft_adjoint_kernel<float>, dim3(blocks), dim3(SOFT_BLOCK), lds32, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (lds64) {
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(soft_adjoint_kernel<double>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, LDS_DYNAMIC_MAX);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(soft_adjoint_kernel<double>, dim3(blocks), dim3(SOFT_BLOCK), lds64, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(soft_finish_kernel, dim3(FFM_SOFT_MAX_XFORMS * (6 + FFM_SOFT_MAX_VARS),
                                              a.n_fits), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace ffm
