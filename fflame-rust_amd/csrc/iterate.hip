// iterate.hip -- the chaos-game walker kernel and its splat, for gfx950.
//
// Replaces the reference's compute dispatch `main`
// (shaders/core/main_template.wgsl:20-1428, @workgroup_size(64), 128
// workgroups, src/renderer/render.rs:350,647-666). Here the launch is a
// persistent grid: each lane walks walkers w = gid, gid + nthreads, ...
// with that walker's own PCG32 stream rng_init(walker, seed), so the
// result does not depend on the launch shape. The flame table and the
// palette live in LDS for the whole launch.
//
// Splat modes (the reference: 4 x u32 atomicAdd per in-frame sample,
// main_template.wgsl:1253-1320):
//   U32_GLOBAL  device-scope u32 atomics into one histogram.
//   U32_XCD     one private histogram per XCD, u32 atomics at workgroup
//               scope (resolved in that XCD's L2), folded by fold_xcd.
//   F32_GLOBAL  f32 atomics ("float4 accumulation").
#include "ffm_device.h"

namespace ffm {

// Wave-level sum then one atomic per wave.
__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, unsigned long long v) {
  if (!dst) return;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v)
    __hip_atomic_fetch_add(dst, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int MODE, bool EXT>
__global__ __launch_bounds__(256) void iterate_kernel(IterArgs a) {
  extern __shared__ uint32_t lds[];
  LdsFlame L = load_lds<MODE == FFM_SPLAT_F32_GLOBAL>(a, lds);
  const ViewConst v = a.view;
  const int pal_size = a.pal_size;
  uint32_t* hist_u = reinterpret_cast<uint32_t*>(a.hist);
  float* hist_f = reinterpret_cast<float*>(a.hist);
  if (MODE == FFM_SPLAT_U32_XCD) hist_u += (size_t)xcc_id() * a.hist_copy_words;

  unsigned long long n_plotted = 0, n_in = 0;
  const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t wi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; wi < a.walkers;
       wi += nthreads) {
    const uint32_t walker = a.walker0 + (uint32_t)wi;
    Rng rng;
    float x, y, ci;
    uint32_t fuse;
    int prev;
    walker_begin<EXT>(L.fv, a, walker, a.orbits ? a.orbits + wi : nullptr, rng, x, y, ci, fuse,
                      prev);
    for (uint32_t it = 0; it < a.ipt; it++) {
      int xi, texel = 0;
      int32_t px, py;
      int st = chaos_step<EXT>(L.fv, v, pal_size, a.burn_in, rng, x, y, ci, fuse, prev, xi, px,
                               py, texel);
      n_plotted += (st & (ST_RESPAWN | ST_BURN)) ? 0u : 1u;
      if (st & ST_PLOT) {
        n_in++;
        const size_t base = 4 * ((size_t)py * (size_t)v.width + (size_t)px);
        if (MODE == FFM_SPLAT_F32_GLOBAL) {
          const float* c = L.pal_f32 + 3 * texel;
          __hip_atomic_fetch_add(hist_f + base + 0, c[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_fetch_add(hist_f + base + 1, c[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_fetch_add(hist_f + base + 2, c[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_fetch_add(hist_f + base + 3, 100.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          const uint32_t pk = L.pal_u32[texel];
          constexpr int SCOPE =
              MODE == FFM_SPLAT_U32_XCD ? __HIP_MEMORY_SCOPE_WORKGROUP : __HIP_MEMORY_SCOPE_AGENT;
          __hip_atomic_fetch_add(hist_u + base + 0, pk & 0xFFu, __ATOMIC_RELAXED, SCOPE);
          __hip_atomic_fetch_add(hist_u + base + 1, (pk >> 8) & 0xFFu, __ATOMIC_RELAXED, SCOPE);
          __hip_atomic_fetch_add(hist_u + base + 2, (pk >> 16) & 0xFFu, __ATOMIC_RELAXED, SCOPE);
          __hip_atomic_fetch_add(hist_u + base + 3, 100u, __ATOMIC_RELAXED, SCOPE);
        }
      }
    }
    if (a.orbits) a.orbits[wi] = walker_orbit<EXT>(L.fv, x, y, ci, fuse, prev);
  }
  wave_add_u64(a.plotted, n_plotted);
  wave_add_u64(a.inframe, n_in);
}

// Per-step records of one walker (the parity trace): same step function.
template <bool EXT>
__global__ void trace_kernel(IterArgs a, uint32_t walker, uint32_t steps, uint32_t* out) {
  extern __shared__ uint32_t lds[];
  LdsFlame L = load_lds<false>(a, lds);
  if (threadIdx.x != 0) return;
  Rng rng;
  float x, y, ci;
  uint32_t fuse;
  int prev;
  walker_begin<EXT>(L.fv, a, walker, nullptr, rng, x, y, ci, fuse, prev);
  for (uint32_t s = 0; s < steps; s++) {
    int xi, texel = 0;
    int32_t px, py;
    int st = chaos_step<EXT>(L.fv, a.view, a.pal_size, a.burn_in, rng, x, y, ci, fuse, prev, xi,
                             px, py, texel);
    uint32_t* o = out + 8 * (size_t)s;
    o[0] = rng.state;
    o[1] = (uint32_t)xi;
    o[2] = __float_as_uint(x);
    o[3] = __float_as_uint(y);
    o[4] = __float_as_uint(ci);
    o[5] = (uint32_t)px;
    o[6] = (uint32_t)py;
    o[7] = (uint32_t)st;
  }
}

// Sum the per-XCD private copies into the histogram (coalesced uint4).
__global__ void fold_xcd_kernel(uint32_t* __restrict__ hist, const uint32_t* __restrict__ copies,
                                size_t words, int ncopies) {
  const size_t n4 = words / 4;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x) {
    uint4 acc = reinterpret_cast<const uint4*>(copies)[i];
    for (int c = 1; c < ncopies; c++) {
      uint4 v = reinterpret_cast<const uint4*>(copies + (size_t)c * words)[i];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
    reinterpret_cast<uint4*>(hist)[i] = acc;
  }
}

__global__ void clear_u32_kernel(uint4* __restrict__ p, size_t n4) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0, 0, 0, 0);
}

static size_t lds_bytes(const IterArgs& a, int nrows, bool f32pal) {
  size_t w = flame_table_floats(nrows) + (size_t)a.pal_size + (f32pal ? 3 * (size_t)a.pal_size : 0);
  return w * 4;
}

// A 128-transform flame with a 4096-texel palette needs more than the
// default 64 KB of dynamic LDS.
static hipError_t allow_lds(const void* fn) {
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, LDS_DYNAMIC_MAX);
}

template <bool EXT>
static hipError_t launch_iterate_t(const IterArgs& a, int splat_mode, int grid, int block,
                                   hipStream_t s) {
  static const hipError_t e0 =
      allow_lds(reinterpret_cast<const void*>(&iterate_kernel<FFM_SPLAT_U32_GLOBAL, EXT>));
  static const hipError_t e1 =
      allow_lds(reinterpret_cast<const void*>(&iterate_kernel<FFM_SPLAT_U32_XCD, EXT>));
  static const hipError_t e3 =
      allow_lds(reinterpret_cast<const void*>(&iterate_kernel<FFM_SPLAT_F32_GLOBAL, EXT>));
  if (e0 != hipSuccess || e1 != hipSuccess || e3 != hipSuccess) return hipErrorInvalidValue;
  size_t shm = lds_bytes(a, a.n_rows, splat_mode == FFM_SPLAT_F32_GLOBAL);
  switch (splat_mode) {
    case FFM_SPLAT_U32_GLOBAL:
      hipLaunchKernelGGL((iterate_kernel<FFM_SPLAT_U32_GLOBAL, EXT>), dim3(grid), dim3(block), shm,
                         s, a);
      break;
    case FFM_SPLAT_U32_XCD:
      hipLaunchKernelGGL((iterate_kernel<FFM_SPLAT_U32_XCD, EXT>), dim3(grid), dim3(block), shm,
                         s, a);
      break;
    case FFM_SPLAT_F32_GLOBAL:
      hipLaunchKernelGGL((iterate_kernel<FFM_SPLAT_F32_GLOBAL, EXT>), dim3(grid), dim3(block), shm,
                         s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_iterate(const IterArgs& a, int splat_mode, int grid, int block, hipStream_t s) {
  return a.ext ? launch_iterate_t<true>(a, splat_mode, grid, block, s)
               : launch_iterate_t<false>(a, splat_mode, grid, block, s);
}

template <bool EXT>
static hipError_t launch_trace_t(const IterArgs& a, uint32_t walker, uint32_t steps,
                                 uint32_t* out, hipStream_t s) {
  static const hipError_t e0 = allow_lds(reinterpret_cast<const void*>(&trace_kernel<EXT>));
  if (e0 != hipSuccess) return e0;
  size_t shm = lds_bytes(a, a.n_rows, false);
  hipLaunchKernelGGL(trace_kernel<EXT>, dim3(1), dim3(64), shm, s, a, walker, steps, out);
  return hipGetLastError();
}

hipError_t launch_trace(const IterArgs& a, uint32_t walker, uint32_t steps, uint32_t* out,
                        hipStream_t s) {
  return a.ext ? launch_trace_t<true>(a, walker, steps, out, s)
               : launch_trace_t<false>(a, walker, steps, out, s);
}

hipError_t launch_fold_xcd(uint32_t* hist, const uint32_t* copies, size_t words, int ncopies,
                           hipStream_t s) {
  size_t n4 = words / 4;
  int grid = (int)((n4 + 255) / 256);
  if (grid > 8192) grid = 8192;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(fold_xcd_kernel, dim3(grid), dim3(256), 0, s, hist, copies, words, ncopies);
  return hipGetLastError();
}

hipError_t launch_clear_u32(void* p, size_t words, hipStream_t s) {
  size_t n4 = words / 4;
  int grid = (int)((n4 + 255) / 256);
  if (grid > 8192) grid = 8192;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(clear_u32_kernel, dim3(grid), dim3(256), 0, s, reinterpret_cast<uint4*>(p), n4);
  return hipGetLastError();
}

}  // namespace ffm
