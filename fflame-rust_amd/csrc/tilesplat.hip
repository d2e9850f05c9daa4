// tilesplat.hip -- the LDS-privatised screen-tile splat (FFM_SPLAT_U32_LDS).
//
// The reference splats each in-frame sample with four u32 atomicAdd on the
// global histogram (shaders/core/main_template.wgsl:1253-1320). At 4K the
// histogram (132.7 MB) is far larger than any cache, so every sample costs
// a random read-modify-write of a 128-byte line in HBM. Here the splat is
// split in two streaming passes:
//
//  1. bin_iterate_kernel: the same walkers as iterate_kernel, but each
//     in-frame sample becomes a 4-byte record (pixel index inside its
//     128x64 screen tile | palette texel << 13 | tile-in-bin << 25) staged
//     in LDS with its rank among the stage's records of its bin (an LDS
//     atomic on the bin's count when it is staged). Every few iterations
//     the workgroup reserves a run in each bin's region of the record
//     buffer with ONE global atomic per (flush, bin) and writes every
//     staged record to its run at its rank. A bin is one tile up to
//     TILE_MAX tiles per image, else a strip of 2^k tiles of a row.
//  2. tile_accum_kernel: one workgroup per (bin, slice of its records)
//     accumulates the slice, tile by tile, into a 128x64 u32x4 tile held in
//     LDS (LDS atomics) and folds the tile into the histogram: a plain
//     read-add-write when it owns the whole bin, integer atomics when the
//     bin's records are split over several workgroups.
//
// Integer sums commute, so the histogram is bit-identical to the atomic
// splat's. Records that do not fit a tile's region (its capacity comes
// from a count-only pre-pass over a subset of the walkers) fall back to
// direct global atomics, so correctness never depends on the estimate.
#include "ffm_device.h"

namespace ffm {

constexpr int BIN_BLOCK = 1024;
// Staged records per workgroup: what the LDS left over by the flame table
// and palette holds, in steps of 512, up to STAGE_MAX (a rank within a
// bin stays below 2^16).
constexpr int STAGE_MIN = 10240, STAGE_MAX = 16384;
constexpr int MAX_FLUSH_EVERY = 8;  // iterations between fill checks, at most
constexpr uint32_t SLICE = 1u << 20;  // records per tile_accum work item (u32 tiles)
// f32 tiles (the 12288^2 stress render: bins of 16 tiles, counting-sorted per
// item) balance better over smaller items: 2^19 records take its
// accumulation 411 -> 398 ms a render, 2^22 and 2^24 lengthen it to 458 and
// 665 ms (the densest bins then sit in a few workgroups)
constexpr uint32_t SLICE_F32 = 1u << 19;

__host__ __device__ inline uint32_t slice_of(const TileArgs& t) { return t.f32 ? SLICE_F32 : SLICE; }
// accum_run adds two channels per 64-bit LDS atomic: a slice's channel sums
// must not carry out of 32 bits (at most 255, or 100 for density, a record)
static_assert((unsigned long long)SLICE * 255ull < (1ull << 32), "tile channel sums carry");
constexpr size_t STAGE_BYTES_PER_RECORD = 8;  // (record, bin | rank << 16)
// LDS besides the stage entries for nb bin slots: the fill words, then cnt,
// gbase (u32) and lim (u16) per bin
static size_t stage_fixed(int nb) { return (size_t)nb * (4 + 4 + 2) + 8; }

// Walker workgroups of WALK_LANES lanes, as many per CU as the LDS holds
// with a useful stage: three (52 KB each; 6 waves a SIMD, the occupancy the
// walker kernels are built for) when the flame table, palette and a stage of
// at least STAGE_THIRD_MIN records fit in a third of the CU's LDS, else two
// (78 KB, a stage of at least STAGE_HALF_MIN), else one of BIN_BLOCK lanes
// with the whole budget. The walk is latency-bound at 4 waves a SIMD: 3
// workgroups a CU walk single4k 1.2x faster than 2, and while one
// workgroup's flush waits at its barriers the others keep the SIMDs issuing.
// Every shape stages at least 6 x its lanes, so the fill check (at most
// every stage / lanes - 1 iterations) never lets the stage overflow.
constexpr int WALK_LANES = 512;
constexpr size_t LDS_THIRD = 52 * 1024;
constexpr size_t LDS_HALF = 78 * 1024;
constexpr int STAGE_THIRD_MIN = 3072;
constexpr int STAGE_HALF_MIN = 4096;
static_assert(STAGE_THIRD_MIN >= 6 * WALK_LANES && STAGE_MIN >= 6 * BIN_BLOCK / 2,
              "every walker shape stages several fill checks' worth of records");
// The walker kernels are compiled for 6 waves a SIMD (at most 80 VGPRs), so
// three workgroups a CU are resident together (the one-flame kernel strides
// its walkers over the whole grid). The batched kernel spills a few
// registers at that limit and still walks the gallery 1.13x faster than
// with two workgroups a CU (650 vs 578 flames/s).
#define FFM_WALKER_WAVES __attribute__((amdgpu_waves_per_eu(6)))

// Words of LDS before the stage: the flame table and the palette, rounded up
// to an even count so the stage's 8-byte entries are aligned.
__host__ __device__ inline size_t stage_base_words(int n_rows, int pal_size) {
  return (flame_table_floats(n_rows) + (size_t)pal_size + 1) & ~(size_t)1;
}

static size_t flame_lds(int n_rows, int pal_size) { return stage_base_words(n_rows, pal_size) * 4; }

static int stage_fit(size_t budget, size_t used) {
  const size_t fit = used < budget ? (budget - used) / STAGE_BYTES_PER_RECORD / 512 * 512 : 0;
  return (int)std::min<size_t>(fit, STAGE_MAX);
}

void bin_shape(int n_rows, int pal_size, int nb, int* block, int* stage, int* per_cu,
               int max_per_cu) {
  const size_t used = flame_lds(n_rows, pal_size) + stage_fixed(nb);
  const int third = stage_fit(LDS_THIRD, used);
  if (max_per_cu >= 3 && third >= STAGE_THIRD_MIN) {
    *block = WALK_LANES;
    *stage = third;
    *per_cu = 3;
    return;
  }
  const int half = stage_fit(LDS_HALF, used);
  if (half >= STAGE_HALF_MIN) {
    *block = WALK_LANES;
    *stage = half;
    *per_cu = 2;
    return;
  }
  *block = BIN_BLOCK;
  *stage = std::max(stage_fit((size_t)LDS_DYNAMIC_MAX, used), STAGE_MIN);
  *per_cu = 1;
}

size_t bin_lds_bytes(int n_rows, int pal_size, int stage, int nb) {
  return flame_lds(n_rows, pal_size) + stage_fixed(nb) + (size_t)stage * STAGE_BYTES_PER_RECORD;
}

size_t accum_lds_bytes(int pal_size, bool f32, int tile_px) {
  return (size_t)tile_px * 16 + (size_t)pal_size * (f32 ? 12 : 4) + 2 * 129 * 4;
}

// Exclusive scan of a[0..n) (n <= 2 * blockDim.x) in place; all threads.
__device__ void block_exclusive_scan(uint32_t* a, int n, uint32_t* wave_sums, uint32_t* total) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  const int i0 = 2 * tid, i1 = i0 + 1;
  const uint32_t v0 = i0 < n ? a[i0] : 0u, v1 = i1 < n ? a[i1] : 0u;
  const uint32_t s = v0 + v1;
  uint32_t x = s;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wave_sums[wave] = x;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < nw ? wave_sums[lane] : 0u;
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t y = __shfl_up(w, off, 64);
      if (lane >= off) w += y;
    }
    if (lane < nw) wave_sums[lane] = w;
    if (lane == nw - 1 && total) *total = w;
  }
  __syncthreads();
  const uint32_t base = (wave > 0 ? wave_sums[wave - 1] : 0u) + x - s;
  if (i0 < n) a[i0] = base;
  if (i1 < n) a[i1] = base + v0;
  __syncthreads();
}

// The stage holds (record, bin | rank among the stage's records of the bin
// << 16) per staged sample: as one 8-byte entry (WIDE: one LDS write a
// record; the one-flame walk is 0.6% faster so) or as two arrays of u32
// (the batched walk: 5% faster so than with 8-byte entries, measured with
// the same tuning and the same instruction counts otherwise).
struct BinLds {
  uint32_t* fill;   // 2: the stage's fill, then one word of scan total
  uint2* st;        // WIDE stage entries
  uint32_t* rec;    // the stage's records, when not WIDE
  uint32_t* tr;     // and their bin | rank << 16
  uint32_t* cnt;    // bins_lds: the stage's records per bin (COUNT_ONLY: the walk's)
  uint32_t* gbase;  // bins_lds: the low 32 bits of the bin's first record slot of this flush
  uint16_t* lim;    // bins_lds: slots left in the bin's region (at most 2^15 - 1), bit 15 the
                    // slot's bit 32
};

// Tile column of a bin's sub-tile.
__device__ __forceinline__ int bin_tile_x(const TileArgs& t, int bin, int sub) {
  return ((bin % t.bins_x) << t.group_shift) + sub;
}

__device__ __forceinline__ uint32_t rec_texel(uint32_t rec) {
  return (rec >> REC_TEXEL_SHIFT) & REC_TEXEL_MASK;
}

// Splat one record straight into the histogram (region overflow).
__device__ __forceinline__ void overflow_splat(const TileArgs& t, const uint32_t* pal, int bin,
                                               uint32_t rec) {
  const int tx = bin_tile_x(t, bin, (int)(rec >> REC_SUB_SHIFT)), ty = bin / t.bins_x;
  const int local = (int)(rec & REC_LOCAL_MASK);
  const int px = (tx << t.tile_wl) + (local & ((1 << t.tile_wl) - 1));
  const int py = (ty << t.tile_hl) + (local >> t.tile_wl);
  const size_t base = 4 * ((size_t)py * t.width + px);
  if (t.f32) {
    const float* c = t.pal_f32 + 3 * rec_texel(rec);
    float* hf = reinterpret_cast<float*>(t.hist) + base;
    __hip_atomic_fetch_add(hf + 0, c[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(hf + 1, c[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(hf + 2, c[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(hf + 3, 100.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const uint32_t pk = pal[rec_texel(rec)];
  uint32_t* h = t.hist + base;
  __hip_atomic_fetch_add(h + 0, pk & 0xFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_fetch_add(h + 1, (pk >> 8) & 0xFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_fetch_add(h + 2, (pk >> 16) & 0xFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_fetch_add(h + 3, 100u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Write the staged records to their bins' regions. Called by every thread
// of the workgroup, after a barrier that follows the last staging.
template <bool WIDE>
__device__ void flush_stage(const TileArgs& t, const BinLds& L, const uint32_t* pal) {
  const int tid = threadIdx.x;
  const int n = (int)min(*L.fill, (uint32_t)t.stage);
  // reserve each bin's run, and keep where it starts and how much of it
  // the region holds, so the record loop below reads LDS only; the counts
  // start over for the next stage
  for (int i = tid; i < t.n_tiles; i += blockDim.x) {
    const uint32_t c = L.cnt[i];
    uint64_t base = 0u;
    uint32_t room = 0u;
    if (c) {
      const uint32_t old =
          __hip_atomic_fetch_add(t.tile_fill + i, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t cap = t.cap[i];
      if (cap > old) {
        base = t.region[i] + old;
        room = min(cap - old, 0x7FFFu);
      }
      L.cnt[i] = 0u;
    }
    L.gbase[i] = (uint32_t)base;
    L.lim[i] = (uint16_t)(room | (uint32_t)(base >> 32) << 15);
  }
  __syncthreads();
  // ranks below 2^14 (the stage), so `rank < room` is `fill + rank < cap`
  for (int j = tid; j < n; j += blockDim.x) {
    const uint2 e = WIDE ? L.st[j] : make_uint2(L.rec[j], L.tr[j]);
    const int tile = (int)(e.y & 0xFFFFu);
    const uint32_t rank = e.y >> 16;
    const uint32_t lim = L.lim[tile];
    if (rank < (lim & 0x7FFFu))
      t.records[((size_t)(lim >> 15) << 32) + L.gbase[tile] + rank] = e.x;
    else
      overflow_splat(t, pal, tile, e.x);
  }
  __syncthreads();
  if (tid == 0) *L.fill = 0u;
  __syncthreads();
}

// Walk walkers base + rd * stride + tid (< limit) for a.ipt iterations,
// staging every in-frame sample (or, COUNT_ONLY, counting it per tile).
// Rounds are uniform across the workgroup so every barrier is reached by all
// threads; lanes past `limit` step nothing.
template <bool COUNT_ONLY, bool EXT, bool WIDE>
__device__ void run_walkers(const IterArgs& a, const TileArgs& t, const LdsFlame& F,
                            const BinLds& L, uint64_t base, uint64_t stride, uint64_t limit,
                            unsigned long long& n_plotted, unsigned long long& n_in) {
  const int tid = threadIdx.x;
  // Uniform values that only vector instructions read, on every plotted
  // iteration, held in vector registers: the scalar file is full, and a
  // value spilled from it costs a v_readlane at every use.
  ViewConst v = a.view;
  uint32_t bins_x = (uint32_t)t.bins_x, hmask = (1u << t.tile_hl) - 1u,
           wmask = (1u << t.tile_wl) - 1u, gmask = (1u << t.group_shift) - 1u;
  FlameView fv = F.fv;
  asm volatile("" : "+v"(v.cy), "+v"(v.width), "+v"(v.height), "+v"(bins_x), "+v"(hmask),
               "+v"(wmask), "+v"(gmask), "+v"(fv.sel_byte));
  const uint64_t rounds = limit > base ? (limit - base + stride - 1) / stride : 0;
  // a fill check every `every` iterations; at most `every` x blockDim
  // records arrive between checks, so flushing at stage - that never
  // overflows (every is capped so that stays positive: the stage holds at
  // least 6 x blockDim records)
  const uint32_t every = (uint32_t)max(
      min(min(max(t.flush_every, 1), MAX_FLUSH_EVERY), t.stage / (int)blockDim.x - 1), 1);
  const uint32_t flush_at = (uint32_t)t.stage - every * blockDim.x;
  for (uint64_t rd = 0; rd < rounds; rd++) {
    const uint64_t wi = base + rd * stride + tid;
    const bool active = wi < limit;
    const uint32_t walker = a.walker0 + (uint32_t)wi;
    Rng rng;
    float x, y, ci;
    uint32_t fuse;
    int prev;
    walker_begin<EXT>(fv, a, walker, active && a.orbits ? a.orbits + wi : nullptr, rng, x, y,
                      ci, fuse, prev);
    uint32_t r_plotted = 0u, r_in = 0u;  // this round's counts (at most ipt each)
    uint32_t to_check = every;           // iterations left until the next fill check
    for (uint32_t it = 0; it < a.ipt; it++) {
      // the pass-sharing bits re-tested every iteration (a mask derived once
      // would hold a scalar pair over the whole loop)
      asm volatile("" : "+s"(fv.share_tp));
      if (active) {
        int xi, texel = 0;
        int32_t px, py;
        int st = chaos_step<EXT>(fv, v, a.pal_size, a.burn_in, rng, x, y, ci, fuse, prev, xi,
                                 px, py, texel);
        r_plotted += (st & (ST_RESPAWN | ST_BURN)) ? 0u : 1u;
        if (st & ST_PLOT) {
          r_in++;
          // in-frame: px, py >= 0, so tile arithmetic is unsigned shifts and masks
          const uint32_t ux = (uint32_t)px, uy = (uint32_t)py, tx = ux >> t.tile_wl;
          const uint32_t tile = __umul24(uy >> t.tile_hl, bins_x) + (tx >> t.group_shift);
          const uint32_t r = atomicAdd(&L.cnt[tile], 1u);  // the record's rank in its bin
          if (!COUNT_ONLY) {
            const uint32_t local =
                (uy & hmask) << t.tile_wl | (ux & wmask) | (tx & gmask) << REC_SUB_SHIFT;
            // the append to the stage: the compiler's atomic optimizer makes
            // the wave's increments of one address one LDS atomic by its
            // first active lane, each lane taking base + its rank in the wave
            const uint32_t b = atomicAdd(L.fill, 1u);
            const uint32_t rec = local | ((uint32_t)texel << REC_TEXEL_SHIFT), tr = tile | r << 16;
            if (WIDE) {
              L.st[b] = make_uint2(rec, tr);
            } else {
              L.rec[b] = rec;
              L.tr[b] = tr;
            }
          }
        }
      }
      if (!COUNT_ONLY && (--to_check == 0u || it + 1 == a.ipt)) {
        to_check = every;
        __syncthreads();
        const uint32_t f = *L.fill;
        __syncthreads();
        if (f >= flush_at) flush_stage<WIDE>(t, L, F.pal_u32);
      }
    }
    n_plotted += r_plotted;
    n_in += r_in;
    if (!COUNT_ONLY && active && a.orbits)
      a.orbits[wi] = walker_orbit<EXT>(fv, x, y, ci, fuse, prev);
  }
}

// The staging layout for `stage` records (a multiple of 512) and `nb` bin
// slots (a multiple of 64).
template <bool WIDE>
__device__ __forceinline__ BinLds bin_lds_views(uint32_t* p, int stage, int nb) {
  BinLds L;
  L.fill = p;
  L.st = reinterpret_cast<uint2*>(p + 2);
  L.rec = p + 2;
  L.tr = p + 2 + stage;
  L.cnt = p + 2 + 2 * stage;
  L.gbase = L.cnt + nb;
  L.lim = reinterpret_cast<uint16_t*>(L.gbase + nb);
  return L;
}

// Finish a workgroup's walking: publish per-tile counts (COUNT_ONLY) or
// flush what is still staged. All threads.
template <bool COUNT_ONLY, bool WIDE>
__device__ void finish_walkers(const TileArgs& t, const BinLds& L, const uint32_t* pal) {
  const int tid = threadIdx.x;
  __syncthreads();
  if (COUNT_ONLY) {
    for (int i = tid; i < t.n_tiles; i += blockDim.x) {
      if (L.cnt[i])
        __hip_atomic_fetch_add(t.tile_fill + i, L.cnt[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      L.cnt[i] = 0u;
    }
    __syncthreads();
  } else {
    if (*L.fill > 0u) flush_stage<WIDE>(t, L, pal);
  }
}

// One dispatch of one flame: walkers grid-strided over the whole grid.
template <bool COUNT_ONLY, bool EXT>
__global__ __launch_bounds__(BIN_BLOCK) FFM_WALKER_WAVES void bin_iterate_kernel(IterArgs a,
                                                                                 TileArgs t) {
  extern __shared__ __align__(16) uint32_t lds[];
  LdsFlame F = load_lds<false>(a, lds);
  const BinLds L =
      bin_lds_views<true>(lds + stage_base_words(a.n_rows, a.pal_size), t.stage, t.bins_lds);
  const int tid = threadIdx.x;
  for (int i = tid; i < t.n_tiles; i += blockDim.x) L.cnt[i] = 0u;
  if (tid == 0) *L.fill = 0u;
  __syncthreads();
  unsigned long long n_plotted = 0, n_in = 0;
  run_walkers<COUNT_ONLY, EXT, true>(a, t, F, L, (uint64_t)blockIdx.x * blockDim.x,
                               (uint64_t)gridDim.x * blockDim.x, a.walkers, n_plotted, n_in);
  finish_walkers<COUNT_ONLY, true>(t, L, F.pal_u32);
  if (!COUNT_ONLY) {
    for (int off = 32; off > 0; off >>= 1) {
      n_plotted += __shfl_xor(n_plotted, off, 64);
      n_in += __shfl_xor(n_in, off, 64);
    }
    if ((tid & 63) == 0) {
      if (a.plotted && n_plotted)
        __hip_atomic_fetch_add(a.plotted, n_plotted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (a.inframe && n_in)
        __hip_atomic_fetch_add(a.inframe, n_in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Exclusive scan of value(i), i < n, into out[] by one workgroup, 2048 at a
// time with a running carry; returns the total. All threads.
template <typename Fn>
__device__ uint32_t block_scan_n(int n, Fn value, uint32_t* out, uint32_t* buf, uint32_t* wsum,
                                 uint32_t* total) {
  uint32_t carry = 0;
  for (int c0 = 0; c0 < n; c0 += 2 * BIN_BLOCK) {
    const int m = min(2 * BIN_BLOCK, n - c0);
    for (int i = threadIdx.x; i < m; i += blockDim.x) buf[i] = value(c0 + i);
    __syncthreads();
    block_exclusive_scan(buf, m, wsum, total);
    for (int i = threadIdx.x; i < m; i += blockDim.x) out[c0 + i] = buf[i] + carry;
    carry += *total;
    __syncthreads();
  }
  return carry;
}

// Capacities from the pre-pass counts: cap = est * scale * 1.25 + margin,
// regions laid out back to back (64-bit starts: a buffer may pass 2^32
// records), scaled down to fit `buffer_records` if needed; then the fill
// counters are cleared for the real pass. One workgroup; any number of tiles.
__global__ __launch_bounds__(BIN_BLOCK) void tile_capacity_kernel(TileArgs t, float scale,
                                                                  uint32_t margin,
                                                                  uint64_t buffer_records) {
  __shared__ unsigned long long part[BIN_BLOCK];
  __shared__ unsigned long long sum64;
  const int tid = threadIdx.x;
  auto cap_of = [&](int i) -> uint32_t {
    const uint32_t f = t.tile_fill[i];
    return f ? (uint32_t)fminf((float)f * scale * 1.25f + (float)margin, 2.0e9f) : margin / 4u;
  };
  if (tid == 0) sum64 = 0ull;
  __syncthreads();
  unsigned long long mine = 0;
  for (int i = tid; i < t.n_tiles; i += blockDim.x) mine += cap_of(i);
  atomicAdd(&sum64, mine);
  __syncthreads();
  const double f = sum64 > buffer_records ? (double)buffer_records / (double)sum64 : 1.0;
  auto scaled = [&](int i) -> uint32_t {
    const uint32_t c = cap_of(i);
    return f < 1.0 ? (uint32_t)((double)c * f) : c;
  };
  // exclusive scan of the capacities: each thread a contiguous run of
  // tiles, the runs' sums scanned across the workgroup
  const int per = (t.n_tiles + (int)blockDim.x - 1) / (int)blockDim.x;
  const int i0 = min(tid * per, t.n_tiles), i1 = min(i0 + per, t.n_tiles);
  unsigned long long run = 0ull;
  for (int i = i0; i < i1; i++) {
    const uint32_t c = scaled(i);
    t.cap[i] = c;
    run += c;
  }
  part[tid] = run;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    const unsigned long long v = tid >= off ? part[tid - off] : 0ull;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  run = tid > 0 ? part[tid - 1] : 0ull;
  for (int i = i0; i < i1; i++) {
    t.region[i] = run;
    run += t.cap[i];
  }
  for (int i = tid; i < t.n_tiles; i += blockDim.x) t.tile_fill[i] = 0u;
}

// Work list for the tile accumulation: tile i gets ceil(stored_i / SLICE)
// items; item_start[n_tiles] is the total.
__global__ __launch_bounds__(BIN_BLOCK) void tile_plan_kernel(TileArgs t) {
  __shared__ uint32_t buf[2 * BIN_BLOCK];
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t total;
  const uint32_t all = block_scan_n(
      t.n_tiles,
      [&](int i) { return (min(t.tile_fill[i], t.cap[i]) + slice_of(t) - 1) / slice_of(t); },
      t.item_start, buf, wsum, &total);
  if (threadIdx.x == 0) t.item_start[t.n_tiles] = all;
}

// A u32 record's (r, g) and (b, density) channel pairs, each pair one
// little-endian u64 of the tile's u32 quad.
__device__ __forceinline__ void rec_channels(uint32_t pk, unsigned long long& rg,
                                             unsigned long long& bd) {
  rg = (unsigned long long)(pk & 0xFFu) | (unsigned long long)((pk >> 8) & 0xFFu) << 32;
  bd = (unsigned long long)((pk >> 16) & 0xFFu) | 100ull << 32;
}

// LDS tile layouts, blocked so that a wave's atomics on scattered pixels
// spread over all 64 LDS banks (an interleaved u32x4 or float4 per pixel
// puts every pixel's first channel in one of 16 banks):
// u32: blocks of 32 pixels, each the 32 (r, g) u64 pairs then the 32
//      (b, density) pairs; the (r, g) pair of pixel px is u64 rg_slot(px),
//      its (b, density) pair 32 further;
// f32: blocks of 64 pixels, each 64 floats per channel in r, g, b, density
//      order; channel c of pixel px is float f32_slot(px) + 64 c.
__device__ __forceinline__ uint32_t rg_slot(uint32_t px) { return (px >> 5 << 6) | (px & 31u); }
__device__ __forceinline__ uint32_t f32_slot(uint32_t px) { return (px >> 6 << 8) | (px & 63u); }

// Add one record into the LDS tile (LDS atomics).
template <bool F32>
__device__ __forceinline__ void accum_record(uint32_t v, uint32_t* tile, const uint32_t* pal,
                                             const float* palf) {
  float* tilef = reinterpret_cast<float*>(tile);
  {
    const uint32_t local = v & REC_LOCAL_MASK;
    if (F32) {
      const float* c = palf + 3 * rec_texel(v);
      float* q = tilef + f32_slot(local);
      atomicAdd(q + 0, c[0]);
      atomicAdd(q + 64, c[1]);
      atomicAdd(q + 128, c[2]);
      atomicAdd(q + 192, 100.0f);
    } else {
      // Two 64-bit LDS atomics add (r, g) and (b, density) to the pixel's
      // little-endian u32 quad. No carry crosses a channel: a run holds at
      // most SLICE = 2^20 records and a channel adds at most 255 a record
      // (2^28 in all), so each 32-bit half stays below 2^32.
      unsigned long long* t64 = reinterpret_cast<unsigned long long*>(tile);
      unsigned long long rg, bd;
      rec_channels(pal[rec_texel(v)], rg, bd);
      unsigned long long* q = t64 + rg_slot(local);
      atomicAdd(q, rg);
      atomicAdd(q + 32, bd);
    }
  }
}

// Add records run[0..n) into the LDS tile. All threads. Eight records a
// thread per round, their loads issued together, so each lane keeps eight
// reads in flight instead of waiting out one at a time (integer sums
// commute: the order of the atomics does not change the tile).
template <bool F32>
__device__ __forceinline__ void accum_run(const uint32_t* run, uint32_t n, uint32_t* tile,
                                          const uint32_t* pal, const float* palf) {
  const uint32_t bd = blockDim.x;
  uint32_t r = threadIdx.x;
  for (; r + 7 * bd < n; r += 8 * bd) {
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = run[r + k * bd];
    // a round whose lanes never repeat a pixel between their first two
    // records (nearly every round of a spread attractor) adds record by record
    if (F32 || !__ballot((v[0] & REC_LOCAL_MASK) == (v[1] & REC_LOCAL_MASK))) {
#pragma unroll
      for (int k = 0; k < 8; k++) accum_record<F32>(v[k], tile, pal, palf);
    } else {
      // A lane sums its eight records of the round in registers while they
      // hit the same pixel, and adds each such run with one pair of atomics.
      // On an attractor that puts most samples in a few pixels (some
      // gallery flames hold 90% in one), every lane of a wave would
      // otherwise hit the same LDS address with every atomic, which the
      // hardware serialises; elsewhere a run is one record and the atomics
      // are those of one record at a time.
      unsigned long long* t64 = reinterpret_cast<unsigned long long*>(tile);
      uint32_t cur = v[0] & REC_LOCAL_MASK;
      unsigned long long rg, bdn;
      rec_channels(pal[rec_texel(v[0])], rg, bdn);
#pragma unroll
      for (int k = 1; k < 8; k++) {
        const uint32_t key = v[k] & REC_LOCAL_MASK;
        unsigned long long rg1, bd1;
        rec_channels(pal[rec_texel(v[k])], rg1, bd1);
        if (key == cur) {
          rg += rg1;
          bdn += bd1;
        } else {
          atomicAdd(t64 + rg_slot(cur), rg);
          atomicAdd(t64 + rg_slot(cur) + 32, bdn);
          cur = key;
          rg = rg1;
          bdn = bd1;
        }
      }
      atomicAdd(t64 + rg_slot(cur), rg);
      atomicAdd(t64 + rg_slot(cur) + 32, bdn);
    }
  }
  for (; r < n; r += bd) accum_record<F32>(run[r], tile, pal, palf);
}

// Fold the LDS tile (tile column tx, row ty) into the histogram: a
// read-add-write when the slice is its bin's only one, atomics otherwise.
template <bool F32>
__device__ __forceinline__ void fold_tile(const TileArgs& t, int tx, int ty, uint32_t n_slices,
                                          const uint32_t* tile) {
  const float* tilef = reinterpret_cast<const float*>(tile);
  const int wmask = (1 << t.tile_wl) - 1;
  for (int q = threadIdx.x; q < 1 << (t.tile_wl + t.tile_hl); q += blockDim.x) {
    const int px = (tx << t.tile_wl) + (q & wmask), py = (ty << t.tile_hl) + (q >> t.tile_wl);
    if (px >= t.width || py >= t.height) continue;
    const uint32_t rgw = 2 * rg_slot((uint32_t)q), fq = f32_slot((uint32_t)q);
    // no hit: a zero density (0u is +0.0f too)
    if ((F32 ? tile[fq + 192] : tile[rgw + 65]) == 0u) continue;
    const size_t p = (size_t)py * t.width + px;
    if (F32) {
      const float4 v = make_float4(tilef[fq], tilef[fq + 64], tilef[fq + 128], tilef[fq + 192]);
      float4* h = reinterpret_cast<float4*>(t.hist) + p;
      if (n_slices == 1) {
        float4 o = *h;
        o.x += v.x;
        o.y += v.y;
        o.z += v.z;
        o.w += v.w;
        *h = o;
      } else {
        float* hw = reinterpret_cast<float*>(h);
        __hip_atomic_fetch_add(hw + 0, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(hw + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(hw + 2, v.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(hw + 3, v.w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else {
      const uint4 v = make_uint4(tile[rgw], tile[rgw + 1], tile[rgw + 64], tile[rgw + 65]);
      uint4* h = reinterpret_cast<uint4*>(t.hist) + p;
      if (n_slices == 1) {
        uint4 o = *h;
        o.x += v.x;
        o.y += v.y;
        o.z += v.z;
        o.w += v.w;
        *h = o;
      } else {
        uint32_t* hw = reinterpret_cast<uint32_t*>(h);
        __hip_atomic_fetch_add(hw + 0, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(hw + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(hw + 2, v.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(hw + 3, v.w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// Accumulate records [r0, r1) of bin `ti` (of the flame `t` describes) and
// fold them into the histogram, tile by tile. A bin of several tiles first
// counting-sorts the slice by tile into t.sorted (same offsets as the
// records), so each tile's records are one contiguous run read once.
// `sort` holds 2 x 129 u32 of LDS. All threads; F32: float4 accumulation.
template <bool F32>
__device__ void accum_slice(const TileArgs& t, int ti, uint32_t r0, uint32_t r1, uint32_t n_slices,
                            uint32_t* tile, const uint32_t* pal, const float* palf,
                            uint32_t* sort) {
  const int tid = threadIdx.x;
  const uint32_t* rec = t.records + t.region[ti] + r0;
  const uint32_t n = r1 - r0;
  const int ty = ti / t.bins_x;
  const int group = 1 << t.group_shift;
  uint32_t* fill = sort;         // group: counts, then next free slot per tile
  uint32_t* start = sort + 129;  // group + 1: first slot per tile
  const uint32_t* runs = rec;
  if (group > 1) {
    for (int i = tid; i < group; i += blockDim.x) fill[i] = 0u;
    __syncthreads();
    for (uint32_t r = tid; r < n; r += blockDim.x) atomicAdd(&fill[rec[r] >> REC_SUB_SHIFT], 1u);
    __syncthreads();
    if (tid == 0) {
      uint32_t acc = 0;
      for (int g = 0; g < group; g++) {
        start[g] = acc;
        acc += fill[g];
        fill[g] = start[g];
      }
      start[group] = acc;
    }
    __syncthreads();
    uint32_t* out = t.sorted + t.region[ti] + r0;
    for (uint32_t r = tid; r < n; r += blockDim.x) {
      const uint32_t v = rec[r];
      out[atomicAdd(&fill[v >> REC_SUB_SHIFT], 1u)] = v;
    }
    __syncthreads();  // the workgroup's global stores are visible to it after the barrier
    runs = out;
  } else if (tid == 0) {
    start[0] = 0u;
    start[1] = n;
  }
  for (int sub = 0; sub < group; sub++) {
    const int tx = bin_tile_x(t, ti, sub);
    if (tx >= t.tiles_x) break;
    for (int i = tid; i < 4 << (t.tile_wl + t.tile_hl); i += blockDim.x) tile[i] = 0u;
    __syncthreads();
    const uint32_t s0 = start[sub], s1 = start[sub + 1];
    if (s1 > s0) {  // uniform
      accum_run<F32>(runs + s0, s1 - s0, tile, pal, palf);
      __syncthreads();
      fold_tile<F32>(t, tx, ty, n_slices, tile);
    }
    __syncthreads();
  }
}

// The largest i < n with start[i] <= x (start non-decreasing). Uniform.
__device__ __forceinline__ int find_range(const uint32_t* start, int n, uint32_t x) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= x) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One flame's tiles: persistent grid over the work list.
template <bool F32>
__global__ __launch_bounds__(BIN_BLOCK) void tile_accum_kernel(TileArgs t, const uint32_t* pal_g,
                                                               int pal_size) {
  extern __shared__ uint32_t lds[];
  uint32_t* tile = lds;                                 // tile pixels x 4 (u32 or f32 bits)
  uint32_t* pal = lds + (4 << (t.tile_wl + t.tile_hl));  // pal_size u32, or 3 x pal_size f32
  float* palf = reinterpret_cast<float*>(pal);
  uint32_t* sort = pal + (F32 ? 3 : 1) * pal_size;  // 2 x 129
  const int tid = threadIdx.x;
  if (F32)
    for (int i = tid; i < 3 * pal_size; i += blockDim.x) palf[i] = t.pal_f32[i];
  else
    for (int i = tid; i < pal_size; i += blockDim.x) pal[i] = pal_g[i];
  const uint32_t n_items = t.item_start[t.n_tiles];
  for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int ti = find_range(t.item_start, t.n_tiles, item);
    const uint32_t first = t.item_start[ti], n_slices = t.item_start[ti + 1] - first;
    const uint32_t stored = min(t.tile_fill[ti], t.cap[ti]);
    const uint32_t sl = slice_of(t);
    const uint32_t r0 = (item - first) * sl, r1 = min(r0 + sl, stored);
    accum_slice<F32>(t, ti, r0, r1, n_slices, tile, pal, palf, sort);
  }
}

// ------------------------------------------------------------ batches

// A copy of *p read through the constant address space: with a uniform p,
// scalar loads into SGPRs (a global load would hold the words in VGPRs).
template <class T>
__device__ __forceinline__ T scalar_copy(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "whole words");
  typedef const __attribute__((address_space(4))) uint32_t KWord;
  T v;
  uint32_t* d = reinterpret_cast<uint32_t*>(&v);
  for (int i = 0; i < (int)(sizeof(T) / 4); i++) d[i] = ((KWord*)p)[i];
  return v;
}

// A flame's view of the global tile arrays.
__device__ __forceinline__ TileArgs flame_tiles(const BatchArgs& b, const BatchFlame& f) {
  TileArgs t = b.g;
  t.tile_fill += f.tile_base;
  t.region += f.tile_base;
  t.cap += f.tile_base;
  t.hist = b.g.hist + 4 * f.hist_off;
  t.tiles_x = f.tiles_x;
  t.n_tiles = f.n_tiles;
  t.bins_x = f.tiles_x;  // batched flames have at most TILE_MAX tiles: one tile per bin
  t.group_shift = 0;
  t.width = f.view.width;
  t.height = f.view.height;
  return t;
}

// Many flames in one persistent launch: each workgroup takes work items
// (flame, walker range), loads that flame's table and palette into LDS and
// walks the range with the same staging and flushes as one flame.
template <bool COUNT_ONLY, bool EXT>
__global__ __launch_bounds__(BIN_BLOCK) FFM_WALKER_WAVES void bin_batch_kernel(BatchArgs b) {
  extern __shared__ __align__(16) uint32_t lds[];
  const BinLds L =
      bin_lds_views<false>(lds + stage_base_words(b.max_rows, b.max_pal), b.g.stage, b.g.bins_lds);
  const int tid = threadIdx.x;
  for (int i = tid; i < b.g.bins_lds; i += blockDim.x) L.cnt[i] = 0u;
  if (tid == 0) *L.fill = 0u;
  __syncthreads();
  unsigned long long n_plotted = 0, n_in = 0;
  __shared__ int next;
  for (;;) {
    // dynamic work stealing: items differ in cost (transform counts, sizes)
    if (tid == 0) next = (int)atomicAdd(b.next_item, 1u);
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(next);
    __syncthreads();
    if (item >= b.n_items) break;
    // the item and its flame by scalar loads: held in VGPRs, their thirty
    // words would push the walk into scratch
    const BatchItem it = scalar_copy(b.items + item);
    const BatchFlame f = scalar_copy(b.flames + it.flame);
    const LdsFlame F = load_table<false>(b.tables + f.table_off, b.pal_u32 + f.pal_off, nullptr,
                                         f.pal_size, lds);
    IterArgs a{};
    a.view = f.view;
    a.pal_size = f.pal_size;
    a.ipt = f.ipt;
    a.burn_in = f.burn_in;
    a.seed = f.seed;
    const TileArgs t = flame_tiles(b, f);
    uint64_t limit = it.w1;
    if (COUNT_ONLY) {
      limit = it.w0 + max((it.w1 - it.w0) / b.count_div, 1u);
      // the plotted share of the walk, shortened: the tiles' shares of an
      // attractor's samples do not depend on how long it is walked
      if (f.ipt > f.burn_in)
        a.ipt = f.burn_in + (f.ipt - f.burn_in + b.count_iter_div - 1) / b.count_iter_div;
    }
    run_walkers<COUNT_ONLY, EXT, false>(a, t, F, L, it.w0, blockDim.x, limit, n_plotted, n_in);
    finish_walkers<COUNT_ONLY, false>(t, L, F.pal_u32);
    __syncthreads();  // the next item reloads the flame table
  }
}

// Tile accumulation over every flame's tiles: global tile -> its flame by
// the flames' tile bases; the palette is reloaded when the flame changes.
__global__ __launch_bounds__(BIN_BLOCK) void tile_accum_batch_kernel(BatchArgs b) {
  extern __shared__ uint32_t lds[];
  uint32_t* tile = lds;
  uint32_t* pal = lds + (4 << (b.g.tile_wl + b.g.tile_hl));
  const int tid = threadIdx.x;
  const TileArgs& g = b.g;
  const uint32_t n_items = g.item_start[g.n_tiles];
  int cur = -1;
  for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int gt = find_range(g.item_start, g.n_tiles, item);
    int lo = 0, hi = b.n_flames - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (b.flames[mid].tile_base <= (uint32_t)gt) lo = mid; else hi = mid - 1;
    }
    const BatchFlame f = b.flames[lo];
    if (lo != cur) {
      for (int i = tid; i < f.pal_size; i += blockDim.x) pal[i] = b.pal_u32[f.pal_off + i];
      __syncthreads();
      cur = lo;
    }
    const TileArgs t = flame_tiles(b, f);
    const int ti = gt - (int)f.tile_base;
    const uint32_t first = g.item_start[gt], n_slices = g.item_start[gt + 1] - first;
    const uint32_t stored = min(g.tile_fill[gt], g.cap[gt]);
    const uint32_t r0 = (item - first) * SLICE, r1 = min(r0 + SLICE, stored);
    accum_slice<false>(t, ti, r0, r1, n_slices, tile, pal, nullptr, pal + b.max_pal);
  }
}

// Kernels here use more than the default 64 KB of dynamic LDS.
static hipError_t allow_lds(const void* fn) {
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, LDS_DYNAMIC_MAX);
}

template <bool COUNT_ONLY, bool EXT>
static hipError_t launch_bin_iterate_t(const IterArgs& a, const TileArgs& t, int grid,
                                       hipStream_t s) {
  static const hipError_t e =
      allow_lds(reinterpret_cast<const void*>(&bin_iterate_kernel<COUNT_ONLY, EXT>));
  if (e != hipSuccess) return e;
  const int blk = t.walk_block;
  const size_t shm = bin_lds_bytes(a.n_rows, a.pal_size, t.stage, t.bins_lds);
  hipLaunchKernelGGL((bin_iterate_kernel<COUNT_ONLY, EXT>), dim3(grid * t.walk_per_cu), dim3(blk),
                     shm, s, a, t);
  return hipGetLastError();
}

hipError_t launch_bin_iterate(const IterArgs& a, const TileArgs& t, bool count_only, int grid,
                              hipStream_t s) {
  if (count_only)
    return a.ext ? launch_bin_iterate_t<true, true>(a, t, grid, s)
                 : launch_bin_iterate_t<true, false>(a, t, grid, s);
  return a.ext ? launch_bin_iterate_t<false, true>(a, t, grid, s)
               : launch_bin_iterate_t<false, false>(a, t, grid, s);
}

template <bool COUNT_ONLY, bool EXT>
static hipError_t launch_bin_batch_t(const BatchArgs& b, int grid, hipStream_t s) {
  static const hipError_t e =
      allow_lds(reinterpret_cast<const void*>(&bin_batch_kernel<COUNT_ONLY, EXT>));
  if (e != hipSuccess) return e;
  const int blk = b.g.walk_block;
  const size_t shm = bin_lds_bytes(b.max_rows, b.max_pal, b.g.stage, b.g.bins_lds);
  hipLaunchKernelGGL((bin_batch_kernel<COUNT_ONLY, EXT>), dim3(grid * b.g.walk_per_cu),
                     dim3(blk), shm, s, b);
  return hipGetLastError();
}

hipError_t launch_bin_batch(const BatchArgs& b, bool count_only, int grid, hipStream_t s) {
  if (count_only)
    return b.ext ? launch_bin_batch_t<true, true>(b, grid, s)
                 : launch_bin_batch_t<true, false>(b, grid, s);
  return b.ext ? launch_bin_batch_t<false, true>(b, grid, s)
               : launch_bin_batch_t<false, false>(b, grid, s);
}

hipError_t launch_tile_capacity_n(const TileArgs& t, float scale, uint32_t margin,
                                  uint64_t buffer_records, hipStream_t s) {
  hipLaunchKernelGGL(tile_capacity_kernel, dim3(1), dim3(BIN_BLOCK), 0, s, t, scale, margin,
                     buffer_records);
  return hipGetLastError();
}

// Accumulating workgroups a CU: as many as the tile's LDS lets be resident
// (two for a 64 KB tile, one for 128 KB).
static int accum_per_cu(const TileArgs& t) {
  return (t.tile_wl + t.tile_hl) <= 12 ? 2 : 1;
}

hipError_t launch_tile_accum_batch(const BatchArgs& b, int grid, hipStream_t s) {
  static const hipError_t e0 = allow_lds(reinterpret_cast<const void*>(&tile_accum_batch_kernel));
  if (e0 != hipSuccess) return e0;
  hipLaunchKernelGGL(tile_plan_kernel, dim3(1), dim3(BIN_BLOCK), 0, s, b.g);
  hipLaunchKernelGGL(tile_accum_batch_kernel, dim3(grid * accum_per_cu(b.g)), dim3(BIN_BLOCK),
                     accum_lds_bytes(b.max_pal, false, 1 << (b.g.tile_wl + b.g.tile_hl)), s, b);
  return hipGetLastError();
}

hipError_t launch_tile_capacity(const TileArgs& t, float scale, uint32_t margin,
                                uint64_t buffer_records, hipStream_t s) {
  hipLaunchKernelGGL(tile_capacity_kernel, dim3(1), dim3(BIN_BLOCK), 0, s, t, scale, margin,
                     buffer_records);
  return hipGetLastError();
}

hipError_t launch_tile_accum(const TileArgs& t, const uint32_t* pal, int pal_size, int grid,
                             hipStream_t s) {
  static const hipError_t e0 = allow_lds(reinterpret_cast<const void*>(&tile_accum_kernel<false>));
  static const hipError_t e1 = allow_lds(reinterpret_cast<const void*>(&tile_accum_kernel<true>));
  if (e0 != hipSuccess) return e0;
  if (e1 != hipSuccess) return e1;
  hipLaunchKernelGGL(tile_plan_kernel, dim3(1), dim3(BIN_BLOCK), 0, s, t);
  grid *= accum_per_cu(t);
  if (t.f32)
    hipLaunchKernelGGL(tile_accum_kernel<true>, dim3(grid), dim3(BIN_BLOCK),
                       accum_lds_bytes(pal_size, true, 1 << (t.tile_wl + t.tile_hl)), s, t, pal,
                       pal_size);
  else
    hipLaunchKernelGGL(tile_accum_kernel<false>, dim3(grid), dim3(BIN_BLOCK),
                       accum_lds_bytes(pal_size, false, 1 << (t.tile_wl + t.tile_hl)), s, t, pal,
                       pal_size);
  return hipGetLastError();
}

}  // namespace ffm
