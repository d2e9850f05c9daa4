// runtime.cpp -- host side of libffm.so: the C ABI of include/ffm.h.
//
// Stands in for the reference's FlameRenderer (src/renderer/compute_kernel.rs)
// and render loop (src/renderer/render.rs:371-800): it packs the flame and
// palette into device tables, owns the per-context stream and buffers, runs
// the dispatch / accumulate / tonemap cadence, and reports errors as status
// codes plus a message, as RenderError and pyfae's fae_call do.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "ffm_internal.h"
#include "ffm_soft.h"

using namespace ffm;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

}  // namespace

// The estimator's filters for one (max radius, min radius, curve): the
// squared half-width and normaliser of each (de_table), kept by the
// context so repeated renders upload them once.
struct DeTable {
  float max_r = -1.0f, min_r = -1.0f, curve = -1.0f;
  int R = 0;
  std::vector<float> h2, inv;
};

struct ffm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::atomic<int> cancel{0};
  int num_cu = 0;
  // flame
  bool has_flame = false;
  ffm_flame_desc flame;
  int n_rows = 0;
  uint32_t refresh_period = 0;  // xaos flames whose walk splits orbits: 4
  bool ext = false;             // the flame needs the EXT kernels
  DevBuf table;
  // palette
  bool has_palette = false;
  int pal_size = 0;
  DevBuf pal_u32, pal_f32;
  // render buffers
  DevBuf hist, copies, scratch, accum, accum2, orbits, counters, rgba, de_h2, de_inv, count;
  DeTable de;  // the estimator filters de_h2 and de_inv hold on the device
  DevBuf rgba_out;               // a supersampled render's boxed-down picture
  DevBuf fx_a, fx_b, fx_rgba;    // effect chain ping-pong buffers
  DevBuf fx_pyr;                 // bloom's pyramids, the distance field's coordinates
  ffm_effects fx{};              // effects renders apply (none by default)
  // LDS tile splat: per-tile fill/region/capacity, work list, record buffer
  DevBuf t_fill, t_region, t_cap, t_items, t_records, t_sorted;
  int t_W = 0, t_H = 0, t_nrows = 0, t_pal = 0;
  uint64_t t_walkers = 0;
  uint32_t t_ipt = 0;
  double t_capq = 1.0;         // capacity_scale() the plan was made with
  bool t_planned = false;
  ViewConst t_view{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // event pairs around individual kernel launches of one dispatch
  std::vector<hipEvent_t> kev;
  std::vector<int> kev_kind;  // 0 walker kernel, 1 tile accumulate
  size_t kev_used = 0;
  std::vector<hipEvent_t> dev;  // event pairs around whole dispatches
  size_t dev_used = 0;
  struct ffm_batch_state* batch = nullptr;
  // differentiable soft splat: the fits set last
  std::vector<SoftFit> s_fits;
  DevBuf s_fits_d, s_tables, s_ckpt, s_part, s_xcd;
  uint64_t s_theta = 0, s_pixels = 0, s_ckpt_bytes = 0, s_part_n = 0;
  uint32_t s_blocks = 0;
  size_t s_fwd_lds = 0, s_adj_lds32 = 0, s_adj_lds64 = 0;
  bool s_any32 = false, s_any64 = false;
};

namespace {

int fail(ffm_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHECK(c, expr)                                                               \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail((c), e_ == hipErrorOutOfMemory ? FFM_ERR_OUT_OF_MEMORY : FFM_ERR_DEVICE, \
                  "%s: %s", #expr, hipGetErrorString(e_));                              \
  } while (0)

int ensure(ffm_ctx* c, DevBuf& b, size_t bytes, const char* what) {
  if (b.bytes >= bytes && b.p) return FFM_OK;
  b.release();
  if (bytes == 0) return FFM_OK;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return fail(c, FFM_ERR_OUT_OF_MEMORY,
                "the GPU ran out of memory (%s, %zu bytes); try a smaller size", what, bytes);
  }
  b.bytes = bytes;
  return FFM_OK;
}

hipStream_t pick(ffm_ctx* c, void* stream) {
  return stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
}

bool supported_var(int id) {
  switch (id) {
    case FFM_VAR_LINEAR: case FFM_VAR_SINUSOIDAL: case FFM_VAR_SPHERICAL: case FFM_VAR_SWIRL:
    case FFM_VAR_HORSESHOE: case FFM_VAR_POLAR: case FFM_VAR_DISC: case FFM_VAR_HYPERBOLIC:
    case FFM_VAR_JULIA: case FFM_VAR_WAVES: case FFM_VAR_JULIAN: case FFM_VAR_BUBBLE:
    case FFM_VAR_CURL: case FFM_VAR_BLOB: case FFM_VAR_NOISE: case FFM_VAR_BLUR:
    case FFM_VAR_GAUSSIAN_BLUR: case FFM_VAR_PRE_BLUR:
      return true;
    default:
      return false;
  }
}

// Whether a flame needs the EXT kernels: xaos, no variation slots, weights
// the selection buckets cannot serve, final chains that differ by transform,
// or a variation beyond the thirteen of the workloads (the plain kernels
// leave those paths out).
bool flame_ext(const ffm_flame_desc* f) {
  if (f->has_xaos || f->n_vars < 1) return true;
  // the plain kernels pick through the selection buckets, usable while the
  // cumulative weights (the sequential f32 sums the table holds) are finite
  // and never decrease
  float cum = 0.0f;
  for (int i = 0; i < f->n_xforms; i++) {
    const float next = cum + f->xforms[i].weight;
    if (!std::isfinite(next) || (i > 0 && next < cum)) return true;
    cum = next;
  }
  if (f->n_finals > 0 && f->per_transform_finals)  // chains that differ by transform
    for (int i = 1; i < f->n_xforms; i++) {
      const ffm_xform &x0 = f->xforms[0], &x = f->xforms[i];
      if (x.n_final_attach != x0.n_final_attach ||
          !std::equal(x.final_attach, x.final_attach + x.n_final_attach, x0.final_attach))
        return true;
    }
  for (int k = 0; k < f->n_vars; k++)
    switch (f->var_id[k]) {
      case FFM_VAR_BLOB: case FFM_VAR_NOISE: case FFM_VAR_BLUR: case FFM_VAR_GAUSSIAN_BLUR:
      case FFM_VAR_PRE_BLUR:
        return true;
      default:
        break;
    }
  return false;
}

// Validation and derived values shared by every flame entry point.
int check_flame(ffm_ctx* c, const ffm_flame_desc* f) {
  if (!f) return fail(c, FFM_ERR_INVALID_ARGUMENT, "flame is NULL");
  if (f->n_xforms < 1) return fail(c, FFM_ERR_INVALID_ARGUMENT, "a flame needs at least one transform");
  if (f->n_xforms > FFM_MAX_XFORMS)
    return fail(c, FFM_ERR_TOO_LARGE, "%d transforms: at most %d fit", f->n_xforms, FFM_MAX_XFORMS);
  if (f->n_finals < 0 || f->n_finals > FFM_MAX_FINALS)
    return fail(c, FFM_ERR_TOO_LARGE, "%d final transforms: at most %d fit", f->n_finals, FFM_MAX_FINALS);
  if (f->n_vars < 0 || f->n_vars > FFM_MAX_VARS)
    return fail(c, FFM_ERR_TOO_LARGE, "%d active variations: at most %d fit", f->n_vars, FFM_MAX_VARS);
  for (int k = 0; k < f->n_vars; k++)
    if (!supported_var(f->var_id[k]))
      return fail(c, FFM_ERR_INVALID_ARGUMENT, "variation id %d is not supported", f->var_id[k]);
  if (f->per_transform_finals)
    for (int i = 0; i < f->n_xforms; i++) {
      const ffm_xform& x = f->xforms[i];
      if (x.n_final_attach < 0 || x.n_final_attach > FFM_MAX_FINALS)
        return fail(c, FFM_ERR_INVALID_ARGUMENT, "transform %d attaches %d finals (0..%d)", i,
                    x.n_final_attach, FFM_MAX_FINALS);
      for (int j = 0; j < x.n_final_attach; j++)
        if (x.final_attach[j] < 0 || x.final_attach[j] >= f->n_finals)
          return fail(c, FFM_ERR_INVALID_ARGUMENT,
                      "transform %d attaches final %d; the flame has %d", i, x.final_attach[j],
                      f->n_finals);
    }
  if (f->has_xaos && f->n_xforms > FFM_MAX_XAOS)
    return fail(c, FFM_ERR_TOO_LARGE, "xaos on %d transforms: at most %d", f->n_xforms,
                FFM_MAX_XAOS);
  return FFM_OK;
}

// Flame::xaos_walk_splits_orbits (src/scene/xaos.rs:64-124): with xaos, the
// transforms a first pick can reach, the closed groups among them (every
// transform a member reaches reaches it back), and whether one first pick
// can end in more than one group. Such a flame restarts one walker in
// ORBIT_REFRESH_PERIOD = 4 each dispatch (src/renderer/compute_kernel.rs:
// 46-60); otherwise 0 (no refresh).
uint32_t xaos_refresh_period(const ffm_flame_desc* f) {
  const int n = f->n_xforms;
  if (!f->has_xaos || n == 0) return 0;
  std::vector<float> w(n);
  for (int i = 0; i < n; i++) w[i] = std::max(f->xforms[i].weight, 0.0f);
  std::vector<std::vector<int>> next(n);
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++)
      if (w[j] * f->xaos[i][j] > 0.0f) next[i].push_back(j);
    if (next[i].empty()) next[i].push_back(0);
  }
  std::vector<std::vector<char>> reach(n, std::vector<char>(n, 0));
  for (int s = 0; s < n; s++) {
    std::vector<int> stack{s};
    reach[s][s] = 1;
    while (!stack.empty()) {
      int i = stack.back();
      stack.pop_back();
      for (int j : next[i])
        if (!reach[s][j]) {
          reach[s][j] = 1;
          stack.push_back(j);
        }
    }
  }
  std::vector<char> live(n, 0);
  for (int i = 0; i < n; i++)
    if (w[i] > 0.0f)
      for (int j = 0; j < n; j++) live[j] |= reach[i][j];
  std::vector<int> closed;
  for (int i = 0; i < n; i++) {
    if (!live[i]) continue;
    bool all_back = true, lowest = true;
    for (int j = 0; j < n; j++) {
      if (reach[i][j] && !reach[j][i]) all_back = false;
      if (j < i && reach[i][j] && reach[j][i]) lowest = false;
    }
    if (all_back && lowest) closed.push_back(i);
  }
  if (closed.size() < 2) return 0;
  for (int i = 0; i < n; i++) {
    if (!(w[i] > 0.0f)) continue;
    int groups = 0;
    for (int g : closed) groups += reach[i][g] ? 1 : 0;
    if (groups > 1) return 4;
  }
  return 0;
}

// init_julian (src/variations/defs/advanced.rs:467-474), in place.
void prepare_flame(ffm_flame_desc* f) {
  for (int k = 0; k < f->n_vars; k++) {
    if (f->var_id[k] != FFM_VAR_JULIAN) continue;
    for (int i = 0; i < f->n_xforms + f->n_finals; i++) {
      ffm_xform* xf = i < f->n_xforms ? &f->xforms[i] : &f->finals[i - f->n_xforms];
      float power = xf->var_param[k][0], dist = xf->var_param[k][1];
      xf->var_param[k][2] = dist / std::max(std::fabs(power), 1e-30f) / 2.0f;
    }
  }
}

// Rows of a flame's packed table: the transforms' and finals' rows, then the
// xaos selection table or the selection buckets (the layout depends on the
// counts only, so a batch can lay tables out before their contents).
int packed_rows(const ffm_flame_desc* f) {
  return f->n_xforms + f->n_finals + (f->has_xaos ? xaos_rows(f->n_xforms) : SEL_ROWS);
}

// The pick of select_transform_const (shaders/core/utilities.wgsl:35-55) for
// the 24-bit draw d: the first i with u * total <= cum_i, else the last.
// With non-decreasing sums the pick never decreases with d, so a scan over
// increasing draws starts from the previous pick (`from`).
static int pick_for_draw(const float* cum, int n, uint32_t d, int from) {
  const float u = (float)d * (1.0f / 16777216.0f);
  const float target = u * cum[n - 1];
  for (int i = from; i < n - 1; i++)
    if (target <= cum[i]) return i;
  return n - 1;
}

// Pack the flame into the device table layout of ffm_internal.h.
std::vector<float> pack_flame(const ffm_flame_desc* f) {
  const int rows = f->n_xforms + f->n_finals;
  const int table_rows = packed_rows(f);
  std::vector<float> t(flame_table_floats(table_rows), 0.0f);
  FlameHeader h{};
  h.table_rows = table_rows;
  h.xaos = f->has_xaos ? 1 : 0;
  h.xaos_off = rows * XF_STRIDE;
  h.n_xforms = f->n_xforms;
  h.n_finals = f->n_finals;
  h.n_vars = f->n_vars;
  // Packed slot k holds the descriptor's slot src[k]: pre-phase variations
  // first, then the rest, each group in local-index order, so one pass over
  // the slots runs the pre phase before the weighted sum.
  int src[FFM_MAX_VARS], ns = 0;
  for (int k = 0; k < f->n_vars; k++)
    if (f->var_id[k] == FFM_VAR_PRE_BLUR) src[ns++] = k;
  for (int k = 0; k < f->n_vars; k++)
    if (f->var_id[k] != FFM_VAR_PRE_BLUR) src[ns++] = k;
  for (int k = 0; k < f->n_vars; k++) h.var_id[k] = f->var_id[src[k]];
  // polar and disc both start from atan2(x, y) / pi: share it in a pass
  // whose transforms use both
  auto uses = [&](const ffm_xform& x, int id) {
    for (int k = 0; k < f->n_vars; k++)
      if (f->var_id[k] == id && x.var_weight[k] != 0.0f) return true;
    return false;
  };
  bool polar = false, disc = false;
  for (int i = 0; i < f->n_xforms; i++) {
    polar = polar || uses(f->xforms[i], FFM_VAR_POLAR);
    disc = disc || uses(f->xforms[i], FFM_VAR_DISC);
  }
  h.share_tp = polar && disc ? 1 : 0;
  for (int i = 0; i < f->n_finals; i++)
    if (uses(f->finals[i], FFM_VAR_POLAR) && uses(f->finals[i], FFM_VAR_DISC))
      h.share_tp |= 2 << i;
  for (int i = 0; i < f->n_finals; i++)
    for (int k = 0; k < f->n_vars; k++)
      if (f->finals[i].var_weight[src[k]] != 0.0f)
        h.fin_act[i][h.fin_nact[i]++] = k | h.var_id[k] << 8;
  // each normal transform's final chain (check_flame validated the lists)
  auto chain = [&](int i, int* out) {
    if (!f->per_transform_finals) {
      for (int j = 0; j < f->n_finals; j++) out[j] = j;
      return f->n_finals;
    }
    const ffm_xform& x = f->xforms[i];
    for (int j = 0; j < x.n_final_attach; j++) out[j] = x.final_attach[j];
    return (int)x.n_final_attach;
  };
  h.fin_uniform = 1;
  h.fin_chain_n = chain(0, h.fin_chain);
  for (int i = 1; i < f->n_xforms; i++) {
    int c[FFM_MAX_FINALS];
    const int n = chain(i, c);
    if (n != h.fin_chain_n || !std::equal(c, c + n, h.fin_chain)) h.fin_uniform = 0;
  }
  std::memcpy(t.data(), &h, sizeof h);
  float* r0 = t.data() + sizeof(FlameHeader) / 4;
  float cum = 0.0f;
  for (int i = 0; i < rows; i++) {
    const ffm_xform& x = i < f->n_xforms ? f->xforms[i] : f->finals[i - f->n_xforms];
    float* r = r0 + (size_t)i * XF_STRIDE;
    r[XF_A] = x.a; r[XF_B] = x.b; r[XF_C] = x.c; r[XF_D] = x.d; r[XF_E] = x.e; r[XF_F] = x.f;
    r[XF_PA] = x.post_a; r[XF_PB] = x.post_b; r[XF_PC] = x.post_c;
    r[XF_PD] = x.post_d; r[XF_PE] = x.post_e; r[XF_PF] = x.post_f;
    if (i < f->n_xforms) {
      cum += x.weight;  // the reference's sequential f32 sum
      r[XF_CUM] = cum;
    }
    // (1 + s) / 2 and colour * (1 - s) / 2, as main_template.wgsl:496-503
    r[XF_C1] = (1.0f + x.color_speed) / 2.0f;
    r[XF_C2] = x.color * (1.0f - x.color_speed) / 2.0f;
    r[XF_OPACITY] = x.opacity;
    r[XF_POST] = x.post_enabled ? 1.0f : 0.0f;
    if (i < f->n_xforms) {
      int c[FFM_MAX_FINALS];
      const int n = chain(i, c);
      r[XF_NFIN] = (float)n;
      for (int j = 0; j < n; j++) r[XF_FIN0 + j] = (float)c[j];
    }
    for (int k = 0; k < f->n_vars; k++) {
      r[XF_W0 + k] = x.var_weight[src[k]];
      for (int q = 0; q < FFM_MAX_VAR_PARAMS; q++)
        r[XF_P0 + k * FFM_MAX_VAR_PARAMS + q] = x.var_param[src[k]][q];
    }
  }
  if (!f->has_xaos) {
    // the buckets, when the cumulative sums are finite and never decrease
    std::vector<float> cum(f->n_xforms);
    bool usable = true;
    for (int i = 0; i < f->n_xforms; i++) {
      cum[i] = r0[(size_t)i * XF_STRIDE + XF_CUM];
      if (!std::isfinite(cum[i]) || (i > 0 && cum[i] < cum[i - 1])) usable = false;
    }
    if (usable) {
      const int off = rows * XF_STRIDE;
      uint8_t* sel = reinterpret_cast<uint8_t*>(r0 + off);
      const uint32_t span = 1u << (24 - SEL_BITS);
      int hi = 0;
      for (uint32_t b = 0; b < (1u << SEL_BITS); b++) {
        const int lo = pick_for_draw(cum.data(), f->n_xforms, b * span, hi);
        hi = pick_for_draw(cum.data(), f->n_xforms, b * span + span - 1, lo);
        sel[b] = (uint8_t)(lo | (lo == hi ? 0 : 0x80));
      }
      reinterpret_cast<FlameHeader*>(t.data())->sel_off = off;
    }
  }
  if (f->has_xaos) {
    // select_transform_xaos's running sums of weight[j] * xaos[i][j], one
    // row per previous transform i (f32, in order, no contraction)
    const int n = f->n_xforms;
    float* xc = r0 + h.xaos_off;
    for (int i = 0; i < n; i++) {
      float acc = 0.0f;
      for (int j = 0; j < n; j++) {
        float prod = f->xforms[j].weight * f->xaos[i][j];
        acc += prod;
        xc[i * n + j] = acc;
      }
    }
  }
  return t;
}

// Palette table (src/scene/palette.rs:265-360 sample/generate; f16 storage
// src/gpu/buffers.rs:3215-3221; srgb_to_linear utilities.wgsl:143-146).
void palette_sample(const float* s, int n, float t, float out[3]) {
  if (n <= 0) {
    out[0] = out[1] = out[2] = 1.0f;
    return;
  }
  t = std::min(std::max(t, 0.0f), 1.0f);
  if (t < s[0]) {
    for (int c = 0; c < 3; c++) out[c] = s[1 + c];
    return;
  }
  int idx = 0;
  for (int i = 0; i < n; i++) {
    if (s[4 * i] > t) break;
    idx = i;
  }
  if (idx == n - 1) {
    for (int c = 0; c < 3; c++) out[c] = s[4 * idx + 1 + c];
    return;
  }
  const float* a = s + 4 * idx;
  const float* b = a + 4;
  float range = b[0] - a[0];
  if (range < 1e-6f) {
    for (int c = 0; c < 3; c++) out[c] = a[1 + c];
    return;
  }
  float lt = (t - a[0]) / range;
  for (int c = 0; c < 3; c++) out[c] = a[1 + c] + (b[1 + c] - a[1 + c]) * lt;
}

// Rust's `f32 as usize`: truncation toward zero, saturating (NaN -> 0).
size_t sat_index(float v, int size) {
  if (!(v > 0.0f)) return 0;
  if (v >= (float)size) return (size_t)size - 1;
  return std::min((size_t)v, (size_t)size - 1);
}

// geometric_squeeze_lookup (src/scene/palette.rs:607-640): octave n of the
// input covers r^n (1 - r) and maps onto the whole palette.
float geometric_squeeze(float t, float r) {
  if (r >= 0.999f) return t;
  if (r <= 0.001f) return std::min(std::max(t / std::max(r, 1e-6f), 0.0f), 1.0f);
  if (t <= 0.0f) return 0.0f;
  float start = 0.0f, rn = 1.0f;
  for (int k = 0; k < 64; k++) {
    float width = rn * (1.0f - r);
    float end = start + width;
    if (t < end) {
      if (width <= 0.0f) return 0.0f;
      return std::min(std::max((t - start) / width, 0.0f), 1.0f);
    }
    start = end;
    rn *= r;
    if (rn < 1e-10f) return 1.0f;
  }
  return 1.0f;
}

// log_redistribute_lookup (src/scene/palette.rs:653-671)
float log_redistribute(float t, float strength) {
  if (std::fabs(strength) < 1.1920929e-7f) return t;
  if (strength > 0.0f) {
    float den = std::expm1(strength);
    if (den == 0.0f) return t;
    return std::min(std::max(std::expm1(strength * t) / den, 0.0f), 1.0f);
  }
  float s = -strength;
  float den = std::expm1(s);
  if (den == 0.0f) return t;
  return std::min(std::max(1.0f - std::expm1(s * (1.0f - t)) / den, 0.0f), 1.0f);
}

// transform_lookup_table (src/scene/palette.rs:703-800) as an index map:
// entry i of the transformed table is entry src[i] of the sampled one.
std::vector<int> palette_index_map(int size, const ffm_palette_xf* xf) {
  std::vector<int> idx(size);
  for (int i = 0; i < size; i++) idx[i] = i;
  if (!xf) return idx;
  auto remap = [&](auto src_t_of) {
    std::vector<int> out(size);
    for (int i = 0; i < size; i++) {
      float t = (float)i / (float)size;
      out[i] = idx[sat_index(src_t_of(t) * (float)size, size)];
    }
    idx.swap(out);
  };
  if (xf->squeeze_mode == 0 && xf->squeeze_factor != 1.0f) {
    float f = xf->squeeze_factor;
    remap([f](float t) {
      float x = t * f;
      return x - std::trunc(x);  // f32::fract
    });
  } else if (xf->squeeze_mode == 1) {
    float r = xf->squeeze_factor;
    remap([r](float t) { return geometric_squeeze(t, r); });
  }
  if (!(std::fabs(xf->log_strength) < 1.1920929e-7f)) {
    float sgt = xf->log_strength;
    remap([sgt](float t) { return log_redistribute(t, sgt); });
  }
  if (xf->rotation != 0.0f) {
    int amount = (int)std::round(xf->rotation * (float)size);  // half away from zero
    std::vector<int> out(size);
    for (int i = 0; i < size; i++) {
      int k = (int)(((int64_t)i + amount) % size);
      if (k < 0) k += size;
      out[i] = idx[k];
    }
    idx.swap(out);
  }
  if (xf->reverse) std::reverse(idx.begin(), idx.end());
  return idx;
}

void palette_tables(const float* stops, int n, int size, std::vector<uint32_t>& pk,
                    std::vector<float>& lin, const ffm_palette_xf* xf = nullptr) {
  pk.assign(size, 0);
  lin.assign(3 * (size_t)size, 0.0f);
  std::vector<int> src = palette_index_map(size, xf);
  for (int i = 0; i < size; i++) {
    float t = (float)src[i] / (float)(size - 1);
    float c[3];
    palette_sample(stops, n, t, c);
    uint32_t packed = 0;
    for (int k = 0; k < 3; k++) {
      float v = std::min(std::max(c[k], 0.0f), 1.0f);
      float h = (float)(_Float16)v;  // round-to-nearest-even binary16
      float l = std::pow(std::max(h, 0.0f), 2.2f);
      lin[3 * i + k] = std::min(std::max(l, 0.0f), 1.0f) * 100.0f;
      packed |= ((uint32_t)(std::min(std::max(l, 0.0f), 1.0f) * 100.0f)) << (8 * k);
    }
    pk[i] = packed;
  }
}

ViewConst make_view(const ffm_view& v) {
  ViewConst c;
  c.pan_x = v.pan_x;
  c.pan_y = v.pan_y;
  c.cos_r = std::cos(v.rotation);
  c.sin_r = std::sin(v.rotation);
  c.zoom = v.zoom;
  c.scale = (float)std::min(v.width, v.height) * 0.25f;
  c.cx = (float)v.width * 0.5f;
  c.cy = (float)v.height * 0.5f;
  c.width = v.width;
  c.height = v.height;
  return c;
}

int check_view(ffm_ctx* c, const ffm_view* v) {
  if (!v) return fail(c, FFM_ERR_INVALID_ARGUMENT, "view is NULL");
  if (v->width < 1 || v->height < 1)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "image size %dx%d is empty", v->width, v->height);
  if (v->width > 32768 || v->height > 32768)
    return fail(c, FFM_ERR_TOO_LARGE, "image side over 32768 (%dx%d)", v->width, v->height);
  return FFM_OK;
}

IterArgs base_args(ffm_ctx* c, const ffm_view& v) {
  IterArgs a{};
  a.flame = reinterpret_cast<const float*>(c->table.p);
  a.n_rows = c->n_rows;
  a.ext = c->ext ? 1 : 0;
  a.pal_u32 = reinterpret_cast<const uint32_t*>(c->pal_u32.p);
  a.pal_f32 = reinterpret_cast<const float*>(c->pal_f32.p);
  a.pal_size = c->pal_size;
  a.view = make_view(v);
  return a;
}

int iterate_grid(ffm_ctx* c, uint64_t walkers) {
  uint64_t want = (walkers + 255) / 256;
  uint64_t cap = (uint64_t)std::max(c->num_cu, 1) * 8;  // 8 x 256 lanes resident per CU
  return (int)std::max<uint64_t>(1, std::min(want, cap));
}

constexpr int XCD_COPIES = 8;
// Staged samples held at once: up to 2^33 records (32 GiB; a flush keeps a
// run's start in 32 bits plus one bit beside its room). single4k's 2^24
// walkers then walk in one launch and one tile accumulation.
constexpr uint64_t RECORD_BUDGET = (1ull << 33) - (1ull << 24);
// The count pre-passes walk a quarter of the plotted iterations (rounded up).
constexpr uint32_t PRE_ITER_DIV = 4;

// FFM_TILE_CAPACITY_SCALE (a positive number, default 1) multiplies every
// estimated region capacity and its margin. Below 1 it sends the records
// past the smaller regions to the direct-atomic fallback, which is how the
// tests reach that path; the histogram is the same either way.
double capacity_scale() {
  const char* e = std::getenv("FFM_TILE_CAPACITY_SCALE");
  if (!e) return 1.0;
  const double v = std::strtod(e, nullptr);
  return std::isfinite(v) && v > 0.0 ? v : 1.0;
}

uint32_t capacity_margin(double q) { return (uint32_t)std::max(1.0, std::min(1024.0 * q, 1.0e9)); }
constexpr size_t LDS_PER_CU = LDS_DYNAMIC_MAX;

// The LDS tile splat of one dispatch (see tilesplat.hip): walkers in chunks
// whose records fit the buffer; region capacities from a count-only
// pre-pass over a subset of the first chunk (kept while the flame, size and
// dispatch shape stay the same); per chunk: clear fills, binned iterate,
// tile accumulate into `hist`.
// Bracket a kernel launch with a pair of timing events (kind 0: walker
// kernel, 1: tile accumulate), or (kind 2) a dispatch's whole splat phase,
// whose pairs hold the kernels' and so live in a pool of their own; read
// back by collect_timings.
hipError_t mark(ffm_ctx* c, hipStream_t s, int kind, bool begin) {
  if (kind == 2) {
    if (c->dev_used == c->dev.size()) {
      hipEvent_t e;
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      c->dev.push_back(e);
    }
    return hipEventRecord(c->dev[c->dev_used++], s);
  }
  if (begin) {
    while (c->kev.size() < c->kev_used + 2) {
      hipEvent_t e;
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      c->kev.push_back(e);
      c->kev_kind.push_back(0);
    }
    c->kev_kind[c->kev_used] = kind;
    return hipEventRecord(c->kev[c->kev_used++], s);
  }
  return hipEventRecord(c->kev[c->kev_used++], s);
}

// Waits for the last recorded event, sums the pairs by kind (2: a whole
// dispatch's splat phase) and frees the pool for reuse.
hipError_t collect_timings(ffm_ctx* c, double& walker_ms, uint64_t& launches, double& accum_ms,
                           double& dispatch_ms) {
  for (hipEvent_t* last : {c->kev_used ? &c->kev[c->kev_used - 1] : nullptr,
                           c->dev_used ? &c->dev[c->dev_used - 1] : nullptr}) {
    if (!last) continue;
    const hipError_t e = hipEventSynchronize(*last);
    if (e != hipSuccess) return e;
  }
  for (size_t i = 0; i + 1 < c->kev_used; i += 2) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->kev[i], c->kev[i + 1]) != hipSuccess) continue;
    if (c->kev_kind[i] == 0) {
      walker_ms += ms;
      launches++;
    } else {
      accum_ms += ms;
    }
  }
  for (size_t i = 0; i + 1 < c->dev_used; i += 2) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->dev[i], c->dev[i + 1]) == hipSuccess) dispatch_ms += ms;
  }
  c->kev_used = 0;
  c->dev_used = 0;
  return hipSuccess;
}

// Tile tables and record buffers the caller provides (ffm_iterate_noalloc):
// nothing is allocated, the plan is made afresh on every call (the memory
// is shared by every call on the device) and no timing events are recorded.
struct TileScratch {
  uint32_t *fill, *cap, *items, *records;
  uint64_t* region;
  uint64_t records_cap;  // u32 words at `records`
};

// The scratch of ffm_iterate_noalloc laid out: fill, region (two words a
// bin), cap, items, then the records (and, for bins of several tiles, their
// sorted copy).
constexpr size_t NOALLOC_TABLE_WORDS = 5 * (size_t)TILE_MAX + 64;
static_assert(FFM_NOALLOC_SCRATCH_BYTES > NOALLOC_TABLE_WORDS * 4 + (64u << 20),
              "scratch holds the tables and a useful record buffer");

TileScratch noalloc_layout(void* base) {
  uint32_t* w = reinterpret_cast<uint32_t*>(base);
  TileScratch t;
  t.fill = w;
  t.region = reinterpret_cast<uint64_t*>(w + TILE_MAX);
  t.cap = w + 3 * TILE_MAX;
  t.items = w + 4 * TILE_MAX;
  t.records = w + NOALLOC_TABLE_WORDS;
  t.records_cap = FFM_NOALLOC_SCRATCH_BYTES / 4 - NOALLOC_TABLE_WORDS;
  return t;
}

int lds_splat(ffm_ctx* c, IterArgs a, uint32_t* hist, hipStream_t s, bool f32 = false,
              const TileScratch* ext = nullptr) {
  const int W = a.view.width, H = a.view.height;
  const int tiles_x = (W + TILE_W - 1) / TILE_W, tiles_y = (H + TILE_H - 1) / TILE_H;
  // bins of 2^g tiles along a row, as few tiles per bin as keep bins <= TILE_MAX
  int g = 0;
  while (g < 7 && (int64_t)tiles_y * ((tiles_x + (1 << g) - 1) >> g) > TILE_MAX) g++;
  const int bins_x = (tiles_x + (1 << g) - 1) >> g;
  const int n_tiles = bins_x * tiles_y;  // bins
  if (n_tiles > TILE_MAX)
    return fail(c, FFM_ERR_TOO_LARGE, "the LDS tile splat holds %d bins of up to 128 %dx%d tiles; "
                "%dx%d needs more", TILE_MAX, TILE_W, TILE_H, W, H);
  int walk_block = 0, stage = 0, per_cu = 0;
  const int nb = (n_tiles + 63) / 64 * 64;
  bin_shape(a.n_rows, a.pal_size, nb, &walk_block, &stage, &per_cu, 3);
  if (bin_lds_bytes(a.n_rows, a.pal_size, stage, nb) > LDS_PER_CU ||
      accum_lds_bytes(a.pal_size, f32, TILE_PX) > LDS_PER_CU)
    return fail(c, FFM_ERR_TOO_LARGE, "flame table + palette too large for the LDS tile splat");
  const uint64_t per_walker = (uint64_t)a.ipt + 1;
  // records held at once: RECORD_BUDGET in context buffers; in a caller's scratch
  // what it holds (half of it for bins of several tiles, whose records
  // also need a copy sorted by tile)
  const uint64_t budget =
      ext ? (g > 0 ? ext->records_cap / 2 : ext->records_cap) : RECORD_BUDGET;
  if (per_walker > budget * 3 / 4)
    return fail(c, FFM_ERR_TOO_LARGE,
                "%u iterations per walker do not fit the %llu-record splat buffer", a.ipt,
                (unsigned long long)budget);
  uint64_t chunk = std::max<uint64_t>(1, (budget * 3 / 4) / per_walker);
  // whole rounds of the walker grid: only the last chunk ends on a partial
  // round, whose idle lanes leave their CUs' workgroups waiting
  const uint64_t lanes = (uint64_t)std::max(c->num_cu, 1) * per_cu * walk_block;
  if (chunk > lanes) chunk = chunk / lanes * lanes;
  chunk = std::min<uint64_t>(chunk, a.walkers);
  const uint64_t rec_cap = std::min<uint64_t>(budget, chunk * per_walker * 5 / 4 +
                                                      (uint64_t)n_tiles * 4096);
  int st;
  TileArgs t{};
  if (ext) {
    t.tile_fill = ext->fill;
    t.region = ext->region;
    t.cap = ext->cap;
    t.item_start = ext->items;
    t.records = ext->records;
    t.sorted = g > 0 ? ext->records + rec_cap : nullptr;
  } else {
    if ((st = ensure(c, c->t_fill, (size_t)TILE_MAX * 4, "tile fills"))) return st;
    if ((st = ensure(c, c->t_region, (size_t)TILE_MAX * 8, "tile regions"))) return st;
    if ((st = ensure(c, c->t_cap, (size_t)TILE_MAX * 4, "tile capacities"))) return st;
    if ((st = ensure(c, c->t_items, (size_t)(TILE_MAX + 4) * 4, "tile work list"))) return st;
    if ((st = ensure(c, c->t_records, rec_cap * 4, "tile records"))) return st;
    if (g > 0 && (st = ensure(c, c->t_sorted, rec_cap * 4, "tile records sorted by tile")))
      return st;
    t.sorted = g > 0 ? reinterpret_cast<uint32_t*>(c->t_sorted.p) : nullptr;
    t.tile_fill = reinterpret_cast<uint32_t*>(c->t_fill.p);
    t.region = reinterpret_cast<uint64_t*>(c->t_region.p);
    t.cap = reinterpret_cast<uint32_t*>(c->t_cap.p);
    t.item_start = reinterpret_cast<uint32_t*>(c->t_items.p);
    t.records = reinterpret_cast<uint32_t*>(c->t_records.p);
  }
  t.hist = hist;
  t.pal_f32 = a.pal_f32;
  t.f32 = f32 ? 1 : 0;
  t.tiles_x = tiles_x;
  t.n_tiles = n_tiles;
  t.bins_x = bins_x;
  t.group_shift = g;
  t.width = W;
  t.height = H;
  t.flush_every = 4;
  t.stage = stage;
  t.walk_block = walk_block;
  t.walk_per_cu = per_cu;
  t.bins_lds = nb;
  t.tile_wl = TILE_WL;
  t.tile_hl = TILE_HL;
  const int grid = std::max(c->num_cu, 1);
  const bool same = !ext && c->t_planned && c->t_W == W && c->t_H == H && c->t_nrows == a.n_rows &&
                    c->t_pal == a.pal_size && c->t_walkers == chunk && c->t_ipt == a.ipt &&
                    c->t_capq == capacity_scale() &&
                    std::memcmp(&c->t_view, &a.view, sizeof(ViewConst)) == 0;
  if (!same) {
    // 1/32 of the walkers (at least 65536), each for its burn-in and a
    // quarter of its plotted iterations: the tiles' shares of an
    // attractor's samples do not depend on how long it is walked, and a
    // small flame's pre-pass is as long as its walkers' dependent iterations
    IterArgs pre = a;
    pre.walkers = std::min<uint64_t>(chunk, std::max<uint64_t>(chunk / 32, 65536));
    pre.plotted = pre.inframe = nullptr;
    double scale = (double)chunk / (double)pre.walkers;
    if (a.ipt > a.burn_in) {
      pre.ipt = a.burn_in + (a.ipt - a.burn_in + PRE_ITER_DIV - 1) / PRE_ITER_DIV;
      scale *= (double)(a.ipt - a.burn_in) / (double)(pre.ipt - a.burn_in);
    }
    HIPCHECK(c, launch_clear_u32(t.tile_fill, TILE_MAX, s));
    HIPCHECK(c, launch_bin_iterate(pre, t, true, grid, s));
    const double q = capacity_scale();
    HIPCHECK(c, launch_tile_capacity(t, (float)(scale * q), capacity_margin(q), rec_cap, s));
  }
  if (!same && !ext) {
    c->t_planned = true;
    c->t_W = W;
    c->t_H = H;
    c->t_nrows = a.n_rows;
    c->t_pal = a.pal_size;
    c->t_walkers = chunk;
    c->t_ipt = a.ipt;
    c->t_capq = capacity_scale();
    c->t_view = a.view;
  }
  for (uint64_t w0 = 0; w0 < a.walkers; w0 += chunk) {
    IterArgs ch = a;
    ch.walker0 = a.walker0 + (uint32_t)w0;
    ch.walkers = std::min<uint64_t>(chunk, a.walkers - w0);
    if (a.orbits) ch.orbits = a.orbits + w0;
    HIPCHECK(c, launch_clear_u32(t.tile_fill, TILE_MAX, s));
    if (!ext) HIPCHECK(c, mark(c, s, 0, true));
    HIPCHECK(c, launch_bin_iterate(ch, t, false, grid, s));
    if (!ext) HIPCHECK(c, mark(c, s, 0, false));
    if (!ext) HIPCHECK(c, mark(c, s, 1, true));
    HIPCHECK(c, launch_tile_accum(t, a.pal_u32, a.pal_size, grid, s));
    if (!ext) HIPCHECK(c, mark(c, s, 1, false));
  }
  return FFM_OK;
}

// Host twin of OrbitFuses (src/renderer/compute_kernel.rs:73-102).
struct OrbitFuses {
  std::vector<uint32_t> left;
  uint64_t dispatch(uint32_t threads, uint32_t ipt, uint32_t burn_in, uint32_t period,
                    uint32_t seed) {
    if (left.size() < threads) left.resize(threads, burn_in);
    uint64_t plotted = 0;
    for (uint32_t tid = 0; tid < threads; tid++) {
      if (period != 0 && (tid + seed) % period == 0) left[tid] = burn_in;
      uint32_t burnt = std::min(left[tid], ipt);
      plotted += ipt - burnt;
      left[tid] -= burnt;
    }
    return plotted;
  }
};

// flam3 filter table (see ffm_density_estimate in include/ffm.h).
void de_table(float max_r, float min_r, float curve, DeTable& t) {
  double comp_max = (double)max_r + 1.0, comp_min = (double)min_r + 1.0;
  t.h2.clear();
  t.inv.clear();
  const int keep = 100;
  t.R = (int)std::ceil(comp_max);
  for (int i = 0; i < 4096; i++) {
    double adj = i < keep ? (double)i : std::pow((double)(i - keep), 1.0 / curve) + keep;
    double h = comp_max / std::pow(adj + 1.0, curve);
    bool last = false;
    if (h <= comp_min) {
      h = comp_min;
      last = true;
    }
    int r = (int)std::ceil(h);
    double sum = 0.0;
    for (int dy = -r; dy <= r; dy++)
      for (int dx = -r; dx <= r; dx++) {
        double d = std::sqrt((double)(dx * dx + dy * dy)) / h;
        if (d <= 1.0) sum += std::exp(-2.0 * (1.5 * d) * (1.5 * d));
      }
    t.h2.push_back((float)(h * h));
    t.inv.push_back((float)(1.0 / sum));
    if (last) break;
  }
  t.max_r = max_r;
  t.min_r = min_r;
  t.curve = curve;
}

// A tone curve the tone map can sample: 2..FFM_MAX_CURVE_POINTS finite
// points with non-decreasing x (curve_n 0 stands for the linear curve).
bool curve_ok(const ffm_tonemap_params& tp) {
  if (!tp.use_curve || tp.curve_n == 0) return true;
  if (tp.curve_n < 2 || tp.curve_n > FFM_MAX_CURVE_POINTS) return false;
  for (int i = 0; i < tp.curve_n; i++) {
    if (!std::isfinite(tp.curve_x[i]) || !std::isfinite(tp.curve_y[i])) return false;
    if (i > 0 && tp.curve_x[i] < tp.curve_x[i - 1]) return false;
  }
  return true;
}

// v rounded to the nearest f16 value, ties to even, as a float (the
// f16::from_f32 of ToneCurve::generate_lut, src/scene/tonemap.rs:289-303).
float round_f16(float v) {
  if (std::isnan(v)) return v;
  if (!(std::fabs(v) < 65520.0f)) return std::copysign(INFINITY, v);
  int e;
  std::frexp(v, &e);                 // 2^(e-1) <= |v| < 2^e
  const int q = std::max(e - 11, -24);  // f16's spacing there: 10 fraction bits, subnormals 2^-24
  return std::ldexp(std::nearbyint(std::ldexp(v, -q)), q);
}

// ToneCurve::evaluate (src/scene/tonemap.rs:209-286): the cubic Hermite
// through the points, tangents the mean of the neighbouring slopes (the
// segment's own at the ends), limited by Fritsch-Carlson: zero on a flat
// segment, the secant when a tangent opposes it, scaled onto
// alpha^2 + beta^2 = 9 when beyond. Every step in f32.
float curve_eval(const float* px, const float* py, int n, float x) {
  x = std::min(std::max(x, 0.0f), 1.0f);
  for (int i = 0; i + 1 < n; i++) {
    const float x0 = px[i], y0 = py[i], x1 = px[i + 1], y1 = py[i + 1];
    if (!(x >= x0 && x <= x1)) continue;
    const float delta = (y1 - y0) / (x1 - x0);
    const float m0 = i == 0 ? delta : ((y0 - py[i - 1]) / (x0 - px[i - 1]) + delta) * 0.5f;
    const float m1 = i == n - 2 ? delta : (delta + (py[i + 2] - y1) / (px[i + 2] - x1)) * 0.5f;
    float t0, t1;
    if (std::fabs(delta) < 1e-6f) {
      t0 = t1 = 0.0f;
    } else {
      const float alpha = m0 / delta, beta = m1 / delta;
      if (alpha < 0.0f || beta < 0.0f) {
        t0 = t1 = delta;
      } else {
        const float sum_sq = alpha * alpha + beta * beta;
        if (sum_sq > 9.0f) {
          const float tau = 3.0f / std::sqrt(sum_sq);
          t0 = tau * alpha * delta;
          t1 = tau * beta * delta;
        } else {
          t0 = m0;
          t1 = m1;
        }
      }
    }
    const float t = (x - x0) / (x1 - x0), t2 = t * t, t3 = t2 * t;
    const float h00 = 2.0f * t3 - 3.0f * t2 + 1.0f, h10 = t3 - 2.0f * t2 + t;
    const float h01 = -2.0f * t3 + 3.0f * t2, h11 = t3 - t2;
    const float dx = x1 - x0;
    return h00 * y0 + h10 * dx * t0 + h01 * y1 + h11 * dx * t1;
  }
  return x;
}

// The tone-map modes and curve a render or ffm_tonemap accepts.
static int check_tonemap(ffm_ctx* c, const ffm_tonemap_params& tp) {
  if (tp.tonemap_mode < 0 || tp.tonemap_mode > 2 || tp.highlight_mode < 0 ||
      tp.highlight_mode > 3 || tp.transparent_mode < 0 || tp.transparent_mode > 2)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "tonemap mode not supported");
  if (!curve_ok(tp))
    return fail(c, FFM_ERR_INVALID_ARGUMENT,
                "tone curve needs 2..%d finite points with non-decreasing x",
                FFM_MAX_CURVE_POINTS);
  return FFM_OK;
}

TonemapConst make_tonemap(const ffm_tonemap_params& tp, int W, int H, float zoom,
                          double samples_in_buffer) {
  // update_tonemap (src/renderer/compute_kernel.rs:3813-3918) and
  // brightness_scale's uniform terms (shaders/tonemap.wgsl:94-113).
  TonemapConst t{};
  float apo_zoom = std::log2(zoom);
  float ppu = (float)std::min(W, H) * 0.25f * std::pow(2.0f, apo_zoom);
  float area = (float)((uint32_t)W * (uint32_t)H) / (ppu * ppu);
  float total_pixels = std::max((float)W * (float)H, 1.0f);
  float sd = std::max((float)samples_in_buffer / total_pixels, 1e-6f);
  t.mode = tp.tonemap_mode;
  t.highlight = tp.highlight_mode;
  t.levels_enabled = tp.levels_enabled;
  t.white = tp.white_level;
  t.k1 = 1.0f * 2.3f * tp.brightness * 268.0f * 67108864.0f / 256.0f;
  t.k2 = 1.0f / (1.0f * area * tp.white_level * sd);
  t.gamma = tp.gamma == 0.0f ? tp.gamma : 1.0f / tp.gamma;
  t.gamma_threshold = tp.gamma_threshold;
  t.funcval = tp.gamma_threshold != 0.0f ? std::pow(tp.gamma_threshold, t.gamma - 1.0f) : 0.0f;
  t.vib = std::nearbyint(tp.vibrancy * 256.0f);
  t.notvib = 256.0f - t.vib;
  t.exposure = tp.exposure;
  t.density_scale = tp.density_scale;
  t.saturation = tp.saturation;
  t.hue_shift = tp.hue_shift;
  t.alpha_low = tp.alpha_blend_low;
  t.alpha_high = tp.alpha_blend_high;
  t.levels_low = tp.levels_low;
  t.levels_high = tp.levels_high;
  t.levels_gamma = tp.levels_gamma;
  t.sample_density = sd;
  for (int k = 0; k < 3; k++) t.bg_lin[k] = std::pow(std::max(tp.background[k], 0.0f), 2.2f);
  t.inv22 = (float)(1.0 / 2.2);
  t.transparent = tp.transparent_mode;
  t.use_curve = tp.use_curve != 0;
  if (t.use_curve) {
    static const float lin_x[2] = {0.0f, 1.0f};
    const bool lin = tp.curve_n == 0;
    const float* px = lin ? lin_x : tp.curve_x;
    const float* py = lin ? lin_x : tp.curve_y;
    // a caller's curve_n is checked by curve_ok first; the clamp keeps
    // the reads inside curve_x / curve_y whatever it holds
    const int n = lin ? 2 : std::min(std::max(tp.curve_n, 1), FFM_MAX_CURVE_POINTS);
    for (int i = 0; i < 256; i++) t.curve_lut[i] = round_f16(curve_eval(px, py, n, (float)i / 255.0f));
  }
  return t;
}

}  // namespace

void free_batch(ffm_batch_state* b);

extern "C" {

int ffm_abi_version(void) { return FFM_ABI_VERSION; }

int ffm_ctx_create(int device, ffm_ctx** out) {
  if (!out) return FFM_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return FFM_ERR_DEVICE;
  if (device < 0 || device >= n) return FFM_ERR_INVALID_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return FFM_ERR_DEVICE;
  ffm_ctx* c = new ffm_ctx();
  c->fx.pixel_scale = 1.0f;
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamDefault) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return FFM_ERR_DEVICE;
  }
  c->num_cu = prop.multiProcessorCount;
  // this device's copy of ffm_iterate_noalloc's scratch (looked up once per
  // device; a failed lookup is reported by that entry point)
  (void)noalloc_scratch_init(device);
  *out = c;
  return FFM_OK;
}

int ffm_ctx_destroy(ffm_ctx* c) {
  if (!c) return FFM_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (DevBuf* b : {&c->table, &c->pal_u32, &c->pal_f32, &c->hist, &c->copies, &c->scratch,
                    &c->rgba_out, &c->accum, &c->accum2, &c->orbits, &c->counters, &c->rgba,
                    &c->de_h2, &c->de_inv, &c->count, &c->fx_a, &c->fx_b, &c->fx_rgba,
                    &c->fx_pyr, &c->t_fill, &c->t_region, &c->t_cap, &c->t_items,
                    &c->t_records, &c->t_sorted, &c->s_fits_d, &c->s_tables, &c->s_ckpt,
                    &c->s_part, &c->s_xcd})
    b->release();
  free_batch(c->batch);
  for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->dev) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FFM_OK;
}

const char* ffm_last_error(const ffm_ctx* c) { return c ? c->err.c_str() : "no context"; }

int ffm_cancel(ffm_ctx* c) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  c->cancel.store(1);
  return FFM_OK;
}

int ffm_flame_set(ffm_ctx* c, const ffm_flame_desc* f) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  int st = check_flame(c, f);
  if (st) return st;
  (void)hipSetDevice(c->device);
  c->flame = *f;
  prepare_flame(&c->flame);
  std::vector<float> t = pack_flame(&c->flame);
  if ((st = ensure(c, c->table, t.size() * 4, "flame table"))) return st;
  HIPCHECK(c, hipMemcpyAsync(c->table.p, t.data(), t.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->n_rows = reinterpret_cast<const FlameHeader*>(t.data())->table_rows;
  c->refresh_period = xaos_refresh_period(&c->flame);
  c->ext = flame_ext(&c->flame);
  c->has_flame = true;
  c->t_planned = false;  // the attractor, hence the tile capacities, changed
  return FFM_OK;
}

int ffm_palette_set(ffm_ctx* c, const float* stops, int n_stops, int table_size) {
  return ffm_palette_set_xf(c, stops, n_stops, table_size, nullptr);
}

int ffm_palette_set_xf(ffm_ctx* c, const float* stops, int n_stops, int table_size,
                       const ffm_palette_xf* xf) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  if (xf && xf->squeeze_mode != 0 && xf->squeeze_mode != 1)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "palette squeeze mode %d", xf->squeeze_mode);
  if (!stops || n_stops < 1) return fail(c, FFM_ERR_NO_PALETTE, "No palette found");
  if (table_size < FFM_MIN_PALETTE || table_size > FFM_MAX_PALETTE)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "palette size %d outside %d..%d", table_size,
                FFM_MIN_PALETTE, FFM_MAX_PALETTE);
  (void)hipSetDevice(c->device);
  std::vector<uint32_t> pk;
  std::vector<float> lin;
  palette_tables(stops, n_stops, table_size, pk, lin, xf);
  int st;
  if ((st = ensure(c, c->pal_u32, pk.size() * 4, "palette"))) return st;
  if ((st = ensure(c, c->pal_f32, lin.size() * 4, "palette"))) return st;
  HIPCHECK(c, hipMemcpyAsync(c->pal_u32.p, pk.data(), pk.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->pal_f32.p, lin.data(), lin.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->pal_size = table_size;
  c->has_palette = true;
  return FFM_OK;
}

int ffm_iterate(ffm_ctx* c, const ffm_view* view, uint64_t walkers, uint32_t ipt, uint32_t burn_in,
                uint32_t seed, void* orbits, int fresh, void* hist, int splat_mode,
                uint64_t* plotted, uint64_t* inframe, void* stream) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  if (!c->has_flame) return fail(c, FFM_ERR_NO_FLAME, "no flame set");
  if (!c->has_palette) return fail(c, FFM_ERR_NO_PALETTE, "No palette found");
  int st = check_view(c, view);
  if (st) return st;
  if (!hist) return fail(c, FFM_ERR_INVALID_ARGUMENT, "histogram is NULL");
  if (walkers == 0 || walkers > 0xFFFFFFFFull)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "walkers %llu outside 1..2^32-1",
                (unsigned long long)walkers);
  if (splat_mode < FFM_SPLAT_U32_GLOBAL || splat_mode > FFM_SPLAT_F32_LDS)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "splat mode %d not available here", splat_mode);
  (void)hipSetDevice(c->device);
  hipStream_t s = pick(c, stream);
  c->kev_used = 0;
  IterArgs a = base_args(c, *view);
  a.walkers = walkers;
  a.ipt = ipt;
  a.burn_in = burn_in;
  a.seed = seed;
  a.orbits = reinterpret_cast<float4*>(orbits);
  a.fresh = fresh;
  a.refresh_period = orbits ? c->refresh_period : 0u;
  if (orbits && c->flame.has_xaos && burn_in >= (1u << 24))
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "xaos orbits keep burn-in below 2^24 (got %u)",
                burn_in);
  a.plotted = reinterpret_cast<unsigned long long*>(plotted);
  a.inframe = reinterpret_cast<unsigned long long*>(inframe);
  const size_t words = (size_t)view->width * view->height * 4;
  if (splat_mode == FFM_SPLAT_U32_LDS || splat_mode == FFM_SPLAT_F32_LDS)
    return lds_splat(c, a, reinterpret_cast<uint32_t*>(hist), s, splat_mode == FFM_SPLAT_F32_LDS);
  if (splat_mode == FFM_SPLAT_U32_XCD) {
    if ((st = ensure(c, c->copies, words * 4 * XCD_COPIES, "per-XCD histograms"))) return st;
    HIPCHECK(c, launch_clear_u32(c->copies.p, words * XCD_COPIES, s));
    a.hist = c->copies.p;
    a.hist_copy_words = (uint32_t)words;
  } else {
    a.hist = hist;
  }
  HIPCHECK(c, launch_iterate(a, splat_mode, iterate_grid(c, walkers), 256, s));
  if (splat_mode == FFM_SPLAT_U32_XCD)
    HIPCHECK(c, launch_fold_xcd(reinterpret_cast<uint32_t*>(hist),
                                reinterpret_cast<const uint32_t*>(c->copies.p), words, XCD_COPIES, s));
  return FFM_OK;
}

int ffm_iterate_noalloc(ffm_ctx* c, const ffm_view* view, uint64_t walkers, uint32_t ipt,
                        uint32_t burn_in, uint32_t seed, void* orbits, int fresh, void* hist,
                        int splat_mode, uint64_t* plotted, uint64_t* inframe, void* stream) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  if (!c->has_flame) return fail(c, FFM_ERR_NO_FLAME, "no flame set");
  if (!c->has_palette) return fail(c, FFM_ERR_NO_PALETTE, "No palette found");
  int st = check_view(c, view);
  if (st) return st;
  if (!hist) return fail(c, FFM_ERR_INVALID_ARGUMENT, "histogram is NULL");
  if (walkers == 0 || walkers > 0xFFFFFFFFull)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "walkers %llu outside 1..2^32-1",
                (unsigned long long)walkers);
  if (splat_mode != FFM_SPLAT_U32_GLOBAL && splat_mode != FFM_SPLAT_U32_LDS &&
      splat_mode != FFM_SPLAT_F32_GLOBAL && splat_mode != FFM_SPLAT_F32_LDS)
    return fail(c, FFM_ERR_INVALID_ARGUMENT,
                "splat mode %d is not available without allocation (its per-XCD histograms "
                "exceed the %u-byte scratch)", splat_mode, (unsigned)FFM_NOALLOC_SCRATCH_BYTES);
  if (orbits && c->flame.has_xaos && burn_in >= (1u << 24))
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "xaos orbits keep burn-in below 2^24 (got %u)",
                burn_in);
  (void)hipSetDevice(c->device);
  const hipError_t e = noalloc_scratch_init(c->device);
  void* base = noalloc_scratch(c->device);
  if (e != hipSuccess || !base)
    return fail(c, FFM_ERR_DEVICE, "the device scratch of this entry point is not available: %s",
                hipGetErrorString(e));
  hipStream_t s = pick(c, stream);
  IterArgs a = base_args(c, *view);
  a.walkers = walkers;
  a.ipt = ipt;
  a.burn_in = burn_in;
  a.seed = seed;
  a.orbits = reinterpret_cast<float4*>(orbits);
  a.fresh = fresh;
  a.refresh_period = orbits ? c->refresh_period : 0u;
  a.plotted = reinterpret_cast<unsigned long long*>(plotted);
  a.inframe = reinterpret_cast<unsigned long long*>(inframe);
  if (splat_mode == FFM_SPLAT_U32_LDS || splat_mode == FFM_SPLAT_F32_LDS) {
    const TileScratch ts = noalloc_layout(base);
    return lds_splat(c, a, reinterpret_cast<uint32_t*>(hist), s,
                     splat_mode == FFM_SPLAT_F32_LDS, &ts);
  }
  a.hist = hist;
  HIPCHECK(c, launch_iterate(a, splat_mode, iterate_grid(c, walkers), 256, s));
  return FFM_OK;
}

int ffm_downsample2x(ffm_ctx* c, const uint8_t* in, uint8_t* out, int out_w, int out_h,
                     void* stream) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  if (!in || !out) return fail(c, FFM_ERR_INVALID_ARGUMENT, "pixel buffer is NULL");
  if (out_w < 1 || out_h < 1 || out_w > 16384 || out_h > 16384)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "output size %dx%d outside 1..16384", out_w, out_h);
  (void)hipSetDevice(c->device);
  HIPCHECK(c, launch_downsample2x(in, out, out_w, out_h, pick(c, stream)));
  return FFM_OK;
}

int ffm_fold_xcd(ffm_ctx* c, uint32_t* hist, int width, int height, void* stream) {
  if (!c || !hist) return FFM_ERR_INVALID_ARGUMENT;
  size_t words = (size_t)width * height * 4;
  if (c->copies.bytes < words * 4 * XCD_COPIES)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "no per-XCD histograms of that size");
  HIPCHECK(c, launch_fold_xcd(hist, reinterpret_cast<const uint32_t*>(c->copies.p), words,
                              XCD_COPIES, pick(c, stream)));
  return FFM_OK;
}

int ffm_trace(ffm_ctx* c, const ffm_view* view, uint32_t walker, uint32_t seed, uint32_t burn_in,
              uint32_t steps, uint32_t* out, void* stream) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  if (!c->has_flame) return fail(c, FFM_ERR_NO_FLAME, "no flame set");
  if (!c->has_palette) return fail(c, FFM_ERR_NO_PALETTE, "No palette found");
  int st = check_view(c, view);
  if (st) return st;
  if (!out) return fail(c, FFM_ERR_INVALID_ARGUMENT, "trace buffer is NULL");
  (void)hipSetDevice(c->device);
  IterArgs a = base_args(c, *view);
  a.seed = seed;
  a.burn_in = burn_in;
  HIPCHECK(c, launch_trace(a, walker, steps, out, pick(c, stream)));
  return FFM_OK;
}

int ffm_blur(ffm_ctx* c, uint32_t* hist, uint32_t* scratch, int width, int height, float sigma_px,
             float density_sigma, void* stream) {
  if (!c || !hist || !scratch || width < 1 || height < 1) return FFM_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(c->device);
  HIPCHECK(c, launch_blur(hist, scratch, width, height, sigma_px, density_sigma, pick(c, stream)));
  return FFM_OK;
}

int ffm_density_estimate(ffm_ctx* c, const float* in, float* out, int width, int height,
                         float max_radius, float min_radius, float curve, void* stream) {
  if (!c || !in || !out || width < 1 || height < 1) return FFM_ERR_INVALID_ARGUMENT;
  if (!(max_radius > 0.0f) || max_radius > 15.0f || min_radius < 0.0f || !(curve > 0.0f))
    return fail(c, FFM_ERR_INVALID_ARGUMENT,
                "density estimation needs 0 < max_radius <= 15, min_radius >= 0, curve > 0");
  (void)hipSetDevice(c->device);
  hipStream_t s = pick(c, stream);
  DeTable& t = c->de;
  if (t.max_r != max_radius || t.min_r != min_radius || t.curve != curve) {
    // new filters; the device copy of the old ones may still be read by
    // work queued earlier on any stream
    HIPCHECK(c, hipDeviceSynchronize());
    de_table(max_radius, min_radius, curve, t);
    int st;
    if ((st = ensure(c, c->de_h2, t.h2.size() * 4, "estimator table"))) return st;
    if ((st = ensure(c, c->de_inv, t.inv.size() * 4, "estimator table"))) return st;
    HIPCHECK(c, hipMemcpy(c->de_h2.p, t.h2.data(), t.h2.size() * 4, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->de_inv.p, t.inv.data(), t.inv.size() * 4, hipMemcpyHostToDevice));
  }
  HIPCHECK(c, launch_density_estimate(in, out, width, height,
                                      reinterpret_cast<const float*>(c->de_h2.p),
                                      reinterpret_cast<const float*>(c->de_inv.p),
                                      (int)t.h2.size(), t.R, curve, s));
  return FFM_OK;
}

int ffm_accumulate(ffm_ctx* c, const void* hist, int hist_is_f32, const float* prev, float* out,
                   uint32_t* count, int width, int height, int mode, float blend_factor,
                   void* stream) {
  if (!c || !hist || !out || width < 1 || height < 1) return FFM_ERR_INVALID_ARGUMENT;
  if (mode < 0 || mode > 2) return fail(c, FFM_ERR_INVALID_ARGUMENT, "accumulate mode %d", mode);
  if (!prev && mode != 2) return fail(c, FFM_ERR_INVALID_ARGUMENT, "prev is NULL");
  (void)hipSetDevice(c->device);
  HIPCHECK(c, launch_accumulate(hist, hist_is_f32, prev ? prev : out, out, count, width, height,
                                mode, blend_factor, pick(c, stream)));
  return FFM_OK;
}

int ffm_tonemap(ffm_ctx* c, const float* accum, uint8_t* rgba, int width, int height, float zoom,
                double samples_in_buffer, const ffm_tonemap_params* tp, void* stream) {
  if (!c || !accum || !rgba || !tp || width < 1 || height < 1) return FFM_ERR_INVALID_ARGUMENT;
  if (tp->tonemap_mode < 0 || tp->tonemap_mode > 2 || tp->highlight_mode < 0 ||
      tp->highlight_mode > 3 || tp->transparent_mode < 0 || tp->transparent_mode > 2)
    return fail(c, FFM_ERR_INVALID_ARGUMENT,
                "tonemap mode %d / highlight mode %d / transparent mode %d not supported",
                tp->tonemap_mode, tp->highlight_mode, tp->transparent_mode);
  if (!curve_ok(*tp))
    return fail(c, FFM_ERR_INVALID_ARGUMENT,
                "tone curve needs 2..%d finite points with non-decreasing x (curve_n %d)",
                FFM_MAX_CURVE_POINTS, tp->curve_n);
  (void)hipSetDevice(c->device);
  TonemapConst t = make_tonemap(*tp, width, height, zoom, samples_in_buffer);
  HIPCHECK(c, launch_tonemap(accum, rgba, width, height, t, pick(c, stream)));
  return FFM_OK;
}

// ---- the tail's effects ---------------------------------------------------

// 1: a density effect, 2: a colour effect, 0: not an effect this path runs.
static int fx_chain(int32_t type) {
  switch (type) {
    case FFM_FX_DENSITY_BLUR:
    case FFM_FX_BILATERAL_BLUR:
    case FFM_FX_SHARPEN:
      return 1;
    case FFM_FX_VIGNETTE:
    case FFM_FX_FILM_GRAIN:
    case FFM_FX_CHROMATIC_ABERRATION:
    case FFM_FX_HUE_SHIFT:
    case FFM_FX_SOBEL_EDGES:
    case FFM_FX_KALEIDOSCOPE:
    case FFM_FX_PLASMA:
    case FFM_FX_TUNNEL:
    case FFM_FX_JULIA:
    case FFM_FX_WORLEY_NOISE:
    case FFM_FX_SIMPLEX_NOISE:
    case FFM_FX_DOMAIN_WARP:
    case FFM_FX_BLOOM:
    case FFM_FX_DISTANCE_FIELD:
    case FFM_FX_TRAILS:
      return 2;
    default:
      return 0;
  }
}

// The Rhai PARAMS defaults, in declaration order (assets/effects/*.rhai).
int ffm_effect_defaults(int32_t type, ffm_effect* out) {
  if (!out) return FFM_ERR_INVALID_ARGUMENT;
  ffm_effect e{};
  e.type = type;
  e.enabled = 1;
  std::vector<float> v;
  switch (type) {
    case FFM_FX_DENSITY_BLUR: v = {3.0f, 0.25f, 0.5f}; break;          // density_blur.rhai:10-19
    case FFM_FX_BILATERAL_BLUR: v = {3.0f, 3.0f, 0.1f}; break;         // bilateral_blur.rhai
    case FFM_FX_SHARPEN: v = {1.0f, 1.0f}; break;                      // sharpen.rhai
    case FFM_FX_VIGNETTE: v = {0.5f, 0.5f, 0.5f, 0.0f}; break;         // vignette.rhai:9-18
    case FFM_FX_FILM_GRAIN: v = {0.1f, 0.0f, 4.0f}; break;             // film_grain.rhai
    case FFM_FX_CHROMATIC_ABERRATION: v = {2.0f, 1.0f, 1.0f, 0.0f}; break;
    case FFM_FX_HUE_SHIFT: v = {0.0f, 1.0f, 0.0f}; break;              // hue_shift.rhai
    case FFM_FX_SOBEL_EDGES: v = {1.0f, 0.1f, 0.5f, 0.5f, 1.0f}; break;  // sobel_edges.rhai
    case FFM_FX_KALEIDOSCOPE: v = {6.0f, 0.0f, 1.0f, 1.0f, 0.0f, 1.0f, 0.0f}; break;
    case FFM_FX_PLASMA: v = {0.3f, 3.0f, 1.0f, 0.0f, 1.0f, 225.0f}; break;
    case FFM_FX_TUNNEL: v = {1.0f, 0.5f, 0.0f, 0.0f, 1.0f, 0.0f}; break;
    case FFM_FX_JULIA: v = {1.0f, -0.7f, 0.27f, 100.0f, 1.0f, 0.0f, 0.0f, 1.0f, 1.0f, 0.0f}; break;
    case FFM_FX_WORLEY_NOISE: v = {0.5f, 8.0f, 0.0f, 0.0f, 1.0f, 4.0f, 225.0f}; break;
    case FFM_FX_SIMPLEX_NOISE: v = {0.5f, 5.0f, 4.0f, 0.0f, 0.0f, 4.0f, 225.0f}; break;
    case FFM_FX_DOMAIN_WARP: v = {0.1f, 3.0f, 4.0f, 0.0f, 0.0f, 225.0f}; break;
    case FFM_FX_BLOOM: v = {0.45f, 0.3f, 1.5f, 0.7f}; break;  // bloom.rhai (tint white)
    case FFM_FX_DISTANCE_FIELD: v = {0.15f, 0.25f, 0.0f, 1.0f, 12.0f}; break;
    case FFM_FX_TRAILS: v = {0.85f, 1.0f}; break;
    default: return FFM_ERR_INVALID_ARGUMENT;
  }
  for (size_t i = 0; i < v.size(); i++) e.params[i] = v[i];
  *out = e;
  return FFM_OK;
}

// sharpen's tap loop spans ceil(radius * pixel_scale) pixels each way; its
// parameter range is 0.5..5 (sharpen.rhai), and a radius past 64 pixels
// is refused rather than run as a near-endless loop.
constexpr float FX_MAX_SHARPEN_RADIUS = 64.0f;
// julia's escape loop runs up to max_iter steps a pixel (UI range 10..500).
constexpr float FX_MAX_JULIA_ITER = 5000.0f;
static int check_effect_params(ffm_ctx* c, const ffm_effect& e, float pixel_scale) {
  if (e.type == FFM_FX_SHARPEN && !(e.params[1] * pixel_scale <= FX_MAX_SHARPEN_RADIUS))
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "sharpen radius %g x pixel scale %g: at most %g",
                e.params[1], pixel_scale, FX_MAX_SHARPEN_RADIUS);
  if (e.type == FFM_FX_JULIA && !(e.params[3] <= FX_MAX_JULIA_ITER))
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "julia max_iter %g: at most %g", e.params[3],
                FX_MAX_JULIA_ITER);
  return FFM_OK;
}

static int check_effects(ffm_ctx* c, const ffm_effects& fx) {
  if (fx.n_density < 0 || fx.n_density > FFM_MAX_EFFECTS || fx.n_color < 0 ||
      fx.n_color > FFM_MAX_EFFECTS)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "effect counts %d / %d outside 0..%d", fx.n_density,
                fx.n_color, FFM_MAX_EFFECTS);
  if (!(fx.pixel_scale > 0.0f) || !std::isfinite(fx.pixel_scale))
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "pixel_scale must be positive");
  for (int i = 0; i < fx.n_density; i++) {
    if (!fx.density[i].enabled) continue;
    if (fx_chain(fx.density[i].type) != 1)
      return fail(c, FFM_ERR_INVALID_ARGUMENT, "density effect %d: type %d is not a density effect",
                  i, fx.density[i].type);
    if (int st = check_effect_params(c, fx.density[i], fx.pixel_scale)) return st;
  }
  for (int i = 0; i < fx.n_color; i++) {
    if (!fx.color[i].enabled) continue;
    if (fx_chain(fx.color[i].type) != 2)
      return fail(c, FFM_ERR_INVALID_ARGUMENT, "colour effect %d: type %d is not a colour effect",
                  i, fx.color[i].type);
    if (int st = check_effect_params(c, fx.color[i], fx.pixel_scale)) return st;
  }
  return FFM_OK;
}

int ffm_effects_set(ffm_ctx* c, const ffm_effects* fx) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  if (!fx) {
    c->fx = ffm_effects{};
    c->fx.pixel_scale = 1.0f;
    return FFM_OK;
  }
  int st = check_effects(c, *fx);
  if (st) return st;
  c->fx = *fx;
  return FFM_OK;
}

static FxConst fx_const(const ffm_effect& e, int W, int H, float density_mean, float pixel_scale) {
  FxConst k{};
  k.type = e.type;
  for (int i = 0; i < FFM_MAX_EFFECT_PARAMS; i++) k.p[i] = e.params[i];
  k.W = W;
  k.H = H;
  k.density_mean = density_mean;
  k.pixel_scale = pixel_scale;
  return k;
}

int ffm_density_effect(ffm_ctx* c, const ffm_effect* fx, const float* in, float* out, int width,
                       int height, float density_mean, float pixel_scale, void* stream) {
  if (!c || !fx || !in || !out || width < 1 || height < 1) return FFM_ERR_INVALID_ARGUMENT;
  if (in == out) return fail(c, FFM_ERR_INVALID_ARGUMENT, "an effect reads and writes different buffers");
  if (fx_chain(fx->type) != 1)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "type %d is not a density effect", fx->type);
  if (!(pixel_scale > 0.0f)) return fail(c, FFM_ERR_INVALID_ARGUMENT, "pixel_scale must be positive");
  if (int st = check_effect_params(c, *fx, pixel_scale)) return st;
  (void)hipSetDevice(c->device);
  HIPCHECK(c, launch_density_effect(fx_const(*fx, width, height, density_mean, pixel_scale), in,
                                    out, pick(c, stream)));
  return FFM_OK;
}

// One colour effect from `in` to `out`: bloom's passes over its
// pyramids, any other effect's one pass.
static int color_effect_on(ffm_ctx* c, const ffm_effect& e, const uint8_t* in, uint8_t* out, int W,
                           int H, float pixel_scale, hipStream_t s) {
  if (e.type == FFM_FX_TRAILS) {
    // trails.rhai keeps max(c, history * fade) in a buffer that persists
    // across frames; a still starts from a zero history (EffectClock::STILL,
    // src/renderer/effect_passes.rs:18-23), so the trail is the picture and
    // c + (trail - c) * intensity is c.
    HIPCHECK(c, hipMemcpyAsync(out, in, (size_t)W * H * 4, hipMemcpyDeviceToDevice, s));
    return FFM_OK;
  }
  if (e.type == FFM_FX_DISTANCE_FIELD) {
    int st;
    if ((st = ensure(c, c->fx_pyr, (size_t)W * H * 16, "distance field coordinates"))) return st;
    DistArgs d{};
    d.in = in;
    d.out = out;
    d.a = reinterpret_cast<float*>(c->fx_pyr.p);
    d.b = d.a + (size_t)W * H * 2;
    d.W = W;
    d.H = H;
    d.threshold = e.params[0];
    d.spread = e.params[1];
    d.mode = e.params[2];
    d.intensity = e.params[3];
    d.band_count = e.params[4];
    HIPCHECK(c, launch_distance_field(d, s));
    return FFM_OK;
  }
  if (e.type != FFM_FX_BLOOM) {
    HIPCHECK(c, launch_color_effect(fx_const(e, W, H, 0.0f, pixel_scale), in, out, s));
    return FFM_OK;
  }
  BloomArgs b{};
  b.in = in;
  b.out = out;
  b.W = W;
  b.H = H;
  b.levels = bloom_levels(W, H, b.lw, b.lh);
  size_t texels = 0;
  for (int l = 0; l < b.levels; l++) texels += (size_t)b.lw[l] * b.lh[l];
  int st;
  if ((st = ensure(c, c->fx_pyr, 2 * texels * 16, "bloom pyramids"))) return st;
  float* p = reinterpret_cast<float*>(c->fx_pyr.p);
  for (int l = 0; l < b.levels; l++) {
    b.down[l] = p;
    p += (size_t)b.lw[l] * b.lh[l] * 4;
  }
  for (int l = 0; l < b.levels; l++) {
    b.up[l] = p;
    p += (size_t)b.lw[l] * b.lh[l] * 4;
  }
  b.threshold = e.params[0];
  b.softness = e.params[1];
  b.intensity = e.params[2];
  b.spread = e.params[3];
  HIPCHECK(c, launch_bloom(b, s));
  return FFM_OK;
}

int ffm_color_effect(ffm_ctx* c, const ffm_effect* fx, const uint8_t* in, uint8_t* out, int width,
                     int height, float pixel_scale, void* stream) {
  if (!c || !fx || !in || !out || width < 1 || height < 1) return FFM_ERR_INVALID_ARGUMENT;
  if (in == out) return fail(c, FFM_ERR_INVALID_ARGUMENT, "an effect reads and writes different buffers");
  if (fx_chain(fx->type) != 2)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "type %d is not a colour effect", fx->type);
  if (!(pixel_scale > 0.0f)) return fail(c, FFM_ERR_INVALID_ARGUMENT, "pixel_scale must be positive");
  if (int st = check_effect_params(c, *fx, pixel_scale)) return st;
  (void)hipSetDevice(c->device);
  return color_effect_on(c, *fx, in, out, width, height, pixel_scale, pick(c, stream));
}

// The density chain over `acc` (each enabled effect reads the previous
// one's output; run_density_effects, src/renderer/effect_chain.rs:737-829):
// returns the buffer the tone map reads.
static int run_density_chain(ffm_ctx* c, const ffm_effects& fx, float* acc, int W, int H,
                             float density_mean, hipStream_t s, float** out) {
  *out = acc;
  const size_t npix = (size_t)W * H;
  float* cur = acc;
  int st;
  for (int i = 0; i < fx.n_density; i++) {
    if (!fx.density[i].enabled) continue;
    if ((st = ensure(c, c->fx_a, npix * 16, "effect buffer")) ||
        (st = ensure(c, c->fx_b, npix * 16, "effect buffer")))
      return st;
    float* a = reinterpret_cast<float*>(c->fx_a.p);
    float* dst = cur == a ? reinterpret_cast<float*>(c->fx_b.p) : a;
    HIPCHECK(c, launch_density_effect(fx_const(fx.density[i], W, H, density_mean, fx.pixel_scale),
                                      cur, dst, s));
    cur = dst;
  }
  *out = cur;
  return FFM_OK;
}

// The colour chain on the tone-mapped pixels `pix` (run_color_effects,
// src/renderer/effect_chain.rs:874-962), ending in `pix`.
static int run_color_chain(ffm_ctx* c, const ffm_effects& fx, uint8_t* pix, int W, int H,
                           hipStream_t s) {
  const size_t npix = (size_t)W * H;
  uint8_t* cur = pix;
  int st;
  for (int i = 0; i < fx.n_color; i++) {
    if (!fx.color[i].enabled) continue;
    if ((st = ensure(c, c->fx_rgba, npix * 4, "effect pixels"))) return st;
    uint8_t* dst = cur == pix ? reinterpret_cast<uint8_t*>(c->fx_rgba.p) : pix;
    if ((st = color_effect_on(c, fx.color[i], cur, dst, W, H, fx.pixel_scale, s))) return st;
    cur = dst;
  }
  if (cur != pix) HIPCHECK(c, hipMemcpyAsync(pix, cur, npix * 4, hipMemcpyDeviceToDevice, s));
  return FFM_OK;
}

// render_with (src/renderer/render.rs:483-800): dispatch until the target,
// accumulate every batch_size dispatches (blur first when asked) and once
// more for a partial batch, stop early after burn_in/ipt + 4 empty
// dispatches; then the density estimator and the tonemap.
// EXACT_COUNT_AFTER_PASSES (src/renderer/compute_kernel.rs:360-372): from
// this accumulate pass on a render counts each pixel's density exactly in 64
// bits (the f32 running sum drifts past ~10^6 passes). FFM_EXACT_COUNT_AFTER
// in the environment overrides it, as FAE_EXACT_COUNT_AFTER does there.
static uint64_t exact_count_after() {
  const char* e = std::getenv("FFM_EXACT_COUNT_AFTER");
  if (e && *e) {
    char* end = nullptr;
    unsigned long long v = std::strtoull(e, &end, 10);
    if (end && *end == 0) return v;
  }
  return 1ull << 18;
}

// `supersample` 2: the job renders at twice the output size and the final
// picture (after the colour effects) is boxed down into rgba_out / rgba_dev,
// which are then (width / 2) x (height / 2) (downsample_2x_firefly,
// src/export/supersample.rs:117-187).
static int render_impl(ffm_ctx* c, const ffm_job* job, uint8_t* rgba_out, float* accum_out,
                       ffm_render_stats* stats, uint8_t* rgba_dev = nullptr,
                       const ffm_effects* fx = nullptr, int supersample = 1) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  if (!job) return fail(c, FFM_ERR_INVALID_ARGUMENT, "job is NULL");
  if (!c->has_flame) return fail(c, FFM_ERR_NO_FLAME, "no flame set");
  if (!c->has_palette) return fail(c, FFM_ERR_NO_PALETTE, "No palette found");
  int st = check_view(c, &job->view);
  if (st) return st;
  if (job->walkers == 0 || job->walkers > 0xFFFFFFFFull || job->ipt == 0)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "walkers and iterations per thread must be positive");
  const int mode = job->splat_mode;
  if (mode < FFM_SPLAT_U32_GLOBAL || mode > FFM_SPLAT_F32_LDS)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "splat mode %d not available here", mode);
  const bool f32h = mode == FFM_SPLAT_F32_GLOBAL || mode == FFM_SPLAT_F32_LDS;
  if (f32h && job->filter_radius > 0.0f)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "the histogram blur reads a u32 histogram");
  if (supersample == 2 && (job->view.width % 2 || job->view.height % 2))
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "a 2x supersampled render needs an even size (%dx%d)",
                job->view.width, job->view.height);
  // the tone map's parameters are checked before any work, also when the
  // render only returns the accumulator
  if ((st = check_tonemap(c, job->tonemap))) return st;
  (void)hipSetDevice(c->device);
  c->cancel.store(0);
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = c->stream;
  const int W = job->view.width, H = job->view.height;
  const size_t npix = (size_t)W * H, words = npix * 4;
  if ((st = ensure(c, c->hist, words * 4, "histogram"))) return st;
  if ((st = ensure(c, c->accum, npix * 16, "accumulator"))) return st;
  if (job->filter_radius > 0.0f && (st = ensure(c, c->scratch, words * 4, "blur scratch"))) return st;
  if (mode == FFM_SPLAT_U32_XCD &&
      (st = ensure(c, c->copies, words * 4 * XCD_COPIES, "per-XCD histograms")))
    return st;
  if (job->persistent_orbits && (st = ensure(c, c->orbits, job->walkers * 16, "orbits"))) return st;
  if ((st = ensure(c, c->counters, 16, "counters"))) return st;
  HIPCHECK(c, hipMemsetAsync(c->accum.p, 0, npix * 16, s));

  IterArgs a = base_args(c, job->view);
  a.walkers = job->walkers;
  a.ipt = job->ipt;
  a.burn_in = job->burn_in;
  a.orbits = job->persistent_orbits ? reinterpret_cast<float4*>(c->orbits.p) : nullptr;
  a.refresh_period = job->persistent_orbits ? c->refresh_period : 0u;
  if (job->persistent_orbits && c->flame.has_xaos && job->burn_in >= (1u << 24))
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "xaos orbits keep burn-in below 2^24 (got %u)",
                job->burn_in);
  if (mode == FFM_SPLAT_U32_XCD) {
    a.hist = c->copies.p;
    a.hist_copy_words = (uint32_t)words;
  } else {
    a.hist = c->hist.p;
  }
  const int grid = iterate_grid(c, job->walkers);
  OrbitFuses fuses;
  const uint64_t exact_after = exact_count_after();
  uint64_t passes = 0;
  uint32_t* exact = nullptr;  // the 64-bit counts, once the switch is made
  uint64_t total = 0, batch_samples = 0, samples_in_buffer = 0, k = 0;
  uint32_t batch_count = 0, empty = 0;
  const uint32_t stall_limit = job->burn_in / std::max(job->ipt, 1u) + 4;
  const uint32_t bs = job->batch_size ? job->batch_size : 4;
  double iterate_ms = 0.0, walker_ms = 0.0, accum_ms = 0.0;
  uint64_t launches = 0;
  c->kev_used = 0;
  c->dev_used = 0;
  // Dispatches are enqueued back to back: the host never waits for one to
  // finish before launching the next (the sample count each adds is the
  // host's OrbitFuses model), only every SYNC_DISPATCHES dispatches, to read
  // the timing events back and to poll for cancellation with little queued.
  constexpr uint64_t SYNC_DISPATCHES = 16;
  for (;;) {
    if (c->cancel.load()) {
      (void)hipStreamSynchronize(s);
      c->kev_used = c->dev_used = 0;
      return fail(c, FFM_ERR_CANCELLED, "Render cancelled");
    }
    if (batch_count == 0) {
      if (mode == FFM_SPLAT_U32_XCD)
        HIPCHECK(c, launch_clear_u32(c->copies.p, words * XCD_COPIES, s));
      else
        HIPCHECK(c, launch_clear_u32(c->hist.p, words, s));
    }
    a.seed = job->seed_base + (uint32_t)k * job->seed_stride;
    a.fresh = k == 0;
    HIPCHECK(c, mark(c, s, 2, true));
    if (mode == FFM_SPLAT_U32_LDS || mode == FFM_SPLAT_F32_LDS) {
      if ((st = lds_splat(c, a, reinterpret_cast<uint32_t*>(c->hist.p), s, f32h))) return st;
    } else {
      HIPCHECK(c, mark(c, s, 0, true));
      HIPCHECK(c, launch_iterate(a, mode, grid, 256, s));
      HIPCHECK(c, mark(c, s, 0, false));
    }
    HIPCHECK(c, mark(c, s, 2, false));
    uint64_t plotted;
    if (job->persistent_orbits) {
      plotted = fuses.dispatch((uint32_t)job->walkers, job->ipt, job->burn_in, a.refresh_period,
                               a.seed);
    } else {
      plotted = job->walkers * (uint64_t)(job->ipt > job->burn_in ? job->ipt - job->burn_in : 0);
    }
    k++;
    total += plotted;
    batch_samples += plotted;
    batch_count++;
    empty = plotted == 0 ? empty + 1 : 0;
    const bool done = total >= job->target_samples || empty >= stall_limit;
    if (batch_count >= bs || done) {
      if (mode == FFM_SPLAT_U32_XCD)
        HIPCHECK(c, launch_fold_xcd(reinterpret_cast<uint32_t*>(c->hist.p),
                                    reinterpret_cast<const uint32_t*>(c->copies.p), words,
                                    XCD_COPIES, s));
      if (job->filter_radius > 0.0f) {
        float pixel_count = (float)W * (float)H;
        float dsig = (((float)batch_samples / pixel_count) * 100.0f) *
                     std::pow(100.0f, 2.0f * job->filter_blur_edges - 1.0f);
        HIPCHECK(c, launch_blur(reinterpret_cast<uint32_t*>(c->hist.p),
                                reinterpret_cast<uint32_t*>(c->scratch.p), W, H, job->filter_radius,
                                dsig, s));
      }
      if (++passes >= exact_after && !exact) {
        // counts start at zero and are seeded from the running sum there
        if ((st = ensure(c, c->count, npix * 8, "exact density counts"))) return st;
        HIPCHECK(c, launch_clear_u32(c->count.p, npix * 2, s));
        exact = reinterpret_cast<uint32_t*>(c->count.p);
      }
      HIPCHECK(c, launch_accumulate(c->hist.p, f32h,
                                    reinterpret_cast<float*>(c->accum.p),
                                    reinterpret_cast<float*>(c->accum.p), exact, W, H, 0, 0.0f, s));
      samples_in_buffer += batch_samples;
      batch_count = 0;
      batch_samples = 0;
    }
    if (done || k % SYNC_DISPATCHES == 0)
      HIPCHECK(c, collect_timings(c, walker_ms, launches, accum_ms, iterate_ms));
    if (done) break;
  }
  float* final_acc = reinterpret_cast<float*>(c->accum.p);
  if (job->de_max_radius > 0.0f) {
    if ((st = ensure(c, c->accum2, npix * 16, "accumulator"))) return st;
    st = ffm_density_estimate(c, final_acc, reinterpret_cast<float*>(c->accum2.p), W, H,
                              job->de_max_radius, job->de_min_radius, job->de_curve, s);
    if (st) return st;
    final_acc = reinterpret_cast<float*>(c->accum2.p);
  }
  const ffm_effects& effects = fx ? *fx : c->fx;
  const TonemapConst tmc = make_tonemap(job->tonemap, W, H, job->view.zoom, (double)samples_in_buffer);
  // density effects measure a pixel's hit count in multiples of the mean
  // density: the tone map's sample density (FlameRenderer::mean_density)
  if ((st = run_density_chain(c, effects, final_acc, W, H, tmc.sample_density, s, &final_acc)))
    return st;
  if (accum_out) {
    HIPCHECK(c, hipMemcpyAsync(accum_out, final_acc, npix * 16, hipMemcpyDeviceToHost, s));
  }
  if (rgba_out || rgba_dev) {
    const bool ss = supersample == 2;
    const size_t out_px = ss ? npix / 4 : npix;
    if ((!rgba_dev || ss) && (st = ensure(c, c->rgba, npix * 4, "pixels"))) return st;
    uint8_t* pix = rgba_dev && !ss ? rgba_dev : reinterpret_cast<uint8_t*>(c->rgba.p);
    HIPCHECK(c, launch_tonemap(final_acc, pix, W, H, tmc, s));
    if ((st = run_color_chain(c, effects, pix, W, H, s))) return st;
    if (ss) {
      uint8_t* dst = rgba_dev;
      if (!dst) {
        if ((st = ensure(c, c->rgba_out, out_px * 4, "downsampled pixels"))) return st;
        dst = reinterpret_cast<uint8_t*>(c->rgba_out.p);
      }
      HIPCHECK(c, launch_downsample2x(pix, dst, W / 2, H / 2, s));
      pix = dst;
    }
    if (rgba_out) {
      hipError_t e = hipMemcpyAsync(rgba_out, pix, out_px * 4, hipMemcpyDeviceToHost, s);
      if (e != hipSuccess)
        return fail(c, FFM_ERR_PIXEL_READ_FAILED, "pixel readback: %s", hipGetErrorString(e));
    }
  }
  HIPCHECK(c, hipStreamSynchronize(s));
  if (stats) {
    stats->total_samples = total;
    stats->dispatches = k;
    stats->iterate_ms = iterate_ms;
    stats->walker_kernel_ms = walker_ms;
    stats->walker_launches = launches;
    stats->tile_accum_ms = accum_ms;
    stats->render_time_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return FFM_OK;
}

int ffm_render(ffm_ctx* c, const ffm_job* job, uint8_t* rgba_out, ffm_render_stats* stats) {
  if (!rgba_out) return fail(c, FFM_ERR_INVALID_ARGUMENT, "pixel buffer is NULL");
  return render_impl(c, job, rgba_out, nullptr, stats);
}

int ffm_render_config(ffm_ctx* c, const char* text, size_t len, int width, int height,
                      uint64_t iterations, uint32_t ipt, uint32_t flags, uint8_t* rgba_out,
                      ffm_render_stats* stats) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  if (!rgba_out) return fail(c, FFM_ERR_INVALID_ARGUMENT, "pixel buffer is NULL");
  std::unique_ptr<ffm_scene> sc(new (std::nothrow) ffm_scene);
  if (!sc) return fail(c, FFM_ERR_OUT_OF_MEMORY, "out of host memory");
  char err[512] = {0};
  int st = ffm_scene_parse(text, len, 0, sc.get(), err, sizeof err);
  if (st) return fail(c, st, "%s", err);
  ffm_job job;
  st = ffm_scene_job(sc.get(), width, height, iterations, ipt, flags, &job, err, sizeof err);
  if (st) return fail(c, st, "%s", err);
  if ((st = ffm_flame_set(c, &sc->flame))) return st;
  if ((st = ffm_palette_set_xf(c, &sc->stops[0][0], sc->n_stops, sc->palette_size,
                               &sc->palette_xf)))
    return st;
  if ((st = check_effects(c, sc->effects))) return st;
  return render_impl(c, &job, rgba_out, nullptr, stats, nullptr, &sc->effects,
                     (flags & FFM_RENDER_SUPERSAMPLE_2X) ? 2 : 1);
}

int ffm_render_to_device(ffm_ctx* c, const ffm_job* job, void* rgba_dev, ffm_render_stats* stats) {
  if (!rgba_dev) return fail(c, FFM_ERR_INVALID_ARGUMENT, "pixel buffer is NULL");
  return render_impl(c, job, nullptr, nullptr, stats, reinterpret_cast<uint8_t*>(rgba_dev));
}

int ffm_render_accum(ffm_ctx* c, const ffm_job* job, float* accum_out, ffm_render_stats* stats) {
  if (!accum_out) return fail(c, FFM_ERR_INVALID_ARGUMENT, "accumulator buffer is NULL");
  return render_impl(c, job, nullptr, accum_out, stats);
}

// ---- batched gallery ---------------------------------------------------

namespace {

// Records staged per wave of a batch: an eighth of the device memory free
// when the batch is laid out (on a 288 GB MI355X: 9 G records, 36 GB), at
// least 2^28 records and no more than the batch needs, and below 2^33 (a
// flush keeps a run's start in 32 bits plus one bit beside its room).
// Fewer, larger waves cut the persistent launches' tails and let the tile
// accumulation fold larger slices: the 512-flame gallery walks in 5 waves.
constexpr uint64_t BATCH_RECORD_MIN = 1ull << 28, BATCH_RECORD_MAX = (1ull << 33) - (1ull << 24);

uint64_t batch_record_budget(uint64_t needed, size_t held_bytes) {
  size_t free_b = 0, total_b = 0;
  uint64_t b = 3ull << 30;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) b = ((uint64_t)free_b + held_bytes) / 8 / 4;
  else (void)hipGetLastError();
  b = std::min(std::max(b, BATCH_RECORD_MIN), BATCH_RECORD_MAX);
  return std::min(b, std::max(needed, BATCH_RECORD_MIN));
}
constexpr int BATCH_PALETTE = 256;
// The pre-pass walks 1/8 of wave 0's walkers, each for its burn-in and a
// quarter of its plotted iterations (rounded up, so the estimate errs high).
constexpr uint32_t BATCH_COUNT_DIV = 8, BATCH_COUNT_ITER_DIV = PRE_ITER_DIV;

struct BatchPlan {
  // what the graph was built for
  std::vector<uint64_t> sig;
  int n = 0, n_waves = 0, max_rows = 0, n_tiles = 0, max_tiles = 0;
  uint64_t total_pix = 0;
  uint32_t item_walkers = 0;
  int flush_every = 4;
  int tile_mode = 0;
  std::vector<uint32_t> wave_start;  // items of wave w: [wave_start[w], wave_start[w+1])
  std::vector<BatchPix> pix;
  std::vector<uint64_t> samples;     // plotted samples per flame
  DevBuf flames, items, tables, pals, pixd, hist, rgba, fill, region, cap, item_start, records,
      counters;
  DevBuf outp;  // per flame, where the finish kernel writes its pixels (set every call)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  void release() {
    drop_graph();
    for (DevBuf* b : {&flames, &items, &tables, &pals, &pixd, &hist, &rgba, &fill, &region, &cap,
                      &item_start, &records, &counters})
      b->release();
  }
};

struct Tuning {
  std::vector<uint64_t> sig;
  uint32_t item_walkers;
  int flush_every;
  int tile_mode;
};

// The gallery's tile shapes (log2 width, log2 height): 0 = 128x64, the one
// flame's; 1 = 64x64, whose LDS tile lets two accumulating workgroups share
// a CU, at twice the bins for the walkers to sort into.
constexpr int BATCH_TILE_MODES = 2;
void tile_dims(int mode, int* wl, int* hl) {
  *wl = mode == 1 ? 6 : TILE_WL;
  *hl = TILE_HL;
}

// Tunings kept across processes: with FFM_TUNE_CACHE naming a file, a
// gallery signature's winner is appended to it as one text line
// "<FNV-1a 64 of the signature> <walkers per item> <flush cadence> <tile
// shape>", and a first call finds it there instead of tuning again.
uint64_t sig_hash(const std::vector<uint64_t>& sig) {
  uint64_t h = 1469598103934665603ull;
  for (uint64_t v : sig)
    for (int b = 0; b < 8; b++) {
      h ^= (v >> (8 * b)) & 0xFFu;
      h *= 1099511628211ull;
    }
  return h;
}

bool tune_cache_find(const std::vector<uint64_t>& sig, int forced, Tuning* out) {
  const char* path = std::getenv("FFM_TUNE_CACHE");
  if (!path) return false;
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  const unsigned long long h = sig_hash(sig);
  unsigned long long fh;
  unsigned iw;
  int fe, mode;
  bool found = false;
  while (std::fscanf(f, "%llx %u %d %d", &fh, &iw, &fe, &mode) == 4) {
    if (fh != h || iw == 0 || fe < 1 || fe > 8 || mode < 0 || mode >= BATCH_TILE_MODES) continue;
    if (forced >= 0 && mode != forced) continue;
    *out = Tuning{sig, iw, fe, mode};
    found = true;  // the last line for a signature wins
  }
  std::fclose(f);
  return found;
}

void tune_cache_add(const Tuning& t) {
  const char* path = std::getenv("FFM_TUNE_CACHE");
  if (!path) return;
  if (FILE* f = std::fopen(path, "a")) {
    std::fprintf(f, "%016llx %u %d %d\n", (unsigned long long)sig_hash(t.sig), t.item_walkers,
                 t.flush_every, t.tile_mode);
    std::fclose(f);
  }
}

}  // namespace

struct ffm_batch_state {
  BatchPlan plan;
  std::vector<Tuning> tuned;
};

namespace {

ffm_batch_state* batch_state(ffm_ctx* c);

int check_batch(ffm_ctx* c, int n, const ffm_flame_desc* flames, const float* const* stops,
                const int* n_stops, const ffm_job* jobs, uint8_t* const* rgba_out) {
  if (n < 1) return fail(c, FFM_ERR_INVALID_ARGUMENT, "an empty batch");
  if (!flames || !stops || !n_stops || !jobs || !rgba_out)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "batch arrays must not be NULL");
  for (int i = 0; i < n; i++) {
    int st = check_flame(c, &flames[i]);
    if (st) return st;
    if (flames[i].has_xaos)
      return fail(c, FFM_ERR_INVALID_ARGUMENT, "flame %d: xaos is not part of the batched path", i);
    if ((st = check_view(c, &jobs[i].view))) return st;
    if (!curve_ok(jobs[i].tonemap))
      return fail(c, FFM_ERR_INVALID_ARGUMENT, "job %d: tone curve needs 2..%d finite points "
                  "with non-decreasing x", i, FFM_MAX_CURVE_POINTS);
    if (!stops[i] || n_stops[i] < 1) return fail(c, FFM_ERR_NO_PALETTE, "flame %d: No palette found", i);
    if (!rgba_out[i]) return fail(c, FFM_ERR_INVALID_ARGUMENT, "flame %d: pixel buffer is NULL", i);
    const ffm_job& j = jobs[i];
    if (j.walkers == 0 || j.walkers > 0xFFFFFFFFull || j.ipt == 0)
      return fail(c, FFM_ERR_INVALID_ARGUMENT, "flame %d: walkers and ipt must be positive", i);
    if (j.filter_radius > 0.0f || j.de_max_radius > 0.0f)
      return fail(c, FFM_ERR_INVALID_ARGUMENT,
                  "flame %d: blur and density estimation are not part of the batched path", i);
    const int tiles = ((j.view.width + TILE_W - 1) / TILE_W) * ((j.view.height + TILE_H - 1) / TILE_H);
    if (tiles > TILE_MAX)
      return fail(c, FFM_ERR_TOO_LARGE, "flame %d: %dx%d is too large for the batched path", i,
                  j.view.width, j.view.height);
  }
  return FFM_OK;
}

std::vector<uint64_t> batch_signature(int n, const ffm_flame_desc* flames, const ffm_job* jobs) {
  std::vector<uint64_t> sig;
  sig.push_back((uint64_t)n);
  for (int i = 0; i < n; i++) {
    sig.push_back(((uint64_t)jobs[i].view.width << 32) | (uint32_t)jobs[i].view.height);
    sig.push_back(jobs[i].walkers);
    sig.push_back(((uint64_t)jobs[i].ipt << 32) | jobs[i].burn_in);
    sig.push_back((uint64_t)(flames[i].n_xforms + flames[i].n_finals) |
                  (uint64_t)flame_ext(&flames[i]) << 32);
  }
  return sig;
}

struct BatchLayout {
  uint64_t tile_base = 0, pix_off = 0, table_off = 0, total_samples = 0;
  int max_rows = 0, max_tiles = 0;
};

// Every flame's descriptor for the kernels (its slices of the tables, tiles
// and histograms, which depend on the batch signature only, and its camera,
// seed and tone map, which do not) and its picture's tone-map constants.
BatchLayout describe_batch(BatchPlan& P, int n, const ffm_flame_desc* flames, const ffm_job* jobs,
                           int wl, int hl, std::vector<BatchFlame>& bf) {
  BatchLayout L;
  bf.assign(n, BatchFlame{});
  P.pix.assign(n, BatchPix{});
  P.samples.assign(n, 0);
  for (int i = 0; i < n; i++) {
    const ffm_job& j = jobs[i];
    const int rows = packed_rows(&flames[i]);
    L.max_rows = std::max(L.max_rows, rows);
    BatchFlame& f = bf[i];
    f.table_off = (uint32_t)L.table_off;
    f.pal_off = (uint32_t)i * BATCH_PALETTE;
    f.n_rows = rows;
    f.pal_size = BATCH_PALETTE;
    f.view = make_view(j.view);
    f.ipt = j.ipt;
    f.burn_in = j.burn_in;
    f.seed = j.seed_base;  // one dispatch per flame: seed_0
    f.tiles_x = (j.view.width + (1 << wl) - 1) >> wl;
    f.n_tiles = f.tiles_x * ((j.view.height + (1 << hl) - 1) >> hl);
    L.max_tiles = std::max(L.max_tiles, f.n_tiles);
    f.tile_base = (uint32_t)L.tile_base;
    f.hist_off = L.pix_off;
    const uint64_t npix = (uint64_t)j.view.width * j.view.height;
    P.pix[i].pix_off = L.pix_off;
    P.pix[i].npix = npix;
    P.samples[i] = j.walkers * (uint64_t)(j.ipt > j.burn_in ? j.ipt - j.burn_in : 0);
    P.pix[i].tc = make_tonemap(j.tonemap, j.view.width, j.view.height, j.view.zoom,
                               (double)P.samples[i]);
    L.tile_base += (uint64_t)f.n_tiles;
    L.pix_off += npix;
    L.table_off += flame_table_floats(rows);
    L.total_samples += j.walkers * (uint64_t)(j.ipt + 1);
  }
  return L;
}

// Lay out the batch (buffers, waves, items) and capture its launches in a
// graph. Flame tables, palettes and the flames' descriptors are uploaded
// separately, every call (a repeat call may move a camera or change a seed).
int build_batch(ffm_ctx* c, BatchPlan& P, int n, const ffm_flame_desc* flames, const ffm_job* jobs,
                uint32_t item_walkers, int flush_every, int tile_mode) {
  P.drop_graph();
  P.n = n;
  P.item_walkers = item_walkers;
  P.flush_every = flush_every;
  P.tile_mode = tile_mode;
  int wl, hl;
  tile_dims(tile_mode, &wl, &hl);
  std::vector<BatchFlame> bf;
  const BatchLayout lay = describe_batch(P, n, flames, jobs, wl, hl, bf);
  const uint64_t tile_base = lay.tile_base, pix_off = lay.pix_off, table_off = lay.table_off,
                 total_samples = lay.total_samples;
  const int max_rows = lay.max_rows, max_tiles = lay.max_tiles;
  P.max_rows = max_rows;
  P.max_tiles = max_tiles;
  P.n_tiles = (int)tile_base;
  P.total_pix = pix_off;
  // waves: each a proportional slice of every flame, sized to the record budget
  // the records buffer already held counts as free: it is reused or replaced
  const uint64_t budget = batch_record_budget(total_samples * 5 / 4, P.records.bytes);
  P.n_waves = (int)std::max<uint64_t>(1, (total_samples * 5 / 4 + budget - 1) / budget);
  std::vector<BatchItem> items;
  P.wave_start.assign(1, 0);
  for (int w = 0; w < P.n_waves; w++) {
    for (int i = 0; i < n; i++) {
      const uint64_t W = jobs[i].walkers;
      const uint64_t a = W * w / P.n_waves, b = W * (w + 1) / P.n_waves;
      for (uint64_t x = a; x < b; x += item_walkers) {
        BatchItem it{(uint32_t)i, (uint32_t)x, (uint32_t)std::min<uint64_t>(b, x + item_walkers), 0u};
        items.push_back(it);
      }
    }
    P.wave_start.push_back((uint32_t)items.size());
  }
  int st;
  if ((st = ensure(c, P.flames, bf.size() * sizeof(BatchFlame), "batch flames"))) return st;
  if ((st = ensure(c, P.items, items.size() * sizeof(BatchItem), "batch items"))) return st;
  if ((st = ensure(c, P.tables, table_off * 4, "batch tables"))) return st;
  if ((st = ensure(c, P.pals, (size_t)n * BATCH_PALETTE * 4, "batch palettes"))) return st;
  if ((st = ensure(c, P.pixd, P.pix.size() * sizeof(BatchPix), "batch images"))) return st;
  if ((st = ensure(c, P.hist, P.total_pix * 16, "batch histograms"))) return st;
  if ((st = ensure(c, P.rgba, P.total_pix * 4, "batch pixels"))) return st;
  if ((st = ensure(c, P.outp, (size_t)n * sizeof(void*), "batch pixel pointers"))) return st;
  const size_t tiles_pad = ((size_t)P.n_tiles + 8) & ~(size_t)3;
  if ((st = ensure(c, P.fill, tiles_pad * 4, "batch tiles"))) return st;
  if ((st = ensure(c, P.region, tiles_pad * 8, "batch tiles"))) return st;
  if ((st = ensure(c, P.cap, tiles_pad * 4, "batch tiles"))) return st;
  if ((st = ensure(c, P.item_start, tiles_pad * 4, "batch tiles"))) return st;
  if ((st = ensure(c, P.records, budget * 4, "batch records"))) return st;
  const size_t n_counters = ((size_t)P.n_waves + 1 + 3) & ~(size_t)3;
  if ((st = ensure(c, P.counters, n_counters * 4, "batch counters"))) return st;
  uint32_t* counters = reinterpret_cast<uint32_t*>(P.counters.p);
  hipStream_t s = c->stream;
  HIPCHECK(c, hipMemcpyAsync(P.flames.p, bf.data(), bf.size() * sizeof(BatchFlame),
                             hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipMemcpyAsync(P.items.p, items.data(), items.size() * sizeof(BatchItem),
                             hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipMemcpyAsync(P.pixd.p, P.pix.data(), P.pix.size() * sizeof(BatchPix),
                             hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipStreamSynchronize(s));

  // the launch sequence, captured once
  BatchArgs b{};
  b.flames = reinterpret_cast<const BatchFlame*>(P.flames.p);
  b.n_flames = n;
  b.tables = reinterpret_cast<const float*>(P.tables.p);
  b.pal_u32 = reinterpret_cast<const uint32_t*>(P.pals.p);
  b.max_rows = max_rows;
  b.max_pal = BATCH_PALETTE;
  b.count_div = BATCH_COUNT_DIV;
  b.count_iter_div = BATCH_COUNT_ITER_DIV;
  for (int i = 0; i < n; i++) b.ext |= flame_ext(&flames[i]) ? 1 : 0;
  b.g.tile_fill = reinterpret_cast<uint32_t*>(P.fill.p);
  b.g.region = reinterpret_cast<uint64_t*>(P.region.p);
  b.g.cap = reinterpret_cast<uint32_t*>(P.cap.p);
  b.g.item_start = reinterpret_cast<uint32_t*>(P.item_start.p);
  b.g.records = reinterpret_cast<uint32_t*>(P.records.p);
  b.g.hist = reinterpret_cast<uint32_t*>(P.hist.p);
  b.g.n_tiles = P.n_tiles;
  b.g.flush_every = flush_every;
  b.g.tile_wl = wl;
  b.g.tile_hl = hl;
  // the largest flame's bins: a workgroup sorts one flame's records at a time
  b.g.bins_lds = (P.max_tiles + 63) / 64 * 64;
  bin_shape(max_rows, BATCH_PALETTE, b.g.bins_lds, &b.g.walk_block, &b.g.stage, &b.g.walk_per_cu,
            3);
  const int grid = std::max(c->num_cu, 1);
  const BatchItem* items_d = reinterpret_cast<const BatchItem*>(P.items.p);
  HIPCHECK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const char* step = "start";
  auto body = [&]() -> hipError_t {
    hipError_t e;
    if ((e = launch_clear_u32(P.hist.p, P.total_pix * 4, s)) != hipSuccess) return e;
    if ((e = launch_clear_u32(P.fill.p, tiles_pad, s)) != hipSuccess) return e;
    if ((e = launch_clear_u32(counters, n_counters, s)) != hipSuccess) return e;
    BatchArgs w0 = b;
    w0.items = items_d;
    w0.n_items = (int32_t)P.wave_start[1];
    w0.next_item = counters + P.n_waves;
    step = "count pre-pass";
    if ((e = launch_bin_batch(w0, true, grid, s)) != hipSuccess) return e;
    step = "capacities";
    const double q = capacity_scale();
    if ((e = launch_tile_capacity_n(b.g, (float)(BATCH_COUNT_DIV * BATCH_COUNT_ITER_DIV * q),
                                    capacity_margin(q), budget, s)) !=
        hipSuccess)
      return e;
    for (int w = 0; w < P.n_waves; w++) {
      BatchArgs bw = b;
      bw.items = items_d + P.wave_start[w];
      bw.n_items = (int32_t)(P.wave_start[w + 1] - P.wave_start[w]);
      bw.next_item = counters + w;
      if (w > 0 && (e = launch_clear_u32(P.fill.p, tiles_pad, s)) != hipSuccess) return e;
      step = "walkers";
      if ((e = launch_bin_batch(bw, false, grid, s)) != hipSuccess) return e;
      step = "tile accumulate";
      if ((e = launch_tile_accum_batch(bw, grid, s)) != hipSuccess) return e;
    }
    step = "accumulate + tonemap";
    return launch_batch_finish(reinterpret_cast<const uint32_t*>(P.hist.p),
                               reinterpret_cast<uint8_t* const*>(P.outp.p),
                               reinterpret_cast<const BatchPix*>(P.pixd.p), n, s);
  };
  hipError_t eb = body();
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(s, &g);
  if (eb != hipSuccess || ee != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(c, FFM_ERR_DEVICE, "capturing the batch (%s): %s", step,
                hipGetErrorString(eb != hipSuccess ? eb : ee));
  }
  P.graph = g;
  HIPCHECK(c, hipGraphInstantiate(&P.exec, g, nullptr, nullptr, 0));
  P.sig = batch_signature(n, flames, jobs);
  return FFM_OK;
}

// Per-call uploads: flame tables and palettes (their content may change
// between calls with the same layout) and where each flame's pixels go: the
// caller's device buffers, or (dev_out NULL) the plan's pixel buffer.
int upload_batch(ffm_ctx* c, BatchPlan& P, const ffm_flame_desc* flames, const ffm_job* jobs,
                 const float* const* stops, const int* n_stops, uint8_t* const* dev_out) {
  int wl, hl;
  tile_dims(P.tile_mode, &wl, &hl);
  std::vector<BatchFlame> bf;
  describe_batch(P, P.n, flames, jobs, wl, hl, bf);
  // every flame's table and palette are independent: packed on up to 16
  // host threads (the selection buckets and palette lookup tables are most
  // of a gallery's host time), then laid end to end in flame order
  std::vector<std::vector<float>> packed((size_t)P.n);
  std::vector<uint32_t> pals((size_t)P.n * BATCH_PALETTE);
  auto pack_range = [&](int i0, int i1) {
    std::vector<float> lin;
    std::vector<uint32_t> pk;
    for (int i = i0; i < i1; i++) {
      std::unique_ptr<ffm_flame_desc> f(new ffm_flame_desc(flames[i]));  // large: not on the stack
      prepare_flame(f.get());
      packed[(size_t)i] = pack_flame(f.get());
      palette_tables(stops[i], n_stops[i], BATCH_PALETTE, pk, lin);
      std::copy(pk.begin(), pk.end(), pals.begin() + (size_t)i * BATCH_PALETTE);
    }
  };
  const int nt = std::max(1, std::min({16, (int)std::thread::hardware_concurrency(), P.n / 16}));
  if (nt == 1) {
    pack_range(0, P.n);
  } else {
    std::vector<std::thread> pool;
    for (int k = 0; k < nt; k++)
      pool.emplace_back(pack_range, (int)((int64_t)P.n * k / nt), (int)((int64_t)P.n * (k + 1) / nt));
    for (std::thread& th : pool) th.join();
  }
  std::vector<float> tables;
  for (const std::vector<float>& t : packed) tables.insert(tables.end(), t.begin(), t.end());
  std::vector<uint8_t*> outp((size_t)P.n);
  for (int i = 0; i < P.n; i++)
    outp[i] = dev_out ? dev_out[i] : reinterpret_cast<uint8_t*>(P.rgba.p) + P.pix[i].pix_off * 4;
  hipStream_t s = c->stream;
  HIPCHECK(c, hipMemcpyAsync(P.tables.p, tables.data(), tables.size() * 4, hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipMemcpyAsync(P.pals.p, pals.data(), pals.size() * 4, hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipMemcpyAsync(P.outp.p, outp.data(), outp.size() * sizeof(void*),
                             hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipMemcpyAsync(P.flames.p, bf.data(), bf.size() * sizeof(BatchFlame),
                             hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipMemcpyAsync(P.pixd.p, P.pix.data(), P.pix.size() * sizeof(BatchPix),
                             hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipStreamSynchronize(s));  // the host vectors die on return
  return FFM_OK;
}

int run_batch_graph(ffm_ctx* c, BatchPlan& P, float* gpu_ms) {
  hipStream_t s = c->stream;
  HIPCHECK(c, hipEventRecord(c->ev0, s));
  HIPCHECK(c, hipGraphLaunch(P.exec, s));
  HIPCHECK(c, hipEventRecord(c->ev1, s));
  HIPCHECK(c, hipEventSynchronize(c->ev1));
  HIPCHECK(c, hipEventElapsedTime(gpu_ms, c->ev0, c->ev1));
  return FFM_OK;
}

}  // namespace

namespace {
// The batched render; `dev_out` names device buffers for the pictures, else
// they are read back into the host buffers `rgba_out`.
int batch_render_impl(ffm_ctx* c, int n, const ffm_flame_desc* flames, const float* const* stops,
                      const int* n_stops, const ffm_job* jobs, uint8_t* const* rgba_out,
                      uint8_t* const* dev_out, ffm_render_stats* stats) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  int st = check_batch(c, n, flames, stops, n_stops, jobs, dev_out ? dev_out : rgba_out);
  if (st) return st;
  (void)hipSetDevice(c->device);
  auto t0 = std::chrono::steady_clock::now();
  ffm_batch_state* B = batch_state(c);
  BatchPlan& P = B->plan;
  const std::vector<uint64_t> sig = batch_signature(n, flames, jobs);
  // FFM_BATCH_TILE=128x64 or 64x64 fixes the tile shape instead of tuning it
  int forced = -1;
  if (const char* e = std::getenv("FFM_BATCH_TILE")) {
    if (std::strcmp(e, "128x64") == 0) forced = 0;
    else if (std::strcmp(e, "64x64") == 0) forced = 1;
    else return fail(c, FFM_ERR_INVALID_ARGUMENT, "FFM_BATCH_TILE is 128x64 or 64x64, not %s", e);
  }
  const Tuning* tune = nullptr;
  for (const Tuning& t : B->tuned)
    if (t.sig == sig && (forced < 0 || t.tile_mode == forced)) tune = &t;
  float gpu_ms = 0.0f;
  Tuning cached;
  if (!tune && tune_cache_find(sig, forced, &cached)) {
    B->tuned.push_back(cached);
    tune = &B->tuned.back();
    P.sig.clear();
  }
  if (!tune) {
    static const uint32_t kItems[] = {2048, 8192, 32768};
    static const int kFlush[] = {2, 4, 8};
    float best = 1e30f;
    const int mode0 = forced < 0 ? 0 : forced;
    if (mode0 != 0) {
      int wl, hl;
      tile_dims(mode0, &wl, &hl);
      for (int i = 0; i < n; i++)
        if ((int64_t)((jobs[i].view.width + (1 << wl) - 1) >> wl) *
                ((jobs[i].view.height + (1 << hl) - 1) >> hl) > TILE_MAX)
          return fail(c, FFM_ERR_TOO_LARGE, "flame %d: %dx%d needs more than %d tiles of %dx%d", i,
                      jobs[i].view.width, jobs[i].view.height, TILE_MAX, 1 << wl, 1 << hl);
    }
    Tuning tb{sig, 65536, 4, mode0};
    for (uint32_t iw : kItems)
      for (int fe : kFlush) {
        if ((st = build_batch(c, P, n, flames, jobs, iw, fe, mode0))) return st;
        if ((st = upload_batch(c, P, flames, jobs, stops, n_stops, dev_out))) return st;
        float ms = 0.0f;
        if ((st = run_batch_graph(c, P, &ms))) return st;
        if (ms < best) {
          best = ms;
          tb.item_walkers = iw;
          tb.flush_every = fe;
        }
      }
    for (int mode = 1; forced < 0 && mode < BATCH_TILE_MODES; mode++) {
      int wl, hl;
      tile_dims(mode, &wl, &hl);
      bool fits = true;
      for (int i = 0; i < n; i++)
        fits = fits && (int64_t)((jobs[i].view.width + (1 << wl) - 1) >> wl) *
                               ((jobs[i].view.height + (1 << hl) - 1) >> hl) <= TILE_MAX;
      if (!fits) continue;
      if ((st = build_batch(c, P, n, flames, jobs, tb.item_walkers, tb.flush_every, mode)))
        return st;
      if ((st = upload_batch(c, P, flames, jobs, stops, n_stops, dev_out))) return st;
      float ms = 0.0f;
      if ((st = run_batch_graph(c, P, &ms))) return st;
      if (ms < best) {
        best = ms;
        tb.tile_mode = mode;
      }
    }
    B->tuned.push_back(tb);
    tune = &B->tuned.back();
    tune_cache_add(tb);
    P.sig.clear();  // rebuild with the winner below
  }
  if (P.sig != sig || P.item_walkers != tune->item_walkers ||
      P.flush_every != tune->flush_every || P.tile_mode != tune->tile_mode || !P.exec) {
    if ((st = build_batch(c, P, n, flames, jobs, tune->item_walkers, tune->flush_every,
                          tune->tile_mode)))
      return st;
  }
  if ((st = upload_batch(c, P, flames, jobs, stops, n_stops, dev_out))) return st;
  if ((st = run_batch_graph(c, P, &gpu_ms))) return st;
  uint64_t total = 0;
  for (int i = 0; i < n; i++) {
    if (!dev_out) {
      hipError_t e = hipMemcpyAsync(rgba_out[i],
                                    reinterpret_cast<uint8_t*>(P.rgba.p) + P.pix[i].pix_off * 4,
                                    P.pix[i].npix * 4, hipMemcpyDeviceToHost, c->stream);
      if (e != hipSuccess)
        return fail(c, FFM_ERR_PIXEL_READ_FAILED, "pixel readback: %s", hipGetErrorString(e));
    }
    total += P.samples[i];
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (stats) {
    stats->total_samples = total;
    stats->dispatches = (uint64_t)P.n_waves;
    stats->iterate_ms = gpu_ms;
    stats->walker_kernel_ms = 0.0;
    stats->walker_launches = (uint64_t)P.n_waves;
    stats->tile_accum_ms = 0.0;
    stats->render_time_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return FFM_OK;
}

// The reference's gallery behaviour: the same flames rendered one after
// another (wasm/render/src/lib.rs:45-53), each a full ffm_render (or
// ffm_render_to_device into dev_out[i]).
int batch_serial_impl(ffm_ctx* c, int n, const ffm_flame_desc* flames, const float* const* stops,
                      const int* n_stops, const ffm_job* jobs, uint8_t* const* rgba_out,
                      uint8_t* const* dev_out, ffm_render_stats* stats) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  int st = check_batch(c, n, flames, stops, n_stops, jobs, dev_out ? dev_out : rgba_out);
  if (st) return st;
  auto t0 = std::chrono::steady_clock::now();
  uint64_t total = 0;
  double it_ms = 0.0;
  for (int i = 0; i < n; i++) {
    if ((st = ffm_flame_set(c, &flames[i]))) return st;
    if ((st = ffm_palette_set(c, stops[i], n_stops[i], BATCH_PALETTE))) return st;
    ffm_job j = jobs[i];
    j.target_samples = 0;  // one dispatch, as the batch renders it
    j.persistent_orbits = 0;
    j.splat_mode = FFM_SPLAT_U32_LDS;
    ffm_render_stats one{};
    if ((st = dev_out ? ffm_render_to_device(c, &j, dev_out[i], &one)
                      : ffm_render(c, &j, rgba_out[i], &one)))
      return st;
    total += one.total_samples;
    it_ms += one.iterate_ms;
  }
  if (stats) {
    stats->total_samples = total;
    stats->dispatches = (uint64_t)n;
    stats->iterate_ms = it_ms;
    stats->walker_kernel_ms = 0.0;
    stats->walker_launches = 0;
    stats->tile_accum_ms = 0.0;
    stats->render_time_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return FFM_OK;
}

}  // namespace

// Many flames in one persistent launch per wave, the whole sequence a
// hipGraph. The first call for a batch layout autotunes walkers per work
// item (walkers per lane = item / 512) and the staging flush cadence on
// the batch itself, then the tile shape with the winner of those, keeping
// the fastest; the graph is reused while the layout (sizes, walker counts,
// transform counts) stays the same.
int ffm_batch_render(ffm_ctx* c, int n, const ffm_flame_desc* flames, const float* const* stops,
                     const int* n_stops, const ffm_job* jobs, uint8_t* const* rgba_out,
                     ffm_render_stats* stats) {
  return batch_render_impl(c, n, flames, stops, n_stops, jobs, rgba_out, nullptr, stats);
}

int ffm_batch_render_to_device(ffm_ctx* c, int n, const ffm_flame_desc* flames,
                               const float* const* stops, const int* n_stops, const ffm_job* jobs,
                               uint8_t* const* rgba_dev, ffm_render_stats* stats) {
  return batch_render_impl(c, n, flames, stops, n_stops, jobs, nullptr, rgba_dev, stats);
}

int ffm_batch_render_serial(ffm_ctx* c, int n, const ffm_flame_desc* flames,
                            const float* const* stops, const int* n_stops, const ffm_job* jobs,
                            uint8_t* const* rgba_out, ffm_render_stats* stats) {
  return batch_serial_impl(c, n, flames, stops, n_stops, jobs, rgba_out, nullptr, stats);
}

int ffm_batch_render_serial_to_device(ffm_ctx* c, int n, const ffm_flame_desc* flames,
                                      const float* const* stops, const int* n_stops,
                                      const ffm_job* jobs, uint8_t* const* rgba_dev,
                                      ffm_render_stats* stats) {
  return batch_serial_impl(c, n, flames, stops, n_stops, jobs, nullptr, rgba_dev, stats);
}

}  // extern "C"

void free_batch(ffm_batch_state* b) {
  if (!b) return;
  b->plan.release();
  delete b;
}

namespace {
ffm_batch_state* batch_state(ffm_ctx* c) {
  if (!c->batch) c->batch = new ffm_batch_state();
  return c->batch;
}
}  // namespace

// ---- differentiable soft splat

namespace {

// Parameters of a fittable flame, or a negative status with the reason.
int soft_params(ffm_ctx* c, const ffm_flame_desc* f) {
  if (!f) return fail(c, FFM_ERR_INVALID_ARGUMENT, "flame is NULL");
  if (f->n_finals != 0)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "soft-splat flames have no final transforms");
  if (f->has_xaos)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "soft-splat flames have no xaos");
  if (f->n_xforms < 1 || f->n_xforms > FFM_SOFT_MAX_XFORMS)
    return fail(c, FFM_ERR_TOO_LARGE, "%d transforms: a soft-splat flame has 1..%d", f->n_xforms,
                FFM_SOFT_MAX_XFORMS);
  if (f->n_vars < 1 || f->n_vars > FFM_SOFT_MAX_VARS)
    return fail(c, FFM_ERR_TOO_LARGE, "%d variations: a soft-splat flame has 1..%d", f->n_vars,
                FFM_SOFT_MAX_VARS);
  for (int k = 0; k < f->n_vars; k++) {
    if (!supported_var(f->var_id[k]))
      return fail(c, FFM_ERR_INVALID_ARGUMENT, "variation id %d is not supported", f->var_id[k]);
    if (f->var_id[k] == FFM_VAR_WAVES)
      return fail(c, FFM_ERR_INVALID_ARGUMENT,
                  "waves reads the affine coefficients itself and has no soft-splat adjoint");
    if (f->var_id[k] == FFM_VAR_BLOB || f->var_id[k] == FFM_VAR_NOISE ||
        f->var_id[k] == FFM_VAR_BLUR || f->var_id[k] == FFM_VAR_GAUSSIAN_BLUR ||
        f->var_id[k] == FFM_VAR_PRE_BLUR)
      return fail(c, FFM_ERR_INVALID_ARGUMENT, "variation id %d has no soft-splat adjoint",
                  f->var_id[k]);
  }
  return f->n_xforms * (6 + f->n_vars);
}

int check_soft_job(ffm_ctx* c, const ffm_soft_job* j) {
  int st = check_view(c, &j->view);
  if (st) return st;
  if (j->walkers < 1 || j->walkers > (1ull << 32))
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "walkers must be 1..2^32 (got %llu)",
                (unsigned long long)j->walkers);
  if (j->ipt < 1) return fail(c, FFM_ERR_INVALID_ARGUMENT, "ipt must be at least 1");
  if (j->window < 1 || j->window > FFM_SOFT_MAX_WINDOW)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "window must be 1..%d (got %u)", FFM_SOFT_MAX_WINDOW,
                j->window);
  if (!std::isfinite(j->density_scale))
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "density_scale is not finite");
  return FFM_OK;
}

SoftArgs soft_args(ffm_ctx* c) {
  SoftArgs a{};
  a.fits = reinterpret_cast<const SoftFit*>(c->s_fits_d.p);
  a.n_fits = (int)c->s_fits.size();
  a.tables = reinterpret_cast<const float*>(c->s_tables.p);
  a.ckpt = reinterpret_cast<unsigned char*>(c->s_ckpt.p);
  a.partials = reinterpret_cast<double*>(c->s_part.p);
  a.total_pixels = c->s_pixels;
  return a;
}

}  // namespace

extern "C" {

int ffm_softsplat_param_count(const ffm_flame_desc* f) { return soft_params(nullptr, f); }

int ffm_softsplat_pack(const ffm_flame_desc* f, float* theta) {
  int P = soft_params(nullptr, f);
  if (P < 0) return P;
  if (!theta) return FFM_ERR_INVALID_ARGUMENT;
  const int n = f->n_xforms, v = f->n_vars;
  for (int i = 0; i < n; i++) {
    const ffm_xform& x = f->xforms[i];
    const float aff[6] = {x.a, x.b, x.c, x.d, x.e, x.f};
    for (int q = 0; q < 6; q++) theta[6 * i + q] = aff[q];
    for (int k = 0; k < v; k++) theta[6 * n + v * i + k] = x.var_weight[k];
  }
  return FFM_OK;
}

int ffm_softsplat_set(ffm_ctx* c, int n, const ffm_flame_desc* flames, const ffm_soft_job* jobs) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  if (n < 1 || !flames || !jobs) return fail(c, FFM_ERR_INVALID_ARGUMENT, "no fits given");
  if (n > 65535) return fail(c, FFM_ERR_TOO_LARGE, "%d fits: at most 65535 in one set", n);
  (void)hipSetDevice(c->device);
  const int B = soft_block();
  std::vector<SoftFit> fits((size_t)n);
  std::vector<float> tables;
  uint64_t theta = 0, pix = 0, ck = 0, part = 0, blocks = 0;
  size_t fwd_lds = 0, adj32 = 0, adj64 = 0;
  for (int i = 0; i < n; i++) {
    int P = soft_params(c, &flames[i]);
    if (P < 0) return fail(c, P, "fit %d: %s", i, c->err.c_str());
    int st = check_soft_job(c, &jobs[i]);
    if (st) return fail(c, st, "fit %d: %s", i, c->err.c_str());
    ffm_flame_desc f = flames[i];
    prepare_flame(&f);
    std::vector<float> t = pack_flame(&f);
    const ffm_soft_job& j = jobs[i];
    SoftFit& F = fits[(size_t)i];
    F.table_off = (uint32_t)tables.size();
    tables.insert(tables.end(), t.begin(), t.end());
    tables.resize((tables.size() + 3) & ~(size_t)3, 0.0f);
    F.n_xforms = f.n_xforms;
    F.n_vars = f.n_vars;
    F.n_params = P;
    F.view = make_view(j.view);
    F.walkers = j.walkers;
    F.ipt = j.ipt;
    F.burn_in = j.burn_in;
    F.seed = j.seed;
    F.window = j.window;
    F.truncate = j.truncate ? 1 : 0;
    F.f64 = j.walk_f64 ? 1 : 0;
    F.density_scale = j.density_scale;
    F.theta_off = theta;
    F.pix_off = pix;
    F.ckpt_off = ck;
    F.part_off = part;
    F.block0 = (uint32_t)blocks;
    F.n_blocks = (uint32_t)((j.walkers + B - 1) / B);
    theta += (uint64_t)P;
    pix += (uint64_t)j.view.width * (uint64_t)j.view.height;
    if (!F.truncate) ck += (soft_ckpt_bytes(j.ipt, j.window, j.walkers, F.f64) + 255) & ~255ull;
    part += (uint64_t)P * F.n_blocks;
    blocks += F.n_blocks;
    if (blocks > 0x7fffffffull)
      return fail(c, FFM_ERR_TOO_LARGE, "the fits need %llu blocks (at most 2^31 - 1)",
                  (unsigned long long)blocks);
    fwd_lds = std::max(fwd_lds, soft_forward_lds_bytes(f.n_xforms));
    size_t& adj = F.f64 ? adj64 : adj32;
    adj = std::max(adj, soft_adjoint_lds_bytes(f.n_xforms, P, (int)j.window, F.f64));
  }
  if (std::max(adj32, adj64) > (size_t)LDS_DYNAMIC_MAX)
    return fail(c, FFM_ERR_TOO_LARGE, "adjoint needs %zu bytes of LDS (at most %d)",
                std::max(adj32, adj64), LDS_DYNAMIC_MAX);
  int st;
  if ((st = ensure(c, c->s_fits_d, fits.size() * sizeof(SoftFit), "soft-splat fits")) ||
      (st = ensure(c, c->s_tables, tables.size() * 4, "soft-splat tables")))
    return st;
  HIPCHECK(c, hipMemcpyAsync(c->s_fits_d.p, fits.data(), fits.size() * sizeof(SoftFit),
                             hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->s_tables.p, tables.data(), tables.size() * 4,
                             hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->s_fits = std::move(fits);
  c->s_theta = theta;
  c->s_pixels = pix;
  c->s_ckpt_bytes = ck;
  c->s_part_n = part;
  c->s_blocks = (uint32_t)blocks;
  c->s_fwd_lds = fwd_lds;
  c->s_adj_lds32 = adj32;
  c->s_adj_lds64 = adj64;
  c->s_any32 = c->s_any64 = false;
  for (const SoftFit& F : c->s_fits) (F.f64 ? c->s_any64 : c->s_any32) = true;
  return FFM_OK;
}

int ffm_softsplat_sizes(const ffm_ctx* c, uint64_t* theta_floats, uint64_t* density_floats) {
  if (!c || !theta_floats || !density_floats) return FFM_ERR_INVALID_ARGUMENT;
  *theta_floats = c->s_theta;
  *density_floats = c->s_pixels;
  return c->s_fits.empty() ? FFM_ERR_NO_FLAME : FFM_OK;
}

int ffm_softsplat_forward(ffm_ctx* c, const float* theta, float* density, void* stream) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  if (c->s_fits.empty()) return fail(c, FFM_ERR_NO_FLAME, "no soft-splat fits set");
  if (!theta || !density) return fail(c, FFM_ERR_INVALID_ARGUMENT, "theta or density is NULL");
  (void)hipSetDevice(c->device);
  hipStream_t s = pick(c, stream);
  int st = ensure(c, c->s_xcd, (size_t)SOFT_XCD_COPIES * c->s_pixels * 4, "XCD density copies");
  if (st) return st;
  SoftArgs a = soft_args(c);
  a.theta = theta;
  a.density = density;
  a.dens_xcd = reinterpret_cast<float*>(c->s_xcd.p);
  HIPCHECK(c, launch_soft_forward(a, c->s_blocks, c->s_fwd_lds, c->s_any32, c->s_any64, s));
  return FFM_OK;
}

int ffm_softsplat_adjoint(ffm_ctx* c, const float* theta, const float* grad_density,
                          float* grad_theta, void* stream) {
  if (!c) return FFM_ERR_INVALID_ARGUMENT;
  if (c->s_fits.empty()) return fail(c, FFM_ERR_NO_FLAME, "no soft-splat fits set");
  if (!theta || !grad_density || !grad_theta)
    return fail(c, FFM_ERR_INVALID_ARGUMENT, "theta, grad_density or grad_theta is NULL");
  (void)hipSetDevice(c->device);
  hipStream_t s = pick(c, stream);
  int st;
  if ((st = ensure(c, c->s_ckpt, std::max<uint64_t>(c->s_ckpt_bytes, 256),
                   "adjoint checkpoints")) ||
      (st = ensure(c, c->s_part, c->s_part_n * 8, "adjoint partial sums")))
    return st;
  SoftArgs a = soft_args(c);
  a.theta = theta;
  a.grad_density = grad_density;
  a.grad_theta = grad_theta;
  HIPCHECK(c, launch_soft_adjoint(a, c->s_blocks, c->s_any32 ? c->s_adj_lds32 : 0,
                                  c->s_any64 ? c->s_adj_lds64 : 0, s));
  return FFM_OK;
}

}  // extern "C"
