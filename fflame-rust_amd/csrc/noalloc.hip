// noalloc.hip -- the device-resident scratch of ffm_iterate_noalloc.
//
// Callers that cannot allocate (code under HIP graph capture, serving loops
// that forbid hipMalloc) get a variant of ffm_iterate whose LDS tile splat
// keeps its per-bin tables and staged records in this file-scope array
// instead of context buffers. Each GPU holds its own copy of the array,
// from the library's load until the process ends; every call on that GPU
// shares it, so calls that overlap on one GPU race (one at a time per
// device, as the context's own calls are).
//
// In host code `scratch` is only a stand-in for the device symbol: its
// device address comes from hipGetSymbolAddress, looked up once per device
// with that device current, and only that address is handed to kernels.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ffm_internal.h"

namespace ffm {

static __device__ __align__(256) uint8_t scratch[FFM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES];
hipError_t g_status[MAX_DEVICES];
bool g_done[MAX_DEVICES];
}  // namespace

// Look up this device's copy of the scratch (once per device; `device` must
// be current). Returns the lookup's status, also on later calls.
hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_done[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_status[device] = e;
    g_ptr[device] = e == hipSuccess ? p : nullptr;
    if (e != hipSuccess) (void)hipGetLastError();
    g_done[device] = true;
  }
  return g_status[device];
}

// The device address of `device`'s scratch, or null before a successful
// noalloc_scratch_init on it.
void* noalloc_scratch(int device) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_mu);
  return g_done[device] ? g_ptr[device] : nullptr;
}

}  // namespace ffm
