// The allocation-free moments readback (C-ABI lbm_moments_noalloc), for
// callers that cannot allocate device memory: code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc.  It computes what
// lbm_moments computes, with the same kernels, but stages rho and u in a
// fixed device array of this file instead of a buffer the library
// allocates.
//
// The array: `scratch`, kScratchBytes (64 MiB) at file scope, one copy on
// each GPU the library runs on, held until the process ends.  In host code
// its name is only a host-side stand-in; the device address comes from
// hipGetSymbolAddress with the device current, once per device (at the
// entry point's first call on that device), and only that address is
// passed to kernels and copies.  Every call on one GPU shares the array
// and it keeps the last call's contents (each call writes every element it
// reads back, so nothing needs clearing).  Calls that overlapped on one
// GPU would race on it, so a per-device lock serialises them: a call holds
// it from its kernel launch until its copies to the host have completed.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "lbm_internal.h"

namespace lbm {
namespace {

constexpr size_t kScratchBytes = 64ull << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_device_lock[kMaxDevices];  // one user of a device's array at a time
void* g_scratch[kMaxDevices] = {};      // its device address, once looked up

// Device address of `scratch` on `device`, looked up with that device
// current on the first call for it and kept.  The caller holds
// g_device_lock[device].
int scratch_address(int device, void** out) {
  if (!g_scratch[device]) {
    int current = 0;
    LBM_HIP(hipGetDevice(&current));
    if (current != device) LBM_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (current != device) (void)hipSetDevice(current);
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(LBM_ERR_HIP, "hipGetSymbolAddress(scratch) gave no address");
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return LBM_OK;
}

}  // namespace

namespace api {

int lbm_moments_noalloc(lbm_ctx* c, float* rho, float* u) {
  if (!c) return fail(LBM_ERR_INVALID, "null context");
  const size_t N = c->grid.n;
  const size_t need = (rho ? 4 * N : 0) + (u ? 12 * N : 0);
  if (need > kScratchBytes)
    return fail(LBM_ERR_INVALID,
                "lbm_moments_noalloc: " + std::to_string(N) + " nodes need " +
                    std::to_string(need) + " bytes of scratch; the maximum is " +
                    std::to_string(kScratchBytes) +
                    " (4,194,304 nodes for rho and u, 5,592,405 for u alone)");
  if (need == 0) return LBM_OK;
  if (c->device < 0 || c->device >= kMaxDevices)
    return fail(LBM_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_device_lock[c->device]);
  void* base = nullptr;
  if (int rc = scratch_address(c->device, &base)) return rc;
  // launches the moments kernels into the array, copies to the host and
  // synchronises the context's stream before the lock is released
  return read_moments(c, base, rho, u);
}

}  // namespace api
}  // namespace lbm
