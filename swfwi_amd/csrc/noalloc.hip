// The allocation-free forward modelling entry point's workspace: one
// file-scope device array per GPU, of a fixed maximum size, held until the
// process ends.  For callers that cannot call hipMalloc (code under graph
// capture, tight serving loops, harnesses that forbid it).
//
// Host code never passes the array's name to a kernel or a copy: the name is
// a host-side stand-in, not a device address.  Each device's address comes
// from hipGetSymbolAddress, taken once per device with that device current,
// and only that address is used.  Every call on one GPU shares that GPU's
// array, and the array keeps the last call's contents; the forward run
// clears its two wavefields itself and writes every recorded sample, so
// nothing else needs to start at zero.  Calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "noalloc.h"

namespace swfwi {

namespace {

static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float* g_ptr[kMaxDevices] = {};

}  // namespace

float* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) fail(kErrArg, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    int cur = -1;
    SWFWI_HIP(hipGetDevice(&cur));
    if (cur != device) SWFWI_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) SWFWI_HIP(hipSetDevice(cur));
    hip_check(e, "hipGetSymbolAddress(scratch)");
    if (!p) fail(kErrHip, "hipGetSymbolAddress(scratch) returned a null address");
    g_ptr[device] = static_cast<float*>(p);
  }
  return g_ptr[device];
}

}  // namespace swfwi
