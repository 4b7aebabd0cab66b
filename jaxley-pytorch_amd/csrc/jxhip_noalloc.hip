// jxhip_fused_forward_noalloc_f64: the float64 fused forward for callers that cannot
// allocate device memory (code under graph capture, serving loops, harnesses that forbid
// hipMalloc). The streaming kernel's scratch (the axial part of every item's residual,
// `work_ax`) lives in a file-scope device array instead of a caller buffer; everything
// else is the caller's, exactly as for jxhip_fused_forward_f64.
#include <atomic>
#include <mutex>

#include "jxhip_fused.h"

namespace {

constexpr int64_t kScratchBytes = int64_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

// The scratch array: each GPU the library runs on holds its own copy for the lifetime of
// the process. In host code `scratch` names a host-side stand-in, never the device
// memory: only the address from hipGetSymbolAddress (below) is passed to kernels.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::atomic<void*> g_scratch[kMaxDevices];
std::mutex g_scratch_mutex;

// The device address of `scratch` on the current device, looked up once per device
// (with that device current) on its first call.
hipError_t scratch_ptr(void** out) {
  int dev = 0;
  if (hipError_t e = hipGetDevice(&dev)) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  void* p = g_scratch[dev].load(std::memory_order_acquire);
  if (p == nullptr) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    p = g_scratch[dev].load(std::memory_order_relaxed);
    if (p == nullptr) {
      if (hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) return e;
      g_scratch[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return hipSuccess;
}

}  // namespace

extern "C" {

int jxhip_fused_forward_noalloc_f64(const jxhip_fused_args* args, int64_t state_elems, void* stream) {
  if (args == nullptr || state_elems < 0) return hipErrorInvalidValue;
  if (args->work_ax != nullptr) return hipErrorInvalidValue;  // the scratch is the library's
  // The scratch holds one float64 per state element (program layout of `v`).
  if (state_elems > kScratchBytes / static_cast<int64_t>(sizeof(double))) return hipErrorInvalidValue;
  void* p = nullptr;
  if (hipError_t e = scratch_ptr(&p)) return e;
  jxhip_fused_args a = *args;
  a.work_ax = p;
  return jxhip::launch_forward<double, double>(&a, static_cast<hipStream_t>(stream));
}

}  // extern "C"
