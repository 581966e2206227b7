// The library's fixed device scratch for callers that cannot allocate
// (code under graph capture, serving loops, harnesses that forbid hipMalloc):
// one file-scope device array per GPU, resolved to its device address once
// per device at set-up (zav1_init).  Host code never passes the array's name
// to a kernel -- on the host it is only a stand-in for the symbol -- but the
// address hipGetSymbolAddress returned for the current device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "zav1_kernels.h"

namespace zav1 {

constexpr size_t kScratchBytes = 32ull << 20;  // 32 MiB per GPU, held until the process ends
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

// With `dev` current: resolve the array's address on that device (once).
hipError_t noalloc_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[dev]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_ptr[dev] = p;
  return hipSuccess;
}

// The device address of dev's scratch, or null before noalloc_setup(dev).
void* noalloc_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[dev];
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

}  // namespace zav1
