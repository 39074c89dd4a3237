// noalloc.hip -- the device array behind the allocation-free entry point.
//
// `scratch` is a file-scope device array: the runtime reserves it on each GPU
// when the library's code object is loaded there, and it stays until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy (in host code that name is only a stand-in); it passes the device
// address that hipGetSymbolAddress returns for the current device, kept per
// device below.
#include <hip/hip_runtime.h>

#include <mutex>

#include "noalloc.hpp"

namespace stqamd {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

int noalloc_scratch(int device, void **ptr) {
  if (!ptr || device < 0 || device >= kMaxDevices) return STQ_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return STQ_ERR_HIP;
    }
    g_scratch[device] = p;
  }
  *ptr = g_scratch[device];
  return STQ_OK;
}

}  // namespace stqamd
