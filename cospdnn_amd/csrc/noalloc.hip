// The fixed device array that cospdnn_run_noalloc works in, and its per-device address.
//
// `scratch` is a file-scope device array: every GPU the library runs on holds its own copy
// from module load until the process ends.  In host code the array's name is only a host-side
// stand-in, not a device address, so it is never passed to a kernel or a copy: the device
// address of each GPU's copy is looked up once per device with hipGetSymbolAddress (with that
// device current) and kept; everything else uses that pointer.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine_internal.hpp"

namespace cospdnn_detail {

namespace {

constexpr size_t kNoallocScratchBytes = static_cast<size_t>(COSPDNN_NOALLOC_SCRATCH_BYTES);

__device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

int noalloc_scratch(int device, void** ptr, size_t* bytes) {
  if (device < 0 || device >= kMaxDevices) return fail(COSPDNN_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != device) HIP_TRY(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || p == nullptr)
      return fail(COSPDNN_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  *bytes = kNoallocScratchBytes;
  return COSPDNN_OK;
}

std::mutex& noalloc_mutex(int device) {
  static std::mutex mu[kMaxDevices];
  return mu[device >= 0 && device < kMaxDevices ? device : 0];
}

}  // namespace cospdnn_detail
