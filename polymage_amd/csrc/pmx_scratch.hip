// Static device scratch of libpmx_rt, for the allocation-free entry points
// (pipeline_<name>_dev_noalloc in every generated pipeline library).
//
// `scratch` is a file-scope device array: each GPU that runs this library
// holds its own copy, at its own address, from the load of the library's
// code object on that device until the process ends.  Host code never uses
// the array's name as a device address (on the host it names a stand-in
// object): pmx_scratch() looks the device address up once per device with
// hipGetSymbolAddress, with that device current, and hands out that
// address.  Every allocation-free call on a GPU shares that GPU's array and
// finds it holding what the previous call left; calls that overlap on one
// GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pmx_rt.h"

static __device__ __align__(256) unsigned char scratch[PMX_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];
}  // namespace

extern "C" {

int pmx_scratch(void** ptr, size_t* bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    pmx_set_error("pmx_scratch: hipGetDevice failed");
    return -(int)e;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    pmx_set_error("pmx_scratch: device index out of range");
    return -22;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      pmx_set_error(e != hipSuccess ? hipGetErrorString(e)
                                    : "pmx_scratch: hipGetSymbolAddress gave no address");
      return e != hipSuccess ? -(int)e : -14;
    }
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  if (bytes) *bytes = (size_t)PMX_SCRATCH_BYTES;
  return 0;
}

int pmx_scratch_peek(void** ptr) {
  int dev = 0;
  *ptr = nullptr;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 0;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  *ptr = g_scratch_ptr[dev];
  return 0;
}

}  // extern "C"
