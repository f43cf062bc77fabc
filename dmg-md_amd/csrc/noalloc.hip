// noalloc.hip -- the fixed device workspace of dmgx_nep_compute_noalloc.
//
// Callers that cannot allocate device memory (tight serving loops, harnesses
// that forbid hipMalloc) evaluate forces with a workspace that lives in a
// file-scope device array: the code object reserves it on every GPU the
// library runs on, and it stays there until the process ends.  Host code
// never passes the array's name to a kernel (in host code it is a stand-in,
// not a device address): the device address of each GPU's copy is taken once
// per device with hipGetSymbolAddress, with that device current, and kept.
#include <hip/hip_runtime.h>

#include <mutex>

#include "noalloc.hpp"

namespace dmgx {

namespace {

static __device__ __align__(256) unsigned char scratch[kFixedWorkspaceBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

}  // namespace

hipError_t fixed_workspace(void** ptr, size_t* bytes)
{
  *ptr = nullptr;
  *bytes = 0;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[dev] = p;
  }
  *ptr = g_scratch[dev];
  *bytes = kFixedWorkspaceBytes;
  return hipSuccess;
}

}  // namespace dmgx
