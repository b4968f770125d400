// The fixed scratch array of the allocation-free batch render and the lookup
// of its per-device address.  In host code `scratch` names a host-side
// stand-in, not device memory: every use on the host goes through the
// address hipGetSymbolAddress returns for the current device.
#include "noalloc.h"

namespace eye {
namespace {

__device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

constexpr int kMaxDevices = 64;
uint32_t* g_scratch[kMaxDevices] = {};

}  // namespace

bool no_alloc_scratch_init(int device, std::string* err) {
  if (device < 0 || device >= kMaxDevices) {
    if (err) *err = "no-alloc scratch: device index out of range";
    return false;
  }
  if (g_scratch[device]) return true;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    if (err) *err = std::string("no-alloc scratch: hipGetSymbolAddress: ") + hipGetErrorString(e);
    return false;
  }
  g_scratch[device] = static_cast<uint32_t*>(p);
  return true;
}

uint32_t* no_alloc_scratch(int device) {
  return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

}  // namespace eye
