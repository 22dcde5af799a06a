// The static device scratch array of the allocation-free entry point.
//
// Built by hipcc (host code plus this one device variable) and linked into
// libgrafeo_hip.so; the kernels that use the array are in the embedded code
// objects and receive its device address as an argument. In host code
// `scratch` names a host-side stand-in, never the device memory: host code
// uses only the address returned by noalloc_scratch_address().
#include "noalloc.hpp"

static __device__ __align__(256) unsigned char scratch[gc::kNoAllocScratchBytes];

namespace gc {

hipError_t noalloc_scratch_address(void** ptr) {
  return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

}  // namespace gc
