// Static device scratch of the allocation-free entry point
// (gc_triangle_count_noalloc).
//
// The array itself lives in noalloc.hip, the one host source built by hipcc,
// because only a HIP translation unit can hold a __device__ variable whose
// address the runtime resolves (hipGetSymbolAddress). Every GPU the library
// opens holds its own copy of the array, at its own address, until the
// process ends.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace gc {

// Size of the static scratch array on each device.
constexpr size_t kNoAllocScratchBytes = 4096;

// Device address of the scratch array on the current device (the device must
// be current: each device has its own copy). Returns the hipGetSymbolAddress
// status.
hipError_t noalloc_scratch_address(void** ptr);

}  // namespace gc
