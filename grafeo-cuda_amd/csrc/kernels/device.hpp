// Device helpers shared by the gfx950 kernel files (wave64 CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GC_WAVE 64

namespace gcd {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (GC_WAVE - 1); }

__device__ __forceinline__ uint64_t global_thread() {
  return static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
}

__device__ __forceinline__ uint64_t grid_threads() {
  return static_cast<uint64_t>(gridDim.x) * blockDim.x;
}

// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = GC_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < GC_WAVE; o <<= 1) {
    uint32_t t = __shfl_up(v, o);
    if (lane >= static_cast<uint32_t>(o)) v += t;
  }
  return v;
}

// For a wave-wide non-decreasing `incl` (one value per lane), the number of
// lanes whose value is <= t, i.e. the first lane j with incl_j > t. Every
// lane must call it (it shuffles).
__device__ __forceinline__ uint32_t wave_upper_bound(uint32_t incl, uint32_t t) {
  uint32_t l = 0;
#pragma unroll
  for (uint32_t step = 32; step > 0; step >>= 1) {
    uint32_t probe = __shfl(incl, static_cast<int>(l + step - 1));
    if (probe <= t) l += step;
  }
  return l;
}

// Makes this wave's LDS writes visible to its own other lanes; waves of a
// workgroup that work independently need no workgroup barrier.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Streaming load that should not displace reused data from the caches.
__device__ __forceinline__ uint32_t load_stream(const uint32_t* p) {
  return __builtin_nontemporal_load(p);
}

// First index in [lo, hi) of the sorted array `a` whose value is >= x.
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi,
                                                    uint32_t x) {
  while (lo < hi) {
    uint32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Source row of CSR entry e: the largest u with offsets[u] <= e.
__device__ __forceinline__ uint32_t row_of_entry(const uint32_t* offsets, uint32_t n, uint32_t e) {
  uint32_t lo = 0, hi = n - 1;  // e < offsets[n], so the answer is in [0, n)
  while (lo < hi) {
    uint32_t mid = lo + ((hi - lo + 1) >> 1);
    if (offsets[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

}  // namespace gcd
