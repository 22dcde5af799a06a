// Small data-movement kernels shared by the algorithm drivers: fills,
// permutation gathers, the forward-neighbour split of a CSR, and an
// order-preserving compaction of flagged CSR entries into (u, v) pairs.
#include "device.hpp"

using namespace gcd;

extern "C" __global__ __launch_bounds__(256) void fill_u32(uint32_t* __restrict__ p, uint64_t n,
                                                           uint32_t value) {
  for (uint64_t i = global_thread(); i < n; i += grid_threads()) p[i] = value;
}

// dst[i] = src[index[i]] for 32-bit words (rank, distance or id vectors).
// When index and dst are 16-byte aligned, a lane takes four entries: one
// 16-byte index load, four gathers in flight, one 16-byte store.
extern "C" __global__ __launch_bounds__(256) void gather_u32(const uint32_t* __restrict__ src,
                                                             const uint32_t* __restrict__ index,
                                                             uint32_t* __restrict__ dst,
                                                             uint64_t n) {
  uint64_t done = 0;
  if (((reinterpret_cast<uintptr_t>(index) | reinterpret_cast<uintptr_t>(dst)) & 15u) == 0) {
    const uint64_t quads = n / 4;
    const uint4* index4 = reinterpret_cast<const uint4*>(index);
    uint4* dst4 = reinterpret_cast<uint4*>(dst);
    for (uint64_t q = global_thread(); q < quads; q += grid_threads()) {
      const uint4 ix = index4[q];
      dst4[q] = make_uint4(src[ix.x], src[ix.y], src[ix.z], src[ix.w]);
    }
    done = 4 * quads;
  }
  for (uint64_t i = done + global_thread(); i < n; i += grid_threads()) dst[i] = src[index[i]];
}

// fwd[u] = first position in row u whose target is > u (rows are sorted).
extern "C" __global__ __launch_bounds__(256) void csr_forward_start(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t n,
    uint32_t* __restrict__ fwd) {
  for (uint64_t u = global_thread(); u < n; u += grid_threads()) {
    uint32_t row = static_cast<uint32_t>(u);
    fwd[u] = lower_bound_u32(targets, offsets[u], offsets[u + 1], row + 1);
  }
}

// ---- order-preserving compaction of flagged CSR entries ----------------
//
// Three passes over `m` byte flags: per-block counts, one exclusive scan of
// the counts, then each block writes its flagged entries as (row, target)
// pairs at its scanned offset, in entry order (and, when `values` is given,
// values[entry] at the same position of `out_values`).

#define COMPACT_THREADS 256
#define COMPACT_PER_THREAD 16
#define COMPACT_PER_BLOCK (COMPACT_THREADS * COMPACT_PER_THREAD)

extern "C" __global__ __launch_bounds__(COMPACT_THREADS) void compact_count(
    const uint8_t* __restrict__ flags, uint64_t m, uint32_t* __restrict__ block_counts) {
  __shared__ uint32_t s_wave[COMPACT_THREADS / GC_WAVE];
  uint64_t base = static_cast<uint64_t>(blockIdx.x) * COMPACT_PER_BLOCK +
                  static_cast<uint64_t>(threadIdx.x) * COMPACT_PER_THREAD;
  uint32_t c = 0;
  for (uint32_t k = 0; k < COMPACT_PER_THREAD; ++k) {
    uint64_t i = base + k;
    if (i < m) c += flags[i] != 0;
  }
  c = wave_sum(c);
  if (lane_id() == 0) s_wave[threadIdx.x / GC_WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (int w = 0; w < COMPACT_THREADS / GC_WAVE; ++w) total += s_wave[w];
    block_counts[blockIdx.x] = total;
  }
}

// Exclusive scan of `count` values in place by one workgroup; *total gets the sum.
extern "C" __global__ __launch_bounds__(1024) void exclusive_scan_small(uint32_t* __restrict__ data,
                                                                        uint32_t count,
                                                                        uint32_t* __restrict__ total) {
  __shared__ uint32_t s_wave[1024 / GC_WAVE];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t chunk = 0; chunk < count; chunk += 1024) {
    uint32_t i = chunk + threadIdx.x;
    uint32_t v = i < count ? data[i] : 0;
    uint32_t incl = wave_inclusive_scan(v);
    uint32_t wave = threadIdx.x / GC_WAVE;
    if (lane_id() == GC_WAVE - 1) s_wave[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t run = 0;
      for (int w = 0; w < 1024 / GC_WAVE; ++w) {
        uint32_t t = s_wave[w];
        s_wave[w] = run;
        run += t;
      }
    }
    __syncthreads();
    uint32_t carry = s_carry;
    if (i < count) data[i] = carry + s_wave[wave] + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) s_carry = carry + s_wave[wave] + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = s_carry;
}

// Writes the flagged entries of block b at block_offsets[b] onwards as
// (source row, target) pairs.
extern "C" __global__ __launch_bounds__(COMPACT_THREADS) void compact_emit_pairs(
    const uint8_t* __restrict__ flags, uint64_t m, const uint32_t* __restrict__ block_offsets,
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t n,
    uint32_t* __restrict__ pairs, const uint32_t* __restrict__ values,
    uint32_t* __restrict__ out_values) {
  __shared__ uint32_t s_wave[COMPACT_THREADS / GC_WAVE];
  uint64_t base = static_cast<uint64_t>(blockIdx.x) * COMPACT_PER_BLOCK +
                  static_cast<uint64_t>(threadIdx.x) * COMPACT_PER_THREAD;
  uint32_t c = 0;
  for (uint32_t k = 0; k < COMPACT_PER_THREAD; ++k) {
    uint64_t i = base + k;
    if (i < m) c += flags[i] != 0;
  }
  uint32_t incl = wave_inclusive_scan(c);
  uint32_t wave = threadIdx.x / GC_WAVE;
  if (lane_id() == GC_WAVE - 1) s_wave[wave] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wave; ++w) before += s_wave[w];
  uint32_t pos = block_offsets[blockIdx.x] + before + incl - c;
  if (c == 0) return;
  uint32_t row = 0;
  bool have_row = false;
  for (uint32_t k = 0; k < COMPACT_PER_THREAD; ++k) {
    uint64_t i = base + k;
    if (i >= m) break;
    if (flags[i] == 0) continue;
    uint32_t e = static_cast<uint32_t>(i);
    if (!have_row) {
      row = row_of_entry(offsets, n, e);
      have_row = true;
    }
    while (offsets[row + 1] <= e) ++row;
    pairs[2 * static_cast<uint64_t>(pos)] = row;
    pairs[2 * static_cast<uint64_t>(pos) + 1] = targets[e];
    if (values != nullptr) out_values[pos] = values[e];
    ++pos;
  }
}

// ---- exclusive prefix sum of u32 values, in place ---------------------------
//
// scan_block_sums: per-block totals of SCAN_PER_BLOCK consecutive values;
// exclusive_scan_small over the totals; scan_apply: each block scans its own
// values and adds its base. data[count] receives nothing; `total` gets the sum.

#define SCAN_THREADS 256
#define SCAN_PER_THREAD 16
#define SCAN_PER_BLOCK (SCAN_THREADS * SCAN_PER_THREAD)

extern "C" __global__ __launch_bounds__(SCAN_THREADS) void scan_block_sums(
    const uint32_t* __restrict__ data, uint64_t count, uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t s_wave[SCAN_THREADS / GC_WAVE];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * SCAN_PER_BLOCK +
                        static_cast<uint64_t>(threadIdx.x) * SCAN_PER_THREAD;
  uint32_t c = 0;
  for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
    const uint64_t i = base + k;
    if (i < count) c += data[i];
  }
  c = wave_sum(c);
  if (lane_id() == 0) s_wave[threadIdx.x / GC_WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < SCAN_THREADS / GC_WAVE; ++w) t += s_wave[w];
    block_sums[blockIdx.x] = t;
  }
}

extern "C" __global__ __launch_bounds__(SCAN_THREADS) void scan_apply(
    uint32_t* __restrict__ data, uint64_t count, const uint32_t* __restrict__ block_bases) {
  __shared__ uint32_t s_wave[SCAN_THREADS / GC_WAVE];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * SCAN_PER_BLOCK +
                        static_cast<uint64_t>(threadIdx.x) * SCAN_PER_THREAD;
  uint32_t v[SCAN_PER_THREAD];
  uint32_t c = 0;
#pragma unroll
  for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
    const uint64_t i = base + k;
    v[k] = i < count ? data[i] : 0u;
    c += v[k];
  }
  const uint32_t incl = wave_inclusive_scan(c);
  const uint32_t wave = threadIdx.x / GC_WAVE;
  if (lane_id() == GC_WAVE - 1) s_wave[wave] = incl;
  __syncthreads();
  uint32_t run = block_bases[blockIdx.x];
  for (uint32_t w = 0; w < wave; ++w) run += s_wave[w];
  run += incl - c;
#pragma unroll
  for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
    const uint64_t i = base + k;
    if (i < count) data[i] = run;
    run += v[k];
  }
}
