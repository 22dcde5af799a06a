// Device CSR construction: a stable LSD radix sort of 64-bit keys and the
// kernels that turn sorted (row, target) keys into a CSR.
//
// Radix sort, one pass per 8-bit digit (the caller runs ceil(bits / 8)
// passes over the significant bits only):
//   rs_histogram  each workgroup counts the digits of its tile of RS_TILE keys
//                 and writes them digit-major, hist[digit * blocks + block],
//                 so that one exclusive scan of the whole array gives every
//                 (digit, tile) pair its first output slot, tiles in order;
//   rs_scatter    each workgroup ranks its keys stably (in input order) and
//                 writes key and value to hist[digit * blocks + block] + rank.
// Stable ranking: the tile is walked in steps of 256 keys, one per lane, in
// input order. Lanes of a wave holding the same digit are found with 8
// ballots (one per digit bit); the rank is the number of such lanes below,
// plus the same digit's count in the lower waves of the step, plus the
// digit's running count over the earlier steps of the tile.
#include "device.hpp"

using namespace gcd;

#define RS_THREADS 256
#define RS_ITEMS 16
#define RS_TILE (RS_THREADS * RS_ITEMS)
#define RS_DIGITS 256
#define RS_WAVES (RS_THREADS / GC_WAVE)

extern "C" __global__ __launch_bounds__(RS_THREADS) void rs_histogram(
    const unsigned long long* __restrict__ keys, uint64_t n, uint32_t shift,
    uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_count[RS_DIGITS];
  s_count[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * RS_TILE;
  for (uint32_t k = 0; k < RS_ITEMS; ++k) {
    const uint64_t i = base + static_cast<uint64_t>(k) * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&s_count[(keys[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[static_cast<uint64_t>(threadIdx.x) * gridDim.x + blockIdx.x] = s_count[threadIdx.x];
}

// Lanes of the wave whose digit equals this lane's (active lanes only).
__device__ __forceinline__ unsigned long long same_digit_lanes(uint32_t digit, bool active) {
  unsigned long long peers = __ballot(active);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const unsigned long long set = __ballot(active && ((digit >> b) & 1u));
    peers &= ((digit >> b) & 1u) ? set : ~set;
  }
  return peers;
}

extern "C" __global__ __launch_bounds__(RS_THREADS) void rs_scatter(
    const unsigned long long* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint64_t n, uint32_t shift, const uint32_t* __restrict__ scanned,
    unsigned long long* __restrict__ keys_out, uint32_t* __restrict__ vals_out) {
  __shared__ uint32_t s_base[RS_DIGITS];               // output slot of the digit's next key
  __shared__ uint32_t s_wave_count[RS_WAVES][RS_DIGITS];
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / GC_WAVE;
  s_base[threadIdx.x] = scanned[static_cast<uint64_t>(threadIdx.x) * gridDim.x + blockIdx.x];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * RS_TILE;
  for (uint32_t k = 0; k < RS_ITEMS; ++k) {
    for (uint32_t w = 0; w < RS_WAVES; ++w) s_wave_count[w][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t i = base + static_cast<uint64_t>(k) * RS_THREADS + threadIdx.x;
    const bool active = i < n;
    unsigned long long key = 0;
    uint32_t val = 0, digit = 0;
    if (active) {
      key = keys_in[i];
      if (vals_in != nullptr) val = vals_in[i];
      digit = static_cast<uint32_t>((key >> shift) & 0xFFu);
    }
    const unsigned long long peers = same_digit_lanes(digit, active);
    const uint32_t below = lanes_below(peers);
    // the lowest lane of each digit group publishes the group's size
    if (active && below == 0) s_wave_count[wave][digit] = static_cast<uint32_t>(__popcll(peers));
    __syncthreads();
    if (active) {
      uint32_t rank = s_base[digit] + below;
      for (uint32_t w = 0; w < wave; ++w) rank += s_wave_count[w][digit];
      keys_out[rank] = key;
      if (vals_in != nullptr) vals_out[rank] = val;
    }
    __syncthreads();
    // advance each digit's base by its count in this step
    uint32_t step = 0;
    for (uint32_t w = 0; w < RS_WAVES; ++w) step += s_wave_count[w][threadIdx.x];
    s_base[threadIdx.x] += step;
    __syncthreads();
  }
}

// ---- CSR from edge pairs --------------------------------------------------------

// keys from (src, dst) pairs: src << id_bits | dst, and, when symmetric,
// dst << id_bits | src for every non-loop pair (slot 2i and 2i+1; a loop's
// second slot repeats its first, removed by the dedup). `bad` counts pairs
// with an endpoint >= n.
extern "C" __global__ __launch_bounds__(256) void csr_make_keys(
    const uint32_t* __restrict__ pairs, uint64_t num_pairs, uint32_t n, uint32_t id_bits,
    uint32_t symmetric, unsigned long long* __restrict__ keys, uint32_t* __restrict__ bad) {
  uint32_t local_bad = 0;
  for (uint64_t p = global_thread(); p < num_pairs; p += grid_threads()) {
    const uint32_t s = pairs[2 * p], d = pairs[2 * p + 1];
    if (s >= n || d >= n) ++local_bad;
    const unsigned long long fwd = (static_cast<unsigned long long>(s) << id_bits) | d;
    if (symmetric) {
      keys[2 * p] = fwd;
      keys[2 * p + 1] = (static_cast<unsigned long long>(d) << id_bits) | s;
    } else {
      keys[p] = fwd;
    }
  }
  local_bad = wave_sum(local_bad);
  if (lane_id() == 0 && local_bad) atomicAdd(bad, local_bad);
}

// ---- oriented CSR from edge pairs -------------------------------------------------
// (the steps of from_edge_pairs_oriented, crates/grafeo-cuda-core/src/csr_export.rs:299-350)

// Canonical undirected keys min << id_bits | max, one per pair (a loop keeps
// its key min == max and is dropped by csr_unique_edges); `bad` counts pairs
// with an endpoint >= n.
extern "C" __global__ __launch_bounds__(256) void csr_undirected_keys(
    const uint32_t* __restrict__ pairs, uint64_t num_pairs, uint32_t n, uint32_t id_bits,
    unsigned long long* __restrict__ keys, uint32_t* __restrict__ bad) {
  uint32_t local_bad = 0;
  for (uint64_t p = global_thread(); p < num_pairs; p += grid_threads()) {
    const uint32_t s = pairs[2 * p], d = pairs[2 * p + 1];
    if (s >= n || d >= n) ++local_bad;
    const unsigned long long lo = min(s, d), hi = max(s, d);
    keys[p] = (lo << id_bits) | hi;
  }
  local_bad = wave_sum(local_bad);
  if (lane_id() == 0 && local_bad) atomicAdd(bad, local_bad);
}

// flags[i] = 1 for the first copy of each sorted undirected key that is not a
// loop (its two ids differ).
extern "C" __global__ __launch_bounds__(256) void csr_unique_edges(
    const unsigned long long* __restrict__ keys, uint64_t count, uint32_t id_bits, uint8_t* __restrict__ flags) {
  const unsigned long long mask = (1ull << id_bits) - 1ull;
  for (uint64_t i = global_thread(); i < count; i += grid_threads()) {
    const unsigned long long k = keys[i];
    const bool first = i == 0 || k != keys[i - 1];
    flags[i] = (first && (k >> id_bits) != (k & mask)) ? 1 : 0;
  }
}

// degree[x] += the unique undirected edges at x (degree cleared by the
// caller). The keys are sorted, so the lower ids come in runs: each run adds
// its length with one atomic (a hub's run would otherwise serialize on one
// address); the higher ids add one each.
extern "C" __global__ __launch_bounds__(256) void csr_edge_degrees(
    const unsigned long long* __restrict__ keys, uint64_t count, uint32_t id_bits,
    uint32_t* __restrict__ degree) {
  const unsigned long long mask = (1ull << id_bits) - 1ull;
  const uint32_t lane = lane_id();
  const uint64_t waves = grid_threads() / GC_WAVE;
  for (uint64_t base = (global_thread() / GC_WAVE) * GC_WAVE; base < count; base += waves * GC_WAVE) {
    const uint64_t i = base + lane;
    const bool valid = i < count;
    const unsigned long long k = valid ? keys[i] : 0ull;
    const uint32_t a = static_cast<uint32_t>(k >> id_bits);
    const uint32_t b = static_cast<uint32_t>(k & mask);
    const uint32_t prev_a = __shfl_up(a, 1);
    // heads of the runs of equal lower ids within the wave
    const bool head = valid && (lane == 0 || prev_a != a);
    const unsigned long long heads = __ballot(head);
    if (head) {
      // the run ends at the next head or at the last valid lane
      const unsigned long long later = lane == GC_WAVE - 1 ? 0ull : heads >> (lane + 1);
      const uint64_t valid_lanes = count - base < GC_WAVE ? count - base : GC_WAVE;
      const uint32_t end = later != 0 ? lane + 1 + static_cast<uint32_t>(__ffsll(static_cast<long long>(later)) - 1)
                                      : static_cast<uint32_t>(valid_lanes);
      atomicAdd(&degree[a], end - lane);
    }
    if (valid) atomicAdd(&degree[b], 1u);
  }
}

// Each unique undirected edge (a, b), a < b, oriented from the endpoint of
// lower degree to the higher, ties from the lower id: a << id_bits | b or
// b << id_bits | a.
extern "C" __global__ __launch_bounds__(256) void csr_orient_keys(
    const unsigned long long* __restrict__ keys, uint64_t count, uint32_t id_bits,
    const uint32_t* __restrict__ degree, unsigned long long* __restrict__ out) {
  const unsigned long long mask = (1ull << id_bits) - 1ull;
  for (uint64_t i = global_thread(); i < count; i += grid_threads()) {
    const unsigned long long k = keys[i];
    const unsigned long long a = k >> id_bits, b = k & mask;
    const uint32_t da = degree[a], db = degree[b];
    out[i] = (da < db || (da == db && a < b)) ? k : ((b << id_bits) | a);
  }
}

// flags[i] = 1 when sorted key i differs from key i-1 (first copies of each key).
extern "C" __global__ __launch_bounds__(256) void csr_first_copies(
    const unsigned long long* __restrict__ keys, uint64_t count, uint8_t* __restrict__ flags) {
  for (uint64_t i = global_thread(); i < count; i += grid_threads())
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// Compacts the flagged (first-copy) sorted keys in order; the per-tile counts
// (tiles of 4096 entries) are scanned by the caller with compact_count and
// exclusive_scan_small of util.hip, which use the same tile.
extern "C" __global__ __launch_bounds__(256) void csr_compact_keys(
    const unsigned long long* __restrict__ keys, const uint8_t* __restrict__ flags, uint64_t count,
    const uint32_t* __restrict__ tile_offsets, unsigned long long* __restrict__ out) {
  __shared__ uint32_t s_wave[256 / GC_WAVE];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * 4096 + static_cast<uint64_t>(threadIdx.x) * 16;
  uint32_t c = 0;
  for (uint32_t k = 0; k < 16; ++k) {
    const uint64_t i = base + k;
    if (i < count) c += flags[i] != 0;
  }
  const uint32_t incl = wave_inclusive_scan(c);
  const uint32_t wave = threadIdx.x / GC_WAVE;
  if (lane_id() == GC_WAVE - 1) s_wave[wave] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wave; ++w) before += s_wave[w];
  uint32_t pos = tile_offsets[blockIdx.x] + before + incl - c;
  for (uint32_t k = 0; k < 16; ++k) {
    const uint64_t i = base + k;
    if (i >= count) break;
    if (flags[i] != 0) out[pos++] = keys[i];
  }
}

// CSR arrays from the unique sorted keys (row << id_bits | target):
// targets[i] = low id_bits of key i; offsets[u] = number of keys whose row is
// below u (a binary search per row), for u = 0..n.
extern "C" __global__ __launch_bounds__(256) void csr_from_keys(
    const unsigned long long* __restrict__ keys, uint64_t count, uint32_t n, uint32_t id_bits,
    uint32_t* __restrict__ targets, uint32_t* __restrict__ offsets) {
  const unsigned long long mask = (1ull << id_bits) - 1ull;
  const uint64_t total = count > static_cast<uint64_t>(n) + 1 ? count : static_cast<uint64_t>(n) + 1;
  for (uint64_t i = global_thread(); i < total; i += grid_threads()) {
    if (i < count) targets[i] = static_cast<uint32_t>(keys[i] & mask);
    if (i <= n) {
      const unsigned long long first = static_cast<unsigned long long>(i) << id_bits;
      uint64_t lo = 0, hi = count;
      while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        if (keys[mid] < first) lo = mid + 1;
        else hi = mid;
      }
      offsets[i] = static_cast<uint32_t>(lo);
    }
  }
}
