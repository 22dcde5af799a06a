// Degree-ordered relabelling on the device (the derived layout of graph.cpp):
// the same arrays as the host preparation, built with the radix sort of
// sort.hip.
//   prep_symmetry      one lane per entry: a self-loop, or a binary search for
//                      the reverse entry; counts loops and entries without a
//                      reverse;
//   prep_degree_keys   (degree << id_bits | id) per vertex; sorting them gives
//                      the host's stable counting-sort order (degree, id);
//   prep_perm          perm / inv from the sorted keys;
//   prep_entry_keys    (inv[u] << id_bits | inv[v]) per entry u -> v; sorting
//                      them gives the relabelled rows, each sorted;
//   prep_stats         per relabelled row: DAG out-degree sums, wedges,
//                      intersection work, largest degrees, rows without
//                      entries, and the end of each in-list (for hub slices).
#include "device.hpp"

using namespace gcd;

extern "C" __global__ __launch_bounds__(256) void prep_symmetry(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t n,
    uint64_t m, uint32_t* __restrict__ counts) {
  uint32_t loops = 0, missing = 0;
  for (uint64_t e = global_thread(); e < m; e += grid_threads()) {
    const uint32_t u = row_of_entry(offsets, n, static_cast<uint32_t>(e));
    const uint32_t v = targets[e];
    if (v == u) {
      ++loops;
      continue;
    }
    const uint32_t a = offsets[v], b = offsets[v + 1];
    const uint32_t p = lower_bound_u32(targets, a, b, u);
    if (p == b || targets[p] != u) ++missing;
  }
  loops = wave_sum(loops);
  missing = wave_sum(missing);
  if (lane_id() == 0) {
    if (loops) atomicAdd(&counts[0], loops);
    if (missing) atomicAdd(&counts[1], missing);
  }
}

extern "C" __global__ __launch_bounds__(256) void prep_degree_keys(
    const uint32_t* __restrict__ offsets, uint32_t n, uint32_t id_bits,
    unsigned long long* __restrict__ keys) {
  for (uint64_t v = global_thread(); v < n; v += grid_threads()) {
    const unsigned long long deg = offsets[v + 1] - offsets[v];
    keys[v] = (deg << id_bits) | v;
  }
}

// perm[i] = old id of new vertex i and inv[old] = i, from the sorted keys.
extern "C" __global__ __launch_bounds__(256) void prep_perm(
    const unsigned long long* __restrict__ keys, uint32_t n, uint32_t id_bits,
    uint32_t* __restrict__ perm, uint32_t* __restrict__ inv) {
  const unsigned long long mask = (1ull << id_bits) - 1ull;
  for (uint64_t i = global_thread(); i < n; i += grid_threads()) {
    const uint32_t old = static_cast<uint32_t>(keys[i] & mask);
    perm[i] = old;
    inv[old] = static_cast<uint32_t>(i);
  }
}

extern "C" __global__ __launch_bounds__(256) void prep_entry_keys(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t n,
    uint64_t m, const uint32_t* __restrict__ inv, uint32_t id_bits,
    unsigned long long* __restrict__ keys) {
  for (uint64_t e = global_thread(); e < m; e += grid_threads()) {
    const uint32_t u = row_of_entry(offsets, n, static_cast<uint32_t>(e));
    keys[e] = (static_cast<unsigned long long>(inv[u]) << id_bits) | inv[targets[e]];
  }
}

// stats[0] DAG entries, [1] intersection work sum (d+(i) + d+(j)) over DAG
// entries i -> j, [2] wedges sum d+(i)(d+(i)-1)/2, [3] rows without entries,
// [4] largest degree, [5] largest DAG out-degree. in_end[i] = end of row i's
// in-list (entries below i, a self-loop excluded).
extern "C" __global__ __launch_bounds__(256) void prep_stats(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ fwd, uint32_t n, unsigned long long* __restrict__ stats,
    uint32_t* __restrict__ in_end) {
  unsigned long long oriented = 0, intersect = 0, wedges = 0, empty = 0, max_deg = 0, max_out = 0;
  for (uint64_t ii = global_thread(); ii < n; ii += grid_threads()) {
    const uint32_t i = static_cast<uint32_t>(ii);
    const uint32_t a = offsets[i], b = offsets[i + 1], f = fwd[i];
    const unsigned long long dout = b - f;
    oriented += dout;
    wedges += dout * (dout ? dout - 1 : 0) / 2;
    empty += a == b;
    max_deg = max(max_deg, static_cast<unsigned long long>(b - a));
    max_out = max(max_out, dout);
    for (uint32_t p = f; p < b; ++p) {
      const uint32_t j = targets[p];
      intersect += dout + (offsets[j + 1] - fwd[j]);
    }
    in_end[i] = (f > a && targets[f - 1] == i) ? f - 1 : f;
  }
  oriented = wave_sum(oriented);
  intersect = wave_sum(intersect);
  wedges = wave_sum(wedges);
  empty = wave_sum(empty);
#pragma unroll
  for (int o = GC_WAVE / 2; o > 0; o >>= 1) {
    max_deg = max(max_deg, static_cast<unsigned long long>(__shfl_xor(max_deg, o)));
    max_out = max(max_out, static_cast<unsigned long long>(__shfl_xor(max_out, o)));
  }
  // one set of six atomics per workgroup (per wave, on one line, they took
  // most of the kernel's 5 ms at R-MAT scale 22)
  __shared__ unsigned long long s_v[6][256 / GC_WAVE];
  const uint32_t wave = threadIdx.x / GC_WAVE;
  if (lane_id() == 0) {
    s_v[0][wave] = oriented;
    s_v[1][wave] = intersect;
    s_v[2][wave] = wedges;
    s_v[3][wave] = empty;
    s_v[4][wave] = max_deg;
    s_v[5][wave] = max_out;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t w = 0; w < blockDim.x / GC_WAVE; ++w) {
      for (int k = 0; k < 4; ++k) t[k] += s_v[k][w];
      t[4] = max(t[4], s_v[4][w]);
      t[5] = max(t[5], s_v[5][w]);
    }
    for (int k = 0; k < 4; ++k)
      if (t[k]) atomicAdd(&stats[k], t[k]);
    if (t[4]) atomicMax(&stats[4], t[4]);
    if (t[5]) atomicMax(&stats[5], t[5]);
  }
}
