// k-truss peeling: per-edge triangle support and removal rounds.
//
// Round semantics of crates/grafeo-cuda-algos/src/k_truss.rs:84-133 and
// kernels/k_truss.cu:46-124: each round computes, for every active entry
// u->v with v > u, support = |{w : u->w and v->w both active}| over the FULL
// lists, copies it to the reverse entry v->u, then removes every active entry
// with support < k-2; rounds repeat until one removes nothing.
//
// Two implementations of one round:
//  - direct (any CSR, the reference's per-entry definition): one thread per
//    forward entry merges the two lists;
//  - DAG (symmetric CSR, degree-ordered relabelling): every triangle is found
//    once, at its middle vertex (wedge.hpp), and adds 1 to the support of its
//    three (forward) edges. On a symmetric graph this
//    equals the reference's support except in round 1, where the reference's
//    merge also counts w = u when u has an active self-loop and w = v when v
//    has one (CHANGELOG.md:76-77); kt_dag_finish adds those two terms in round
//    1. Self-loops get support 0 and are removed in round 1, so later rounds
//    need no correction.
#include "device.hpp"
#include "wedge.hpp"

using namespace gcd;

// ---------------------------------------------------------------- direct --

extern "C" __global__ __launch_bounds__(256) void kt_support_direct(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint8_t* __restrict__ removed, uint32_t n, uint32_t m, uint32_t* __restrict__ support) {
  for (uint64_t i = global_thread(); i < m; i += grid_threads()) {
    const uint32_t e = static_cast<uint32_t>(i);
    if (removed[e]) continue;
    const uint32_t u = row_of_entry(offsets, n, e);
    const uint32_t v = targets[e];
    if (v <= u) continue;
    uint32_t j = offsets[u], je = offsets[u + 1];
    uint32_t k = offsets[v], ke = offsets[v + 1];
    uint32_t count = 0;
    while (j < je && k < ke) {
      if (removed[j]) { ++j; continue; }
      if (removed[k]) { ++k; continue; }
      const uint32_t a = targets[j], b = targets[k];
      if (a == b) { ++count; ++j; ++k; }
      else if (a < b) ++j;
      else ++k;
    }
    support[e] = count;
    const uint32_t r = lower_bound_u32(targets, offsets[v], offsets[v + 1], u);
    if (r < offsets[v + 1] && targets[r] == u && !removed[r]) support[r] = count;
  }
}

extern "C" __global__ __launch_bounds__(256) void kt_filter(const uint32_t* __restrict__ support,
                                                            uint8_t* __restrict__ removed,
                                                            uint32_t m, uint32_t threshold,
                                                            uint32_t* __restrict__ removed_count) {
  uint32_t local = 0;
  for (uint64_t i = global_thread(); i < m; i += grid_threads()) {
    if (!removed[i] && support[i] < threshold) {
      removed[i] = 1;
      ++local;
    }
  }
  local = wave_sum(local);
  if (lane_id() == 0 && local) atomicAdd(removed_count, local);
}

extern "C" __global__ __launch_bounds__(256) void kt_emit_flags_direct(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t n,
    uint32_t m, const uint8_t* __restrict__ removed, uint8_t* __restrict__ flags) {
  for (uint64_t i = global_thread(); i < m; i += grid_threads()) {
    const uint32_t e = static_cast<uint32_t>(i);
    const uint32_t u = row_of_entry(offsets, n, e);
    flags[e] = (targets[e] > u && !removed[e]) ? 1 : 0;
  }
}

// ------------------------------------------------------------------- DAG --

// rev[i] = position of u in row v for every forward entry i = (u -> v).
extern "C" __global__ __launch_bounds__(256) void kt_reverse_index(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ fwd, uint32_t n, uint32_t* __restrict__ rev) {
  for (uint64_t uu = global_thread(); uu < n; uu += grid_threads()) {
    const uint32_t u = static_cast<uint32_t>(uu);
    for (uint32_t i = fwd[u]; i < offsets[u + 1]; ++i) {
      const uint32_t v = targets[i];
      rev[i] = lower_bound_u32(targets, offsets[v], fwd[v], u);
    }
  }
}

// Adds 1 to the support of each forward entry of every triangle whose three
// DAG entries are active (enumeration at the middle vertex, see wedge.hpp).
extern "C" __global__ __launch_bounds__(64) void kt_wedge_support(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ fwd, const uint8_t* __restrict__ removed, uint32_t n,
    const uint32_t* __restrict__ hub_units, uint32_t num_hub_units, uint32_t hub_chunk,
    uint32_t table_bits, uint32_t* __restrict__ counters, uint32_t* __restrict__ support) {
  wedge_body<true>(offsets, targets, fwd, removed, n, hub_units, num_hub_units, hub_chunk, table_bits,
                   counters, nullptr, support);
}

__device__ __forceinline__ uint32_t has_loop(const uint32_t* __restrict__ offsets,
                                             const uint32_t* __restrict__ targets,
                                             const uint32_t* __restrict__ fwd, uint32_t x) {
  const uint32_t f = fwd[x];
  return (f > offsets[x] && targets[f - 1] == x) ? 1u : 0u;
}

// Removes active forward entries (and their reverse entries) whose support
// is below `threshold`; in round 1 also every self-loop, and the support of
// an edge gains the self-loops of its two endpoints first.
extern "C" __global__ __launch_bounds__(256) void kt_dag_finish(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ fwd, const uint32_t* __restrict__ rev,
    uint8_t* __restrict__ removed, uint32_t n, const uint32_t* __restrict__ support,
    uint32_t threshold, uint32_t round1, uint32_t* __restrict__ removed_count) {
  uint32_t local = 0;
  for (uint64_t uu = global_thread(); uu < n; uu += grid_threads()) {
    const uint32_t u = static_cast<uint32_t>(uu);
    const uint32_t f = fwd[u], e = offsets[u + 1];
    uint32_t loop_u = 0;
    if (round1) {
      loop_u = has_loop(offsets, targets, fwd, u);
      if (loop_u && !removed[f - 1]) {
        removed[f - 1] = 1;
        ++local;
      }
    }
    for (uint32_t i = f; i < e; ++i) {
      if (removed[i]) continue;
      uint32_t s = support[i];
      if (round1) s += loop_u + has_loop(offsets, targets, fwd, targets[i]);
      if (s < threshold) {
        removed[i] = 1;
        removed[rev[i]] = 1;
        local += 2;
      }
    }
  }
  local = wave_sum(local);
  if (lane_id() == 0 && local) atomicAdd(removed_count, local);
}

// flags[e] = 1 for original entries u->v, v > u, whose relabelled entry
// inv[u] -> inv[v] survived.
extern "C" __global__ __launch_bounds__(256) void kt_emit_flags_relabelled(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t n,
    uint32_t m, const uint32_t* __restrict__ inv, const uint32_t* __restrict__ r_offsets,
    const uint32_t* __restrict__ r_targets, const uint8_t* __restrict__ removed,
    uint8_t* __restrict__ flags) {
  for (uint64_t i = global_thread(); i < m; i += grid_threads()) {
    const uint32_t e = static_cast<uint32_t>(i);
    const uint32_t u = row_of_entry(offsets, n, e);
    const uint32_t v = targets[e];
    uint8_t flag = 0;
    if (v > u) {
      const uint32_t iu = inv[u], iv = inv[v];
      const uint32_t p = lower_bound_u32(r_targets, r_offsets[iu], r_offsets[iu + 1], iv);
      flag = removed[p] ? 0 : 1;
    }
    flags[e] = flag;
  }
}
