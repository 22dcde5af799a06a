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
//    three (forward) edges, once; the rounds then peel (kt_truss_frontier,
//    kt_truss_peel, kt_truss_commit) and only destroy the triangles of the
//    edges they remove. On a symmetric graph this equals the reference's
//    support except in round 1, where the reference's merge also counts w = u
//    when u has an active self-loop and w = v when v has one
//    (CHANGELOG.md:76-77); the first frontier adds those two terms.
//    Self-loops get support 0 and are removed in round 1, so later rounds
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

// ----------------------------------------------------------- truss numbers --
//
// Per-edge trussness (the largest k whose k-truss holds the edge; 2 for an
// edge in no triangle) of a symmetric graph, by peeling on the relabelled
// CSR. Every undirected edge {a, b}, a < b, is identified by its canonical
// entry: the forward entry a -> b. Level k (threshold t = k - 2) removes, in
// batches, every live edge whose support (triangles of live edges) is below
// t; the edges removed at level k have trussness k - 1. A batch is the
// frontier of edges below t; for each frontier edge the triangles it closes
// with two live edges are found by intersecting the two full rows, and each
// such triangle is destroyed exactly once: an edge of the triangle outside
// the frontier loses one support per destroyed triangle, and when two edges
// of a triangle are in the frontier the one with the smaller entry index
// does the decrement. An edge whose support falls to t - 1 joins the next
// frontier. Triangles with an edge removed in an earlier batch were
// destroyed then and are skipped.

#define KT_ALIVE 0u
#define KT_FRONTIER 1u
#define KT_REMOVED 2u
#define KT_NEXT 3u  // fell below the threshold in the current batch; still live in it
#define KT_NONE 4u  // not a canonical entry (reverse direction or self-loop)

// canon[i] = canonical entry of the undirected edge of entry i: i itself for a
// forward entry, the forward entry of the reverse direction otherwise (self-
// loops map to themselves and are never looked up).
extern "C" __global__ __launch_bounds__(256) void kt_canonical_index(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ fwd, uint32_t n,
    uint32_t* __restrict__ canon, uint8_t* __restrict__ state) {
  for (uint64_t uu = global_thread(); uu < n; uu += grid_threads()) {
    const uint32_t u = static_cast<uint32_t>(uu);
    const uint32_t f = fwd[u], e = offsets[u + 1];
    for (uint32_t i = offsets[u]; i < f; ++i) {
      canon[i] = i;  // overwritten below by the forward entry for non-loops
      state[i] = static_cast<uint8_t>(KT_NONE);
    }
    for (uint32_t i = f; i < e; ++i) {
      canon[i] = i;
      state[i] = static_cast<uint8_t>(KT_ALIVE);
    }
  }
}

// Second pass (after kt_canonical_index has finished everywhere): the
// reverse entry of each forward entry points at it.
extern "C" __global__ __launch_bounds__(256) void kt_canonical_reverse(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ fwd,
    const uint32_t* __restrict__ rev, uint32_t n, uint32_t* __restrict__ canon) {
  for (uint64_t uu = global_thread(); uu < n; uu += grid_threads()) {
    const uint32_t u = static_cast<uint32_t>(uu);
    for (uint32_t i = fwd[u]; i < offsets[u + 1]; ++i) canon[rev[i]] = i;
  }
}

// Starts a level: every live edge with support < t joins the frontier,
// appended to queue from counters[0] on. The smallest support among the edges
// that stay live goes to counters[1], so the host can skip levels that would
// remove nothing. With `loops` (the first k-truss round), an edge's support
// also counts the self-loops of its two endpoints, as the reference's first
// round does.
#define KTF_ITEMS 16  // entries per thread: a workgroup owns 4096 consecutive entries

extern "C" __global__ __launch_bounds__(256) void kt_truss_frontier(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ fwd, uint32_t n, uint32_t loops,
    uint8_t* __restrict__ state, const uint32_t* __restrict__ support, uint64_t m, uint32_t t,
    uint32_t* __restrict__ queue, uint32_t* __restrict__ counters) {
  // one queue atomic and one minimum atomic per workgroup (per wave and step,
  // all on one line, they serialized: several ms per call at R-MAT scale 22)
  __shared__ uint32_t s_wave_count[256 / GC_WAVE];
  __shared__ uint32_t s_wave_min[256 / GC_WAVE];
  __shared__ uint32_t s_base;
  const uint64_t first = static_cast<uint64_t>(blockIdx.x) * 256 * KTF_ITEMS;
  uint32_t local_min = 0xFFFFFFFFu;
  uint32_t flags = 0;  // bit k: item k joins the frontier
  for (uint32_t k = 0; k < KTF_ITEMS; ++k) {
    const uint64_t i = first + static_cast<uint64_t>(k) * 256 + threadIdx.x;
    if (i < m && state[i] == KT_ALIVE) {
      uint32_t sup = support[i];
      if (loops) {
        const uint32_t u = row_of_entry(offsets, n, static_cast<uint32_t>(i));
        sup += has_loop(offsets, targets, fwd, u) + has_loop(offsets, targets, fwd, targets[i]);
      }
      if (sup < t) {
        flags |= 1u << k;
        state[i] = static_cast<uint8_t>(KT_FRONTIER);
      } else {
        local_min = min(local_min, sup);
      }
    }
  }
  const uint32_t mine = static_cast<uint32_t>(__popc(flags));
  const uint32_t incl = wave_inclusive_scan(mine);
  for (int off = GC_WAVE / 2; off > 0; off >>= 1) local_min = min(local_min, __shfl_xor(local_min, off));
  const uint32_t wave = threadIdx.x / GC_WAVE;
  if (lane_id() == GC_WAVE - 1) s_wave_count[wave] = incl;
  if (lane_id() == 0) s_wave_min[wave] = local_min;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0, lowest = 0xFFFFFFFFu;
    for (uint32_t w = 0; w < 256 / GC_WAVE; ++w) {
      total += s_wave_count[w];
      lowest = min(lowest, s_wave_min[w]);
    }
    s_base = total ? atomicAdd(&counters[0], total) : 0u;
    if (lowest != 0xFFFFFFFFu) atomicMin(&counters[1], lowest);
  }
  __syncthreads();
  uint32_t pos = s_base + incl - mine;
  for (uint32_t w = 0; w < wave; ++w) pos += s_wave_count[w];
  while (flags != 0) {
    const uint32_t k = static_cast<uint32_t>(__ffs(flags)) - 1;
    flags &= flags - 1;
    queue[pos++] = static_cast<uint32_t>(first + static_cast<uint64_t>(k) * 256 + threadIdx.x);
  }
}

__device__ __forceinline__ void kt_truss_decrement(uint32_t x, uint32_t t,
                                                   uint32_t* __restrict__ support,
                                                   uint8_t* __restrict__ state,
                                                   uint32_t* __restrict__ next_queue,
                                                   uint32_t* __restrict__ next_count) {
  const uint32_t old = atomicSub(&support[x], 1u);
  if (old == t) {
    state[x] = static_cast<uint8_t>(KT_NEXT);
    next_queue[atomicAdd(next_count, 1u)] = x;
  }
}

// One batch: one wave per frontier edge (u, v) walks the shorter of the two
// full rows and binary-searches each neighbour w in the longer one.
extern "C" __global__ __launch_bounds__(256) void kt_truss_peel(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ canon, uint32_t n, const uint32_t* __restrict__ queue,
    uint32_t queue_size, uint8_t* __restrict__ state, uint32_t* __restrict__ support, uint32_t t,
    uint32_t* __restrict__ next_queue, uint32_t* __restrict__ next_count) {
  const uint32_t lane = lane_id();
  const uint64_t wave = global_thread() / GC_WAVE;
  const uint64_t waves = grid_threads() / GC_WAVE;
  for (uint64_t q = wave; q < queue_size; q += waves) {
    const uint32_t e = queue[q];
    const uint32_t u = row_of_entry(offsets, n, e);
    const uint32_t v = targets[e];
    const uint32_t du = offsets[u + 1] - offsets[u], dv = offsets[v + 1] - offsets[v];
    const uint32_t s = du <= dv ? u : v;
    const uint32_t l = du <= dv ? v : u;
    const uint32_t s_begin = offsets[s], s_end = offsets[s + 1];
    const uint32_t l_begin = offsets[l], l_end = offsets[l + 1];
    for (uint32_t es = s_begin + lane; es < s_end; es += GC_WAVE) {
      const uint32_t w = targets[es];
      if (w == u || w == v) continue;
      const uint32_t p = lower_bound_u32(targets, l_begin, l_end, w);
      if (p == l_end || targets[p] != w) continue;
      const uint32_t c1 = canon[es], c2 = canon[p];
      const uint32_t st1 = state[c1], st2 = state[c2];
      if (st1 == KT_REMOVED || st2 == KT_REMOVED) continue;
      const bool f1 = st1 == KT_FRONTIER, f2 = st2 == KT_FRONTIER;
      if (!f1 && (!f2 || e < c2)) kt_truss_decrement(c1, t, support, state, next_queue, next_count);
      if (!f2 && (!f1 || e < c1)) kt_truss_decrement(c2, t, support, state, next_queue, next_count);
    }
  }
}

// Ends a batch: the frontier is removed with trussness `truss` (stored in its
// support slot), and the next frontier becomes the frontier.
extern "C" __global__ __launch_bounds__(256) void kt_truss_commit(
    const uint32_t* __restrict__ queue, uint32_t queue_size,
    const uint32_t* __restrict__ next_queue, uint32_t next_size, uint8_t* __restrict__ state,
    uint32_t* __restrict__ support, uint32_t truss) {
  const uint64_t total = queue_size > next_size ? queue_size : next_size;
  for (uint64_t i = global_thread(); i < total; i += grid_threads()) {
    if (i < queue_size) {
      const uint32_t x = queue[i];
      state[x] = static_cast<uint8_t>(KT_REMOVED);
      support[x] = truss;
    }
    if (i < next_size) state[next_queue[i]] = static_cast<uint8_t>(KT_FRONTIER);
  }
}

// flags[e] = 1 for original entries u -> v, v > u, whose undirected edge
// {inv[u], inv[v]} is still live (its canonical entry not KT_REMOVED).
extern "C" __global__ __launch_bounds__(256) void kt_emit_flags_state(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t n,
    uint32_t m, const uint32_t* __restrict__ inv, const uint32_t* __restrict__ r_offsets,
    const uint32_t* __restrict__ r_targets, const uint32_t* __restrict__ canon,
    const uint8_t* __restrict__ state, uint8_t* __restrict__ flags) {
  for (uint64_t i = global_thread(); i < m; i += grid_threads()) {
    const uint32_t e = static_cast<uint32_t>(i);
    const uint32_t u = row_of_entry(offsets, n, e);
    const uint32_t v = targets[e];
    uint8_t flag = 0;
    if (v > u) {
      const uint32_t iu = inv[u], iv = inv[v];
      const uint32_t p = lower_bound_u32(r_targets, r_offsets[iu], r_offsets[iu + 1], iv);
      flag = state[canon[p]] == KT_REMOVED ? 0 : 1;
    }
    flags[e] = flag;
  }
}

// flags[e] = 1 and values[e] = trussness for original entries u -> v, v > u;
// the relabelled canonical entry of {inv[u], inv[v]} holds the trussness.
extern "C" __global__ __launch_bounds__(256) void kt_truss_values(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t n,
    uint32_t m, const uint32_t* __restrict__ inv, const uint32_t* __restrict__ r_offsets,
    const uint32_t* __restrict__ r_targets, const uint32_t* __restrict__ truss,
    uint8_t* __restrict__ flags, uint32_t* __restrict__ values) {
  for (uint64_t i = global_thread(); i < m; i += grid_threads()) {
    const uint32_t e = static_cast<uint32_t>(i);
    const uint32_t u = row_of_entry(offsets, n, e);
    const uint32_t v = targets[e];
    uint8_t flag = 0;
    uint32_t value = 0;
    if (v > u) {
      const uint32_t a = min(inv[u], inv[v]), b = max(inv[u], inv[v]);
      const uint32_t p = lower_bound_u32(r_targets, r_offsets[a], r_offsets[a + 1], b);
      flag = 1;
      value = truss[p];
    }
    flags[e] = flag;
    values[e] = value;
  }
}
