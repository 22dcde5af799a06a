// Direction-optimizing level-synchronous BFS.
//
// Same result as crates/grafeo-cuda-algos/kernels/bfs.cu:54-83 (distance =
// BFS level along the CSR lists, -1 if unreachable): a vertex's level does
// not depend on which frontier vertex discovers it, so any schedule below is
// bit-exact with the reference's top-down expansion.
//  - bfs_topdown: the frontier is a queue; a wave takes 64 frontier vertices,
//    prefix-sums their degrees and spreads the edges evenly over its lanes
//    (hubs and leaves in one pass); an unvisited neighbour is claimed with a
//    compare-and-swap on its distance, and winners are appended to the next
//    queue with one atomic per wave (ballot + popcount ranks).
//  - bfs_bottomup: the frontier is a bitmap; each lane owns one unvisited
//    vertex and scans its list from the end (highest degree-ordered ids, the
//    hubs most likely to be in the frontier) until it finds a parent; the
//    wave writes its 64 result bits as one 64-bit bitmap word, so the next
//    bitmap needs no clearing and no atomics.
// Both report the size of the next frontier and the degree sum of its
// vertices, which drive the direction choice.
#include "device.hpp"

using namespace gcd;

struct BfsCounters {
  uint32_t next_size;
  uint32_t pad;
  unsigned long long next_edges;
};

extern "C" __global__ __launch_bounds__(256) void bfs_init(int32_t* __restrict__ dist, uint32_t n,
                                                           uint32_t source, uint32_t* __restrict__ queue) {
  for (uint64_t v = global_thread(); v < n; v += grid_threads()) dist[v] = v == source ? 0 : -1;
  if (global_thread() == 0) queue[0] = source;
}

__device__ __forceinline__ void bfs_append(bool discovered, uint32_t v, uint32_t deg,
                                           uint32_t* __restrict__ next_queue,
                                           BfsCounters* __restrict__ counters) {
  const uint64_t mask = __ballot(discovered);
  if (mask == 0) return;
  const uint32_t leader = __builtin_ctzll(mask);
  uint32_t base = 0;
  unsigned long long edges = wave_sum(static_cast<unsigned long long>(discovered ? deg : 0u));
  if (lane_id() == leader) {
    base = atomicAdd(&counters->next_size, static_cast<uint32_t>(__popcll(mask)));
    atomicAdd(&counters->next_edges, edges);
  }
  base = __shfl(base, static_cast<int>(leader));
  if (discovered && next_queue != nullptr) next_queue[base + lanes_below(mask)] = v;
}

extern "C" __global__ __launch_bounds__(256) void bfs_topdown(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    int32_t* __restrict__ dist, const uint32_t* __restrict__ queue, uint32_t queue_size,
    uint32_t* __restrict__ next_queue, BfsCounters* __restrict__ counters, int32_t next_level) {
  const uint32_t lane = lane_id();
  const uint64_t waves = grid_threads() / GC_WAVE;
  for (uint64_t base = (global_thread() / GC_WAVE) * GC_WAVE; base < queue_size;
       base += waves * GC_WAVE) {
    const uint64_t idx = base + lane;
    uint32_t st = 0, len = 0;
    if (idx < queue_size) {
      const uint32_t u = queue[idx];
      st = offsets[u];
      len = offsets[u + 1] - st;
    }
    const uint32_t incl = wave_inclusive_scan(len);
    const uint32_t total = __shfl(incl, GC_WAVE - 1);
    for (uint32_t t0 = 0; t0 < total; t0 += GC_WAVE) {
      const uint32_t t = t0 + lane;
      const bool active = t < total;
      const uint32_t tt = active ? t : total - 1;
      const uint32_t owner = wave_upper_bound(incl, tt);
      const uint32_t prev = __shfl(incl, static_cast<int>(owner == 0 ? 0 : owner - 1));
      const uint32_t excl = owner == 0 ? 0 : prev;
      const uint32_t ost = __shfl(st, static_cast<int>(owner));
      bool discovered = false;
      uint32_t v = 0, deg = 0;
      if (active) {
        v = targets[ost + (tt - excl)];
        if (dist[v] < 0) {
          discovered = atomicCAS(&dist[v], -1, next_level) == -1;
          if (discovered) deg = offsets[v + 1] - offsets[v];
        }
      }
      bfs_append(discovered, v, deg, next_queue, counters);
    }
  }
}

extern "C" __global__ __launch_bounds__(256) void bfs_bottomup(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    int32_t* __restrict__ dist, const uint32_t* __restrict__ front, uint64_t* __restrict__ next,
    uint32_t n, BfsCounters* __restrict__ counters, int32_t next_level) {
  const uint32_t lane = lane_id();
  const uint64_t waves = grid_threads() / GC_WAVE;
  const uint64_t chunks = (static_cast<uint64_t>(n) + GC_WAVE - 1) / GC_WAVE;
  for (uint64_t chunk = global_thread() / GC_WAVE; chunk < chunks; chunk += waves) {
    const uint64_t v64 = chunk * GC_WAVE + lane;
    bool found = false;
    uint32_t deg = 0;
    if (v64 < n) {
      const uint32_t v = static_cast<uint32_t>(v64);
      if (dist[v] < 0) {
        const uint32_t b = offsets[v], e = offsets[v + 1];
        // four candidates per step, their loads issued together
        for (uint32_t i = e; i > b && !found;) {
          const uint32_t cnt = min(4u, i - b);
          uint32_t u[4], word[4];
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) u[k] = k < cnt ? targets[i - 1 - k] : 0u;
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) word[k] = k < cnt ? front[u[k] >> 5] : 0u;
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) found |= k < cnt && ((word[k] >> (u[k] & 31)) & 1u);
          i -= cnt;
        }
        if (found) {
          dist[v] = next_level;
          deg = e - b;
        }
      }
    }
    const uint64_t mask = __ballot(found);
    if (lane == 0) next[chunk] = mask;
    if (mask != 0) {
      unsigned long long edges = wave_sum(static_cast<unsigned long long>(deg));
      if (lane == 0) {
        atomicAdd(&counters->next_size, static_cast<uint32_t>(__popcll(mask)));
        atomicAdd(&counters->next_edges, edges);
      }
    }
  }
}

// degree[i] = degree of queue[i] (input of the frontier's edge prefix sum).
extern "C" __global__ __launch_bounds__(256) void bfs_frontier_degrees(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ queue, uint32_t size,
    uint32_t* __restrict__ degree) {
  for (uint64_t i = global_thread(); i < size; i += grid_threads()) {
    const uint32_t u = queue[i];
    degree[i] = offsets[u + 1] - offsets[u];
  }
}

#define TD_EDGES_PER_WAVE 256

// Top-down step balanced over the frontier's edges: `start` is the exclusive
// prefix sum of the frontier's degrees (`total` edges in all). Each wave takes
// TD_EDGES_PER_WAVE consecutive edge ranks; a lane finds the frontier vertex
// owning its edge by binary search in `start`, restricted to the owners of
// the wave's first and last edge, so a hub's edges spread over every CU.
extern "C" __global__ __launch_bounds__(256) void bfs_topdown_edges(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    int32_t* __restrict__ dist, const uint32_t* __restrict__ queue, uint32_t queue_size,
    const uint32_t* __restrict__ start, uint32_t total, uint32_t* __restrict__ next_queue,
    BfsCounters* __restrict__ counters, int32_t next_level) {
  const uint32_t lane = lane_id();
  const uint64_t waves = grid_threads() / GC_WAVE;
  for (uint64_t base = (global_thread() / GC_WAVE) * TD_EDGES_PER_WAVE; base < total;
       base += waves * TD_EDGES_PER_WAVE) {
    const uint32_t first = static_cast<uint32_t>(base);
    const uint32_t last = static_cast<uint32_t>(min(base + TD_EDGES_PER_WAVE, static_cast<uint64_t>(total)) - 1);
    // owners of the first and last edge of this wave's range (largest j with start[j] <= t)
    uint32_t probe = lane == 0 ? first : last;
    uint32_t lo = 0, hi = queue_size - 1;
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo + 1) >> 1);
      if (start[mid] <= probe) lo = mid;
      else hi = mid - 1;
    }
    const uint32_t own_lo = __shfl(lo, 0);
    const uint32_t own_hi = __shfl(lo, GC_WAVE - 1);
    for (uint32_t k = 0; k < TD_EDGES_PER_WAVE; k += GC_WAVE) {
      const uint32_t t = first + k + lane;
      const bool active = t <= last;
      bool discovered = false;
      uint32_t v = 0, deg = 0;
      if (active) {
        uint32_t a = own_lo, b = own_hi;
        while (a < b) {
          const uint32_t mid = a + ((b - a + 1) >> 1);
          if (start[mid] <= t) a = mid;
          else b = mid - 1;
        }
        const uint32_t u = queue[a];
        v = targets[offsets[u] + (t - start[a])];
        if (dist[v] < 0) {
          discovered = atomicCAS(&dist[v], -1, next_level) == -1;
          if (discovered) deg = offsets[v + 1] - offsets[v];
        }
      }
      bfs_append(discovered, v, deg, next_queue, counters);
    }
  }
}

// Sets the bitmap bits of the queue's vertices (the bitmap is cleared first).
extern "C" __global__ __launch_bounds__(256) void bfs_queue_to_bitmap(
    const uint32_t* __restrict__ queue, uint32_t size, uint32_t* __restrict__ bitmap) {
  for (uint64_t i = global_thread(); i < size; i += grid_threads()) {
    const uint32_t v = queue[i];
    atomicOr(&bitmap[v >> 5], 1u << (v & 31));
  }
}

// Lists the set bits of a bitmap as a queue (any order).
extern "C" __global__ __launch_bounds__(256) void bfs_bitmap_to_queue(
    const uint64_t* __restrict__ bitmap, uint64_t words, uint32_t* __restrict__ queue,
    uint32_t* __restrict__ size) {
  const uint32_t lane = lane_id();
  const uint64_t waves = grid_threads() / GC_WAVE;
  for (uint64_t base = (global_thread() / GC_WAVE) * GC_WAVE; base < words; base += waves * GC_WAVE) {
    const uint64_t w = base + lane;
    const uint64_t bits = w < words ? bitmap[w] : 0ull;
    const uint32_t c = static_cast<uint32_t>(__popcll(bits));
    const uint32_t incl = wave_inclusive_scan(c);
    const uint32_t total = __shfl(incl, GC_WAVE - 1);
    uint32_t start = 0;
    if (lane == 0 && total) start = atomicAdd(size, total);
    start = __shfl(start, 0);
    uint32_t pos = start + incl - c;
    uint64_t rest = bits;
    while (rest) {
      const uint32_t bit = __builtin_ctzll(rest);
      rest &= rest - 1;
      queue[pos++] = static_cast<uint32_t>(w * 64 + bit);
    }
  }
}
