// Multi-source BFS: up to 128 sources traversed together.
//
// Every vertex carries three 128-bit masks (bit s = source s), in the
// degree-ordered relabelled ids: `seen` (sources that have reached it),
// `visit` (sources whose frontier holds it at the current level) and `next`
// (sources reaching it at the next level). A level is one of
//   - top-down: each frontier vertex v ORs visit[v] & ~seen[w] into next[w]
//     of every neighbour w (64-bit atomic ORs);
//   - bottom-up: each vertex w still missing sources ORs visit[v] over its
//     neighbours v and stops as soon as every missing source is covered (no
//     atomics; needs a symmetric CSR so that w's list holds its parents);
// followed by ms_finalize (new = next & ~seen; seen |= new; visit = new) and
// ms_write_distances (distance = level for each new bit, in the caller's
// ids, one array of n per source). Distances equal those of one BFS per
// source (bfs.hip; crates/grafeo-cuda-algos/kernels/bfs.cu:54-83) whatever the
// direction of each level.
#include "device.hpp"

using namespace gcd;

struct MsCounters {
  uint32_t frontier_vertices;  // vertices with a nonzero new mask
  uint32_t pad;
  unsigned long long frontier_edges;    // sum of their degrees
  unsigned long long unexplored_edges;  // sum of degrees of vertices still missing a source
};

struct Mask {
  unsigned long long lo, hi;
};

__device__ __forceinline__ Mask load_mask(const ulonglong2* p) {
  const ulonglong2 v = *p;
  return Mask{v.x, v.y};
}

__device__ __forceinline__ void store_mask(ulonglong2* p, Mask m) {
  *p = make_ulonglong2(m.lo, m.hi);
}

__device__ __forceinline__ Mask wave_or(Mask m) {
#pragma unroll
  for (int o = GC_WAVE / 2; o > 0; o >>= 1) {
    m.lo |= __shfl_xor(m.lo, o);
    m.hi |= __shfl_xor(m.hi, o);
  }
  return m;
}

// Sets the bit of source s at its (relabelled) vertex; one thread, since two
// sources may share a vertex.
extern "C" __global__ __launch_bounds__(64) void ms_seed(const uint32_t* __restrict__ sources_rl,
                                                         uint32_t count,
                                                         ulonglong2* __restrict__ seen,
                                                         ulonglong2* __restrict__ visit) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (uint32_t s = 0; s < count; ++s) {
    const uint32_t v = sources_rl[s];
    Mask m = load_mask(&seen[v]);
    if (s < 64) m.lo |= 1ull << s;
    else m.hi |= 1ull << (s - 64);
    store_mask(&seen[v], m);
    store_mask(&visit[v], m);
  }
}

// Top-down: one wave per frontier vertex, lanes over its list.
extern "C" __global__ __launch_bounds__(256) void ms_topdown(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ queue, uint32_t queue_size, const ulonglong2* __restrict__ visit,
    const ulonglong2* __restrict__ seen, ulonglong2* __restrict__ next) {
  const uint32_t lane = lane_id();
  const uint64_t waves = grid_threads() / GC_WAVE;
  for (uint64_t q = global_thread() / GC_WAVE; q < queue_size; q += waves) {
    const uint32_t v = queue[q];
    const Mask mv = load_mask(&visit[v]);
    const uint32_t end = offsets[v + 1];
    for (uint32_t i = offsets[v] + lane; i < end; i += GC_WAVE) {
      const uint32_t w = targets[i];
      const Mask sw = load_mask(&seen[w]);
      const unsigned long long add_lo = mv.lo & ~sw.lo, add_hi = mv.hi & ~sw.hi;
      if (add_lo) atomicOr(&next[w].x, add_lo);
      if (add_hi) atomicOr(&next[w].y, add_hi);
    }
  }
}

// Bottom-up for vertices [begin, end) of low degree: one thread per vertex.
extern "C" __global__ __launch_bounds__(256) void ms_bottomup_thread(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t begin,
    uint32_t end, unsigned long long active_lo, unsigned long long active_hi,
    const ulonglong2* __restrict__ visit, const ulonglong2* __restrict__ seen,
    ulonglong2* __restrict__ next) {
  for (uint64_t ww = begin + global_thread(); ww < end; ww += grid_threads()) {
    const uint32_t w = static_cast<uint32_t>(ww);
    const Mask sw = load_mask(&seen[w]);
    const unsigned long long miss_lo = active_lo & ~sw.lo, miss_hi = active_hi & ~sw.hi;
    if ((miss_lo | miss_hi) == 0) continue;
    unsigned long long acc_lo = 0, acc_hi = 0;
    const uint32_t stop = offsets[w + 1];
    for (uint32_t i = offsets[w]; i < stop; ++i) {
      const Mask mv = load_mask(&visit[targets[i]]);
      acc_lo |= mv.lo;
      acc_hi |= mv.hi;
      if ((acc_lo & miss_lo) == miss_lo && (acc_hi & miss_hi) == miss_hi) break;
    }
    acc_lo &= miss_lo;
    acc_hi &= miss_hi;
    if (acc_lo | acc_hi) store_mask(&next[w], Mask{acc_lo, acc_hi});
  }
}

// Bottom-up for vertices [begin, end) of high degree: one wave per vertex,
// up to 512 neighbours per step (up to eight per lane in flight, as many
// groups of 64 as the list has left; a lane past the list's end reads its
// last neighbour again, which ORs nothing new), stopping once the missing
// sources are covered. The masks are ORed across the wave once per step: with
// 128 sources a vertex seldom covers all its missing ones early, and one
// gather per lane and an OR across the wave per 64 neighbours left the level
// bound by their round trips (at R-MAT scale 22 with 128 sources, the five
// bottom-up levels: 6.3 ms that way, 3.97 with four groups per step, 3.64
// with eight).
#define MSB_GROUPS 8  // groups of 64 neighbours per step
extern "C" __global__ __launch_bounds__(256) void ms_bottomup_wave(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t begin,
    uint32_t end, unsigned long long active_lo, unsigned long long active_hi,
    const ulonglong2* __restrict__ visit, const ulonglong2* __restrict__ seen,
    ulonglong2* __restrict__ next) {
  const uint32_t lane = lane_id();
  const uint64_t waves = grid_threads() / GC_WAVE;
  for (uint64_t ww = begin + global_thread() / GC_WAVE; ww < end; ww += waves) {
    const uint32_t w = static_cast<uint32_t>(ww);
    const Mask sw = load_mask(&seen[w]);
    const unsigned long long miss_lo = active_lo & ~sw.lo, miss_hi = active_hi & ~sw.hi;
    if ((miss_lo | miss_hi) == 0) continue;
    Mask acc{0, 0};
    const uint32_t stop = offsets[w + 1];
    for (uint32_t base = offsets[w]; base < stop; base += MSB_GROUPS * GC_WAVE) {
      const uint32_t left = stop - base;  // wave-uniform: the groups of 64 past it are skipped
      uint32_t v[MSB_GROUPS];
#pragma unroll
      for (int q = 0; q < MSB_GROUPS; ++q)
        if (q * GC_WAVE < left) v[q] = targets[min(base + q * GC_WAVE + lane, stop - 1)];
#pragma unroll
      for (int q = 0; q < MSB_GROUPS; ++q) {
        if (q * GC_WAVE < left) {
          const Mask mv = load_mask(&visit[v[q]]);
          acc.lo |= mv.lo;
          acc.hi |= mv.hi;
        }
      }
      acc = wave_or(acc);
      if ((acc.lo & miss_lo) == miss_lo && (acc.hi & miss_hi) == miss_hi) break;
    }
    acc.lo &= miss_lo;
    acc.hi &= miss_hi;
    if (lane == 0 && (acc.lo | acc.hi)) store_mask(&next[w], acc);
  }
}

// Ends a level: new = next & ~seen, seen |= new, visit = new, next = 0; the new
// frontier is appended to `queue` and counted. Workgroup b owns the vertices
// [b * MSF_RANGE, (b + 1) * MSF_RANGE), MSF_ITEMS per thread (coalesced): it
// counts its frontier first, takes its queue slots with one atomic, and adds
// its edge sums with one atomic each. (With one append atomic per wave and
// step and two edge atomics per wave, all on one line, a level took 770 us at
// R-MAT scale 22.)
#define MSF_THREADS 256
#define MSF_ITEMS 16
#define MSF_RANGE (MSF_THREADS * MSF_ITEMS)

extern "C" __global__ __launch_bounds__(MSF_THREADS) void ms_finalize(
    const uint32_t* __restrict__ offsets, uint32_t n, unsigned long long active_lo,
    unsigned long long active_hi, ulonglong2* __restrict__ seen, ulonglong2* __restrict__ visit,
    ulonglong2* __restrict__ next, uint32_t* __restrict__ queue, MsCounters* __restrict__ counters) {
  __shared__ uint32_t s_wave_count[MSF_THREADS / GC_WAVE];
  __shared__ unsigned long long s_f[MSF_THREADS / GC_WAVE], s_u[MSF_THREADS / GC_WAVE];
  __shared__ uint32_t s_base;
  const uint64_t first = static_cast<uint64_t>(blockIdx.x) * MSF_RANGE;
  unsigned long long f_edges = 0, u_edges = 0;
  uint32_t flags = 0;  // bit k: item k is in the new frontier
#pragma unroll 4
  for (uint32_t k = 0; k < MSF_ITEMS; ++k) {
    const uint64_t ww = first + static_cast<uint64_t>(k) * MSF_THREADS + threadIdx.x;
    if (ww >= n) continue;
    const uint32_t w = static_cast<uint32_t>(ww);
    const Mask nx = load_mask(&next[w]);
    Mask sw = load_mask(&seen[w]);
    const Mask nw{nx.lo & ~sw.lo, nx.hi & ~sw.hi};
    const Mask old_visit = load_mask(&visit[w]);
    if (nx.lo | nx.hi) store_mask(&next[w], Mask{0, 0});
    if (nw.lo | nw.hi) {
      sw.lo |= nw.lo;
      sw.hi |= nw.hi;
      store_mask(&seen[w], sw);
    }
    if ((old_visit.lo | old_visit.hi | nw.lo | nw.hi) != 0) store_mask(&visit[w], nw);
    const uint32_t deg = offsets[w + 1] - offsets[w];
    if (nw.lo | nw.hi) {
      flags |= 1u << k;
      f_edges += deg;
    }
    if ((sw.lo & active_lo) != active_lo || (sw.hi & active_hi) != active_hi) u_edges += deg;
  }
  // the workgroup's slots: exclusive prefix of the per-thread counts
  const uint32_t mine = static_cast<uint32_t>(__popc(flags));
  const uint32_t incl = wave_inclusive_scan(mine);
  const uint32_t wave = threadIdx.x / GC_WAVE;
  f_edges = wave_sum(f_edges);
  u_edges = wave_sum(u_edges);
  if (lane_id() == GC_WAVE - 1) s_wave_count[wave] = incl;
  if (lane_id() == 0) {
    s_f[wave] = f_edges;
    s_u[wave] = u_edges;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    unsigned long long tf = 0, tu = 0;
    for (uint32_t w2 = 0; w2 < MSF_THREADS / GC_WAVE; ++w2) {
      total += s_wave_count[w2];
      tf += s_f[w2];
      tu += s_u[w2];
    }
    s_base = total ? atomicAdd(&counters->frontier_vertices, total) : 0u;
    if (tf) atomicAdd(&counters->frontier_edges, tf);
    if (tu) atomicAdd(&counters->unexplored_edges, tu);
  }
  __syncthreads();
  uint32_t pos = s_base + incl - mine;
  for (uint32_t w2 = 0; w2 < wave; ++w2) pos += s_wave_count[w2];
  while (flags != 0) {
    const uint32_t k = static_cast<uint32_t>(__ffs(flags)) - 1;
    flags &= flags - 1;
    queue[pos++] = static_cast<uint32_t>(first + static_cast<uint64_t>(k) * MSF_THREADS + threadIdx.x);
  }
}

// distances[s * n + u] = level for every source s newly reaching caller
// vertex u (relabelled id inv[u]).
extern "C" __global__ __launch_bounds__(256) void ms_write_distances(
    const uint32_t* __restrict__ inv, uint32_t n, const ulonglong2* __restrict__ visit,
    int32_t level, int32_t* __restrict__ dist) {
  for (uint64_t uu = global_thread(); uu < n; uu += grid_threads()) {
    const uint32_t u = static_cast<uint32_t>(uu);
    Mask m = load_mask(&visit[inv[u]]);
    while (m.lo) {
      const uint32_t s = __builtin_ctzll(m.lo);
      m.lo &= m.lo - 1;
      dist[static_cast<uint64_t>(s) * n + u] = level;
    }
    while (m.hi) {
      const uint32_t s = 64u + __builtin_ctzll(m.hi);
      m.hi &= m.hi - 1;
      dist[static_cast<uint64_t>(s) * n + u] = level;
    }
  }
}

// The distances as levels in a byte matrix lv[v * 128 + s] (relabelled v,
// 0xFF = not reached yet), set at each level by one lane per vertex for the
// sources that reached it there: a vertex's bytes share one 128-byte line, so
// a level's stores land on one line per newly reached vertex instead of one
// per (source, vertex) pair scattered over 128 rows of caller ids (2.5 of 10.4
// ms at R-MAT 22 with 128 sources). ms_levels_to_dist then writes every row
// once, coalesced, in caller ids; traversals deeper than MS_MAX_BYTE_LEVEL
// levels switch to ms_write_distances from there on.
#define MS_MAX_BYTE_LEVEL 254

extern "C" __global__ __launch_bounds__(256) void ms_write_levels(uint32_t n, const ulonglong2* __restrict__ visit,
                                                                  uint32_t level, uint8_t* __restrict__ lv) {
  // eight lanes per vertex, sixteen bytes (sources 16c .. 16c + 15) each: a wave rewrites
  // eight consecutive 128-byte rows, only the chunks that gained a source
  const uint32_t c = lane_id() & 7u;
  const uint64_t groups = grid_threads() / 8;
  for (uint64_t vv = global_thread() / 8; vv < n; vv += groups) {
    const Mask m = load_mask(&visit[vv]);
    const uint32_t bits = static_cast<uint32_t>(((c < 4 ? m.lo : m.hi) >> (16u * (c & 3u))) & 0xFFFFu);
    if (bits == 0) continue;
    uint4* chunk = reinterpret_cast<uint4*>(lv + 128ull * vv) + c;
    uint4 q = *chunk;
    uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (uint32_t b = 0; b < 16; ++b) {
      if ((bits >> b) & 1u) {
        const uint32_t sh = 8u * (b & 3u);
        w[b >> 2] = (w[b >> 2] & ~(0xFFu << sh)) | (level << sh);
      }
    }
    *chunk = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// lv bytes of the sources at level 0 (relabelled ids in sources_rl).
extern "C" __global__ __launch_bounds__(64) void ms_seed_levels(const uint32_t* __restrict__ sources_rl,
                                                                uint32_t count, uint8_t* __restrict__ lv) {
  const uint32_t s = threadIdx.x + blockIdx.x * blockDim.x;
  if (s < count) lv[128ull * sources_rl[s] + s] = 0;
}

// dist[s * n + u] for every source s < count and caller vertex u, from the
// bytes of relabelled vertex inv[u] (0xFF: -1). With `vec4` (n a multiple of 4
// and dist 16-byte aligned, checked by the host), a lane takes four consecutive caller vertices and writes the four distances of each
// source with one 16-byte store (the sources in two halves of 64, 64 bytes of
// each row in registers); otherwise one lane per vertex and 4-byte stores.
__device__ __forceinline__ int32_t ms_level_dist(uint32_t word, uint32_t s) {
  const uint32_t b = (word >> (8u * (s & 3u))) & 0xFFu;
  return b == 0xFFu ? -1 : static_cast<int32_t>(b);
}

extern "C" __global__ __launch_bounds__(256) void ms_levels_to_dist(const uint32_t* __restrict__ inv, uint32_t n,
                                                                    uint32_t count, const uint8_t* __restrict__ lv,
                                                                    int32_t* __restrict__ dist, uint32_t vec4) {
  if (vec4) {
    const uint64_t quads = n / 4;
    for (uint64_t qq = global_thread(); qq < quads; qq += grid_threads()) {
      const uint64_t u0 = 4 * qq;
      const uint32_t* rows[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) rows[j] = reinterpret_cast<const uint32_t*>(lv + 128ull * inv[u0 + j]);
#pragma unroll
      for (uint32_t half = 0; half < 2; ++half) {
        if (64 * half >= count) break;
        uint32_t w[4][16];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint4* r = reinterpret_cast<const uint4*>(rows[j] + 16 * half);
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const uint4 x = r[k];
            w[j][4 * k] = x.x;
            w[j][4 * k + 1] = x.y;
            w[j][4 * k + 2] = x.z;
            w[j][4 * k + 3] = x.w;
          }
        }
#pragma unroll
        for (uint32_t t = 0; t < 64; ++t) {
          const uint32_t s = 64 * half + t;
          if (s >= count) break;
          const int4 d = make_int4(ms_level_dist(w[0][t >> 2], t), ms_level_dist(w[1][t >> 2], t),
                                   ms_level_dist(w[2][t >> 2], t), ms_level_dist(w[3][t >> 2], t));
          *reinterpret_cast<int4*>(dist + static_cast<uint64_t>(s) * n + u0) = d;
        }
      }
    }
    return;
  }
  for (uint64_t uu = global_thread(); uu < n; uu += grid_threads()) {
    const uint32_t* row = reinterpret_cast<const uint32_t*>(lv + 128ull * inv[uu]);
    for (uint32_t s = 0; s < count; ++s) dist[static_cast<uint64_t>(s) * n + uu] = ms_level_dist(row[s >> 2], s);
  }
}

// Per source s (workgroups s * blocks_per_source ...): reachable[s] += count
// of distances >= 0 and max_depth[s] = max distance, over the workgroup's
// share of the n vertices.
extern "C" __global__ __launch_bounds__(256) void ms_summarize(const int32_t* __restrict__ dist,
                                                               uint32_t n,
                                                               uint32_t blocks_per_source,
                                                               uint32_t* __restrict__ reachable,
                                                               int32_t* __restrict__ max_depth) {
  const uint32_t s = blockIdx.x / blocks_per_source;
  const uint32_t part = blockIdx.x % blocks_per_source;
  const int32_t* d = dist + static_cast<uint64_t>(s) * n;
  uint32_t count = 0;
  int32_t deepest = 0;
  for (uint64_t u = static_cast<uint64_t>(part) * blockDim.x + threadIdx.x; u < n;
       u += static_cast<uint64_t>(blocks_per_source) * blockDim.x) {
    const int32_t x = d[u];
    if (x >= 0) {
      ++count;
      deepest = max(deepest, x);
    }
  }
  count = wave_sum(count);
#pragma unroll
  for (int o = GC_WAVE / 2; o > 0; o >>= 1) deepest = max(deepest, __shfl_xor(deepest, o));
  // one atomic pair per workgroup (the 128 sources' counters share four lines)
  __shared__ uint32_t s_count[256 / GC_WAVE];
  __shared__ int32_t s_deep[256 / GC_WAVE];
  if (lane_id() == 0) {
    s_count[threadIdx.x / GC_WAVE] = count;
    s_deep[threadIdx.x / GC_WAVE] = deepest;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t c = 0;
    int32_t dd = 0;
    for (uint32_t w = 0; w < blockDim.x / GC_WAVE; ++w) {
      c += s_count[w];
      dd = max(dd, s_deep[w]);
    }
    if (c) atomicAdd(&reachable[s], c);
    if (dd) atomicMax(&max_depth[s], dd);
  }
}
