// Direction-optimizing level-synchronous BFS.
//
// Same result as crates/grafeo-cuda-algos/kernels/bfs.cu:54-83 (distance =
// BFS level along the CSR lists, -1 if unreachable): a vertex's level does
// not depend on which frontier vertex discovers it, so any schedule below is
// bit-exact with the reference's top-down expansion.
//  - bfs_topdown_edges: the frontier is a queue; the frontier's degrees are
//    prefix-summed and each wave takes 256 consecutive frontier edges (hubs
//    and leaves spread evenly); an unvisited neighbour is claimed with a
//    compare-and-swap on its distance, and winners are appended to the next
//    queue with one atomic per wave (ballot + popcount ranks).
//  - bfs_bottomup: the frontier is a bitmap; each lane owns one unvisited
//    vertex and scans its list from the end (highest degree-ordered ids, the
//    hubs most likely to be in the frontier) until it finds a parent; the
//    wave writes its 64 result bits as one 64-bit bitmap word, so the next
//    bitmap needs no clearing and no atomics. The first four candidates come
//    from a transposed copy of every row's four last entries (bfs_top_build),
//    which the wave's lanes read coalesced: most vertices find their parent
//    there and never touch their own (scattered) row.
// Both report the size of the next frontier and the degree sum of its
// vertices, which drive the direction choice.
#include "device.hpp"

using namespace gcd;

struct BfsCounters {
  uint32_t next_size;
  uint32_t pad;
  unsigned long long next_edges;
};

// The level's counters as the host reads them: written to pinned host
// memory by bfs_publish (or by the scheduler's decision kernel, which adds
// its decision), `seq` last, which the host polls instead of waiting for
// the stream and copying.
struct BfsPublished {
  BfsCounters counters;
  uint32_t decision[16];  // the decision kernel's SchedDecision
  uint32_t seq;
  uint32_t pad[3];
};

// Counter slots per set of the bottom-up runs (see block_counters_add_slot).
#define BFS_SLOTS 64

// Also clears the level counters (the BFS_SLOTS slots of the levels run one
// at a time, which every level's publication clears again for the next one,
// and the three sets of the bottom-up runs), sets start[0] = 0 (the degree prefix of the one-vertex first
// frontier, so the first level needs no scan), and sets up the frontier
// bitmaps that top-down levels keep in step with their queues: bits0 = the
// source alone, bits1 = empty (`words` 64-bit words each).
// dist is written four entries (16 bytes) per store, over a grid-stride loop.
extern "C" __global__ __launch_bounds__(256) void bfs_init(int32_t* __restrict__ dist, uint32_t n,
                                                           uint32_t source, uint32_t* __restrict__ queue,
                                                           BfsCounters* __restrict__ counters,
                                                           BfsCounters* __restrict__ bu_counters,
                                                           uint32_t* __restrict__ start,
                                                           uint64_t* __restrict__ bits0,
                                                           uint64_t* __restrict__ bits1, uint64_t words) {
  const uint64_t quads = n / 4;
  int4* d4 = reinterpret_cast<int4*>(dist);
  for (uint64_t q = global_thread(); q < quads; q += grid_threads()) {
    const uint64_t v = 4 * q;
    d4[q] = make_int4(v == source ? 0 : -1, v + 1 == source ? 0 : -1, v + 2 == source ? 0 : -1,
                      v + 3 == source ? 0 : -1);
  }
  for (uint64_t v = 4 * quads + global_thread(); v < n; v += grid_threads()) dist[v] = v == source ? 0 : -1;
  for (uint64_t w = global_thread(); w < words; w += grid_threads()) {
    bits0[w] = w == source / 64 ? 1ull << (source % 64) : 0ull;
    bits1[w] = 0ull;
  }
  if (global_thread() == 0) {
    queue[0] = source;
    start[0] = 0;
  }
  if (global_thread() < BFS_SLOTS) {
    counters[global_thread()].next_size = 0;
    counters[global_thread()].pad = 0;
    counters[global_thread()].next_edges = 0;
  }
  if (global_thread() < 3 * BFS_SLOTS) {
    bu_counters[global_thread()].next_size = 0;
    bu_counters[global_thread()].next_edges = 0;
  }
}

// Copies a finished level's counters, the sum of its `slots` counter slots,
// to `host`, clears the slots for the next level unless `keep`, then sets
// host->seq = seq after a system-scope fence, so the counters are visible
// before the sequence word. Called by one whole wave (slots <= 64).
__device__ __forceinline__ void publish_counters(BfsCounters* __restrict__ counters, uint32_t slots,
                                                 BfsPublished* __restrict__ host, uint32_t seq,
                                                 bool keep = false) {
  const uint32_t lane = lane_id();
  uint32_t c = 0;
  unsigned long long e = 0;
  if (lane < slots) {
    c = counters[lane].next_size;
    e = counters[lane].next_edges;
  }
  c = wave_sum(c);
  e = wave_sum(e);
  if (!keep && lane < slots) {
    counters[lane].next_size = 0;
    counters[lane].pad = 0;
    counters[lane].next_edges = 0;
  }
  if (lane == 0) {
    host->counters.next_size = c;
    host->counters.pad = 0;
    host->counters.next_edges = e;
    __threadfence_system();
    __hip_atomic_store(&host->seq, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Publishes the level's counters (`slots` slots) to the host (one wave).
extern "C" __global__ __launch_bounds__(64) void bfs_publish(BfsCounters* __restrict__ counters, uint32_t slots,
                                                             BfsPublished* __restrict__ host, uint32_t seq) {
  if (blockIdx.x != 0) return;
  publish_counters(counters, slots, host, seq);
}

// top[k * n + v] = the k-th last entry of row v (k < 4); past the row's start,
// v itself: a vertex that bottom-up checks is unvisited, so never in the
// frontier, and its four candidates need no bound from the row's offsets.
extern "C" __global__ __launch_bounds__(256) void bfs_top_build(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t n,
    uint32_t* __restrict__ top) {
  for (uint64_t v = global_thread(); v < n; v += grid_threads()) {
    const uint32_t b = offsets[v], e = offsets[v + 1];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      top[k * static_cast<uint64_t>(n) + v] = e - b > k ? targets[e - 1 - k] : static_cast<uint32_t>(v);
  }
}

// Adds the workgroup's per-lane counts to the level counters with one atomic
// pair per workgroup (every thread of the 256-lane workgroup calls it). One
// pair per wave and chunk serialized on the counters' line: a bottom-up level
// that finds vertices in every chunk took 880 us instead of tens.
__device__ __forceinline__ void block_counters_add(uint32_t count, unsigned long long edges,
                                                   BfsCounters* __restrict__ counters) {
  __shared__ uint32_t s_count[256 / GC_WAVE];
  __shared__ unsigned long long s_edges[256 / GC_WAVE];
  count = wave_sum(count);
  edges = wave_sum(edges);
  if (lane_id() == 0) {
    s_count[threadIdx.x / GC_WAVE] = count;
    s_edges[threadIdx.x / GC_WAVE] = edges;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t c = 0;
    unsigned long long e = 0;
    for (uint32_t w = 0; w < 256 / GC_WAVE; ++w) {
      c += s_count[w];
      e += s_edges[w];
    }
    if (c != 0) atomicAdd(&counters->next_size, c);
    if (e != 0) atomicAdd(&counters->next_edges, e);
  }
}

// As block_counters_add into slot blockIdx.x % slots of `counters`: device-
// scope atomics on one word serialize across the XCDs (a bottom-up level
// whose 2,048 workgroups all found vertices took 40-60 us more than one
// whose workgroups found none), so the bottom-up runs spread them over
// BFS_SLOTS words and their publication adds the slots up.
__device__ __forceinline__ void block_counters_add_slot(uint32_t count, unsigned long long edges,
                                                        BfsCounters* __restrict__ counters, uint32_t slots) {
  block_counters_add(count, edges, counters + blockIdx.x % slots);
}

// A wave takes two chunks of 64 vertices at a time (chunk and chunk + waves),
// their loads interleaved: dist of both, then the four candidates and row
// bounds of both, then the candidates' frontier words, so that each wave has
// twice the loads in flight along the chain (one chunk at a time, the
// level's launches ran at about 0.6 TB/s on three dependent round trips per
// chunk).
extern "C" __global__ __launch_bounds__(256) void bfs_bottomup(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ top, int32_t* __restrict__ dist, const uint32_t* __restrict__ front,
    uint64_t* __restrict__ next, uint32_t n, uint32_t first_chunk, BfsCounters* __restrict__ counters,
    int32_t next_level, BfsCounters* __restrict__ pub_counters, BfsPublished* __restrict__ pub_host,
    uint32_t pub_seq, BfsCounters* __restrict__ clear_counters, uint32_t slots) {
  // A level launched ahead publishes the counters of the level before it
  // (finished, since launches on the stream run in order): no publishing
  // launch of its own between the two levels. The runs rotate three counter
  // sets of `slots` slots each: this level counts into one, publishes the one
  // before it (left as it is, every wave reads it) and clears the one of two
  // levels back, which the next level counts into. If the level before found
  // nothing, the traversal has ended and this level returns at once.
  if (pub_host != nullptr) {
    const uint32_t lane_slot = lane_id();
    const bool any = __ballot(lane_slot < slots && pub_counters[lane_slot].next_size != 0) != 0ull;
    if (blockIdx.x == 0 && threadIdx.x < GC_WAVE) {
      publish_counters(pub_counters, slots, pub_host, pub_seq, true);
      if (clear_counters != nullptr && lane_slot < slots) {
        clear_counters[lane_slot].next_size = 0;
        clear_counters[lane_slot].next_edges = 0;
      }
    }
    if (!any) return;
  }
  const uint32_t lane = lane_id();
  uint32_t found_count = 0;  // this lane's discoveries and their degrees
  unsigned long long found_edges = 0;
  const uint64_t waves = grid_threads() / GC_WAVE;
  const uint64_t chunks = (static_cast<uint64_t>(n) + GC_WAVE - 1) / GC_WAVE;
  // Chunks below first_chunk hold only vertices of degree 0 (sorted first),
  // which no level ever reaches: their words of `next` are only cleared.
  for (uint64_t c = global_thread(); c < first_chunk; c += grid_threads()) next[c] = 0;
  for (uint64_t chunk = first_chunk + global_thread() / GC_WAVE; chunk < chunks; chunk += 2 * waves) {
    uint64_t ch[2] = {chunk, chunk + waves};
    uint32_t v[2];
    bool open[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint64_t v64 = ch[j] * GC_WAVE + lane;
      v[j] = static_cast<uint32_t>(v64 < n ? v64 : 0);
      open[j] = v64 < n;
    }
    int32_t d[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) d[j] = open[j] ? dist[v[j]] : 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) open[j] = open[j] && d[j] < 0;
    // the four candidates (the row's four last entries, from the transposed
    // copy) and the row bounds of both vertices
    uint32_t u[2][4], rb[2], re[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) u[j][k] = open[j] ? top[k * static_cast<uint64_t>(n) + v[j]] : 0u;
      rb[j] = open[j] ? offsets[v[j]] : 0u;
      re[j] = open[j] ? offsets[v[j] + 1] : 0u;
    }
    uint32_t word[2][4];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) word[j][k] = open[j] ? front[u[j][k] >> 5] : 0u;
    bool found[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      found[j] = false;
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) found[j] |= ((word[j][k] >> (u[j][k] & 31)) & 1u) != 0;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      // then four candidates per step, their loads issued together
      const uint32_t b = rb[j], e = re[j];
      if (open[j] && !found[j]) {
        for (uint32_t i = e - min(4u, e - b); i > b && !found[j];) {
          const uint32_t cnt = min(4u, i - b);
          uint32_t w[4], bits[4];
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) w[k] = k < cnt ? targets[i - 1 - k] : 0u;
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) bits[k] = k < cnt ? front[w[k] >> 5] : 0u;
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) found[j] |= k < cnt && ((bits[k] >> (w[k] & 31)) & 1u);
          i -= cnt;
        }
      }
      uint32_t deg = 0;
      if (found[j]) {
        dist[v[j]] = next_level;
        deg = e - b;
      }
      const uint64_t mask = __ballot(found[j]);
      if (lane == 0 && ch[j] < chunks) next[ch[j]] = mask;
      found_count += found[j] ? 1u : 0u;
      found_edges += deg;
    }
  }
  block_counters_add_slot(found_count, found_edges, counters, slots);
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

// For a small frontier (one workgroup's work): start[i] = exclusive prefix
// sum of the degrees of queue[0..size), and the level's counters cleared, in
// one launch instead of a fill, the degree kernel and the three scan kernels
// (about 30 us of launches and kernels per top-down level).
#define FS_THREADS 1024

extern "C" __global__ __launch_bounds__(FS_THREADS) void bfs_frontier_scan_small(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ queue, uint32_t size,
    uint32_t* __restrict__ start, BfsCounters* __restrict__ counters) {
  __shared__ uint32_t s_wave[FS_THREADS / GC_WAVE];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) {
    s_carry = 0;
    counters->next_size = 0;
    counters->pad = 0;
    counters->next_edges = 0;
  }
  __syncthreads();
  const uint32_t wave = threadIdx.x / GC_WAVE;
  for (uint32_t chunk = 0; chunk < size; chunk += FS_THREADS) {
    const uint32_t i = chunk + threadIdx.x;
    uint32_t deg = 0;
    if (i < size) {
      const uint32_t u = queue[i];
      deg = offsets[u + 1] - offsets[u];
    }
    const uint32_t incl = wave_inclusive_scan(deg);
    if (lane_id() == GC_WAVE - 1) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = s_carry;
    for (uint32_t w = 0; w < wave; ++w) before += s_wave[w];
    if (i < size) start[i] = before + incl - deg;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t total = 0;
      for (uint32_t w = 0; w < FS_THREADS / GC_WAVE; ++w) total += s_wave[w];
      s_carry += total;
    }
    __syncthreads();
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
    BfsCounters* __restrict__ counters, int32_t next_level, uint32_t* __restrict__ cur_bits,
    uint32_t* __restrict__ next_bits) {
  // With cur_bits / next_bits (the frontier's bitmap, which holds exactly the
  // queue's vertices, and an empty one), the level also clears its frontier's
  // bits and sets the bits of the vertices it discovers, so that afterwards
  // next_bits holds the next frontier and cur_bits is empty: a switch to
  // bottom-up needs no conversion. Every frontier vertex has an edge (it was
  // reached over one, and the bitmaps are kept only for symmetric graphs),
  // so the lane of its first edge clears its bit.
  // The workgroup's discoveries of one block of 4 x TD_EDGES_PER_WAVE edges are
  // gathered in LDS and appended with one atomic (one per wave and 64 edges,
  // all on one counter, serialized: a 370K-edge level took 368 us) on the
  // size of counter slot 0; the edges go to the BFS_SLOTS counter slots.
  // (Appending only while the queue held fewer than 16,384 entries, with the
  // level's size counted in the slots and larger frontiers left in next_bits,
  // was slower: the test of the queue counter before each append is one more
  // access to its contended line; 0.22 against 0.185 ms per R-MAT 22 source.)
  __shared__ uint32_t s_list[4 * TD_EDGES_PER_WAVE];
  __shared__ uint32_t s_count, s_base;
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / GC_WAVE;
  unsigned long long found_edges = 0;  // degrees of this lane's discoveries
  if (threadIdx.x == 0) s_count = 0;
  __syncthreads();
  for (uint64_t block = blockIdx.x; block * 4 * TD_EDGES_PER_WAVE < total; block += gridDim.x) {
    const uint64_t base = block * 4 * TD_EDGES_PER_WAVE + static_cast<uint64_t>(wave) * TD_EDGES_PER_WAVE;
    if (base < total) {
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
        bool discovered = false;
        uint32_t v = 0;
        if (t <= last) {
          uint32_t a = own_lo, b = own_hi;
          while (a < b) {
            const uint32_t mid = a + ((b - a + 1) >> 1);
            if (start[mid] <= t) a = mid;
            else b = mid - 1;
          }
          const uint32_t u = queue[a];
          v = targets[offsets[u] + (t - start[a])];
          if (cur_bits != nullptr && t == start[a]) atomicAnd(&cur_bits[u >> 5], ~(1u << (u & 31)));
          if (dist[v] < 0) {
            discovered = atomicCAS(&dist[v], -1, next_level) == -1;
            if (discovered) {
              found_edges += offsets[v + 1] - offsets[v];
              if (next_bits != nullptr) atomicOr(&next_bits[v >> 5], 1u << (v & 31));
            }
          }
        }
        const unsigned long long mask = __ballot(discovered);
        if (mask != 0) {
          uint32_t pos = 0;
          if (lane == 0) pos = atomicAdd(&s_count, static_cast<uint32_t>(__popcll(mask)));
          pos = __shfl(pos, 0);
          if (discovered) s_list[pos + lanes_below(mask)] = v;
        }
      }
    }
    __syncthreads();
    const uint32_t found = s_count;
    if (threadIdx.x == 0 && found != 0) s_base = atomicAdd(&counters->next_size, found);
    __syncthreads();
    if (next_queue != nullptr)
      for (uint32_t i = threadIdx.x; i < found; i += blockDim.x) next_queue[s_base + i] = s_list[i];
    __syncthreads();
    if (threadIdx.x == 0) s_count = 0;
    __syncthreads();
  }
  block_counters_add_slot(0u, found_edges, counters, BFS_SLOTS);
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

// ---- device-chained levels (a captured graph of many levels) -------------------
//
// A HIP graph holds a fixed number of levels, each = bfs_chain_expand (push
// from the frontier queue named in the graph into the other one) followed by
// bfs_chain_advance (one lane: the level's counts become the next frontier).
// Both read the frontier size and the stop flag from device memory, so once a
// level ends the traversal (no new vertex) or hands it back to the host (more
// than `cap` vertices or `edge_cap` edges), the remaining nodes of the graph
// return at once. The host replays the graph and reads the state once per
// replay instead of once per level. The expansion walks lists of degree <=
// CH_LANE_DEGREE with one lane each and longer lists with the whole wave, and
// appends discoveries with one atomic per wave.

#define CH_LANE_DEGREE 32

// Appends the lanes' discoveries with one atomic per wave (every lane calls it).
__device__ __forceinline__ void chain_append(bool found, uint32_t w, const uint32_t* __restrict__ offsets,
                                             uint32_t* __restrict__ nxt, uint32_t* count,
                                             unsigned long long* edges) {
  const unsigned long long mask = __ballot(found);
  if (mask == 0) return;
  const unsigned long long e = wave_sum(found ? static_cast<unsigned long long>(offsets[w + 1] - offsets[w]) : 0ull);
  uint32_t base = 0;
  if (lane_id() == 0) {
    base = atomicAdd(count, static_cast<uint32_t>(__popcll(mask)));
    atomicAdd(edges, e);
  }
  base = __shfl(base, 0);
  if (found) nxt[base + lanes_below(mask)] = w;
}

struct BfsChainState {
  uint32_t size;        // current frontier
  int32_t level;        // its level
  uint32_t stop;        // 0 running, 1 finished (empty level), 2 frontier too large
  uint32_t levels;      // levels built since the host last set the state
  uint32_t next_size;
  uint32_t cap;
  unsigned long long next_edges;
  unsigned long long edges;      // degree sum of the current frontier
  unsigned long long all_edges;  // degree sums of every level built
  unsigned long long reached;    // vertices discovered
  unsigned long long edge_cap;
  // Hand-back tests of the fixed rule's chains (0 = off): a frontier of more
  // than avg_cap edges per vertex (the lists a lane or a wave walks alone),
  // and Beamer's switch to bottom-up, next edges * alpha > unexplored edges.
  unsigned long long avg_cap;
  unsigned long long unexplored;  // edges of the vertices not yet reached
  double alpha;
};

extern "C" __global__ __launch_bounds__(256) void bfs_chain_expand(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    int32_t* __restrict__ dist, const uint32_t* __restrict__ cur, uint32_t* __restrict__ nxt,
    BfsChainState* __restrict__ st) {
  if (st->stop != 0) return;
  const uint32_t size = st->size;
  const int32_t next_level = st->level + 1;
  const uint32_t lane = lane_id();
  const uint64_t wave = global_thread() / GC_WAVE;
  const uint64_t waves = grid_threads() / GC_WAVE;
  for (uint64_t base = wave * GC_WAVE; base < size; base += waves * GC_WAVE) {
    const uint64_t q = base + lane;
    uint32_t a = 0, e = 0;
    if (q < size) {
      const uint32_t v = cur[q];
      a = offsets[v];
      e = offsets[v + 1];
    }
    const bool long_list = e - a > CH_LANE_DEGREE;
    const uint32_t my_deg = long_list ? 0u : e - a;
    uint32_t max_deg = my_deg;
#pragma unroll
    for (int o = GC_WAVE / 2; o > 0; o >>= 1) max_deg = max(max_deg, static_cast<uint32_t>(__shfl_xor(max_deg, o)));
    for (uint32_t k = 0; k < max_deg; ++k) {
      bool found = false;
      uint32_t w = 0;
      if (k < my_deg) {
        w = targets[a + k];
        found = dist[w] == -1 && atomicCAS(&dist[w], -1, next_level) == -1;
      }
      chain_append(found, w, offsets, nxt, &st->next_size, &st->next_edges);
    }
    unsigned long long longs = __ballot(long_list);
    while (longs) {
      const int j = __builtin_ctzll(longs);
      longs &= longs - 1;
      const uint32_t ja = __shfl(a, j), je = __shfl(e, j);
      for (uint32_t i0 = ja; i0 < je; i0 += GC_WAVE) {
        const uint32_t i = i0 + lane;
        bool found = false;
        uint32_t w = 0;
        if (i < je) {
          w = targets[i];
          found = dist[w] == -1 && atomicCAS(&dist[w], -1, next_level) == -1;
        }
        chain_append(found, w, offsets, nxt, &st->next_size, &st->next_edges);
      }
    }
  }
}

extern "C" __global__ __launch_bounds__(64) void bfs_chain_advance(BfsChainState* __restrict__ st) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || st->stop != 0) return;
  const uint32_t s = st->next_size;
  const unsigned long long e = st->next_edges;
  st->next_size = 0;
  st->next_edges = 0;
  st->levels += 1;
  st->level += 1;
  st->reached += s;
  st->all_edges += e;
  st->size = s;
  st->edges = e;
  st->unexplored -= min(st->unexplored, e);
  if (s == 0) st->stop = 1;
  else if (s > st->cap || e > st->edge_cap) st->stop = 2;
  else if (st->avg_cap != 0 && e > st->avg_cap * s) st->stop = 2;
  else if (st->alpha > 0.0 && static_cast<double>(e) * st->alpha > static_cast<double>(st->unexplored)) st->stop = 2;
}
