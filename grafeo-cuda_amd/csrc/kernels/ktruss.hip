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

// The row loops of the index kernels below: one row per wave (grid-stride
// over the waves), its entries over the lanes, so a hub row of 1e5 entries is
// not one thread's serial loop (one thread per row took 11.5 ms at R-MAT scale
// 22, the hubs' threads running alone at the end).
__device__ __forceinline__ uint64_t kt_wave() { return global_thread() / GC_WAVE; }
__device__ __forceinline__ uint64_t kt_waves() { return grid_threads() / GC_WAVE; }

// rev[i] = position of u in row v for every forward entry i = (u -> v).
extern "C" __global__ __launch_bounds__(256) void kt_reverse_index(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ fwd, uint32_t n, uint32_t* __restrict__ rev) {
  for (uint64_t uu = kt_wave(); uu < n; uu += kt_waves()) {
    const uint32_t u = static_cast<uint32_t>(uu);
    const uint32_t e = offsets[u + 1];
    for (uint32_t i = fwd[u] + lane_id(); i < e; i += GC_WAVE) {
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

// ---- support on the hub bitmap (kt_bitmap_support) ------------------------
//
// The same supports as kt_wedge_support, from the walk of tc_bitmap (tc.hip)
// over the 32-bit lists: a wave holds L(u) ∩ [hot_base, n) of row u as a
// 32768-bit LDS bitmap (and a Bloom filter of its cold prefix) and walks
// L(v) for every v in L(u); a hit w is the triangle u < v < w. Its three
// forward entries gain one support each:
//  - v->w is the probed entry itself (one global atomic per triangle, at a
//    position the walk already holds);
//  - u->v takes the list's hits at once (one global atomic per list);
//  - u->w is counted in a 16-bit LDS counter per hot entry of L(u) (the
//    entry's rank: the rank of the first entry in w's bitmap word, kept per
//    word, plus the bits below w's in that word), added to the support once
//    per row; rows with more than KTB_COUNTERS hot entries, and cold w (found
//    by a search of the prefix), take a global atomic.
// The wedge enumeration probes all 7.2e9 DAG wedges at R-MAT scale 22 (295
// ms); the bitmap walk tests 2.9e10 list entries, but at one LDS read each.
// Rows and parts are scheduled as in tc_bitmap (heavy rows one at a time,
// the others in chunks; rows of at least split_min entries walked once per
// middle-vertex part, each part's counters on a line of their own).
#define KTB_WAVES 4
#define KTB_HOT_BITS 32768
#define KTB_HOT_WORDS (KTB_HOT_BITS / 32)
#define KTB_BLOOM_LOG2 12
#define KTB_BLOOM_WORDS ((1 << KTB_BLOOM_LOG2) / 32)
#define KTB_BLOOM_BASE (KTB_HOT_WORDS + 1)  // after the bitmap's zero word
#define KTB_BITS_WORDS (KTB_BLOOM_BASE + KTB_BLOOM_WORDS)
#define KTB_COUNTERS 2048  // 16-bit u->w counters per wave, two per word
#define KTB_COUNTER_STRIDE 64

struct KtbLds {
  uint32_t bits[KTB_BITS_WORDS];
  uint16_t wrank[KTB_HOT_WORDS];  // rank among L(u)'s hot entries of the first one in each word
  uint32_t cnt[KTB_COUNTERS / 2];
};

__device__ __forceinline__ uint32_t ktb_bloom_hash(uint32_t w) {
  return (w * 0x9E3779B1u) >> (32 - KTB_BLOOM_LOG2);
}

// Hits of entries [s1, s1 + lj) of targets (the list L(v)) among the row's
// sets, with their supports: v->w at the entry, u->w in the row's counters
// (or at A's entry, base + position), per lane.
__device__ __forceinline__ uint32_t ktb_walk(const uint32_t* __restrict__ targets, uint32_t s1, uint32_t lj,
                                             KtbLds& L, uint32_t hot_base, const uint32_t* __restrict__ A,
                                             uint32_t c, uint32_t row_e, bool lds_cnt,
                                             uint32_t* __restrict__ support) {
  const uint32_t lane = lane_id();
  uint32_t local = 0;
  for (uint32_t kk = 0; kk < lj; kk += 4 * GC_WAVE) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = targets[s1 + min(kk + q * GC_WAVE + lane, lj - 1)];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t idx = kk + q * GC_WAVE + lane;
      const uint32_t y = w[q] - hot_base;
      const uint32_t word = L.bits[min(y >> 5, static_cast<uint32_t>(KTB_HOT_WORDS))];
      const uint32_t bit = y & 31u;
      if (idx < lj && ((word >> bit) & 1u)) {
        atomicAdd(&support[s1 + idx], 1u);
        const uint32_t r = L.wrank[y >> 5] + static_cast<uint32_t>(__popc(word & ((1u << bit) - 1u)));
        if (lds_cnt) atomicAdd(&L.cnt[r >> 1], 1u << (16u * (r & 1u)));
        else atomicAdd(&support[row_e + c + r], 1u);
        ++local;
      } else if (c != 0 && idx < lj && w[q] < hot_base) {
        const uint32_t h = ktb_bloom_hash(w[q]);
        if ((L.bits[KTB_BLOOM_BASE + (h >> 5)] >> (h & 31u)) & 1u) {
          const uint32_t p = lower_bound_u32(A, 0, c, w[q]);
          if (p < c && A[p] == w[q]) {
            atomicAdd(&support[s1 + idx], 1u);
            atomicAdd(&support[row_e + p], 1u);
            ++local;
          }
        }
      }
    }
  }
  return local;
}

// lo[x] .. hi[x]: the forward entries of row x (the relabelled CSR's fwd and
// offsets + 1); rows [0, n), taken from the top. support[e] (zeroed) gains
// the triangles of every forward entry e. counters: parts x
// KTB_COUNTER_STRIDE zeroed words. The edges u->v whose bit is set in
// midbits (null: none) are left to kt_mid_support.
extern "C" __global__ __launch_bounds__(KTB_WAVES * GC_WAVE) void kt_bitmap_support(
    const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi, const uint32_t* __restrict__ targets,
    uint32_t n, uint32_t hot_base, uint32_t heavy, uint32_t light_chunk, uint32_t parts, uint32_t split_min,
    uint32_t* __restrict__ counters, uint32_t* __restrict__ support, const uint32_t* __restrict__ midbits) {
  __shared__ KtbLds s_all[KTB_WAVES];
  KtbLds& L = s_all[threadIdx.x / GC_WAVE];
  const uint32_t lane = lane_id();
  for (uint32_t j = lane; j < KTB_BITS_WORDS; j += GC_WAVE) L.bits[j] = 0;
  for (uint32_t j = lane; j < KTB_COUNTERS / 2; j += GC_WAVE) L.cnt[j] = 0;
  wave_lds_sync();
  const uint32_t home = blockIdx.x % parts;
  for (uint32_t qp = 0; qp < parts; ++qp) {
    const uint32_t part = (home + qp) % parts;
    uint32_t* part_counters = counters + KTB_COUNTER_STRIDE * part;
    bool heavy_left = heavy > 0;
    for (;;) {
      uint32_t kb = n, ke = n;
      if (lane == 0) {
        if (heavy_left) {
          const uint32_t k = atomicAdd(&part_counters[0], 1u);
          if (k < heavy) {
            kb = k;
            ke = k + 1;
          } else {
            heavy_left = false;
          }
        }
        if (!heavy_left && kb == n) {
          const uint32_t cc = atomicAdd(&part_counters[1], light_chunk);
          const uint64_t b = static_cast<uint64_t>(heavy) + cc;
          if (b < n) {
            kb = static_cast<uint32_t>(b);
            const uint64_t e = b + light_chunk;
            ke = e < n ? static_cast<uint32_t>(e) : n;
          }
        }
      }
      kb = __shfl(kb, 0);
      ke = __shfl(ke, 0);
      if (kb >= n) break;
      for (uint32_t kc = kb; kc < ke; kc += GC_WAVE) {
        uint32_t row_lo_e = 0, row_a = 0;
        bool take = false;
        if (kc + lane < ke) {
          const uint32_t u = n - 1 - (kc + lane);
          row_lo_e = lo[u];
          row_a = hi[u] - row_lo_e;
          take = row_a != 0 && (row_a >= split_min || u % parts == part);
        }
        unsigned long long rows = __ballot(take);
        while (rows != 0) {
          const int jr = __ffsll(static_cast<long long>(rows)) - 1;
          rows &= rows - 1;
          const uint32_t a = static_cast<uint32_t>(__builtin_amdgcn_readlane(row_a, jr));
          const uint32_t row_e = static_cast<uint32_t>(__builtin_amdgcn_readlane(row_lo_e, jr));
          const uint32_t* A = targets + row_e;
          const bool whole = a < split_min;
          // the cold prefix [0, c) (entries below hot_base) and the hot suffix's bits and ranks
          const uint32_t c = __builtin_amdgcn_readfirstlane(lower_bound_u32(A, 0, a, hot_base));
          for (uint32_t j = c + lane; j < a; j += GC_WAVE) {
            const uint32_t y = A[j] - hot_base;
            atomicOr(&L.bits[y >> 5], 1u << (y & 31u));
            if (j == c || ((A[j - 1] - hot_base) >> 5) != (y >> 5)) L.wrank[y >> 5] = static_cast<uint16_t>(j - c);
          }
          for (uint32_t j = lane; j < c; j += GC_WAVE) {
            const uint32_t h = ktb_bloom_hash(A[j]);
            atomicOr(&L.bits[KTB_BLOOM_BASE + (h >> 5)], 1u << (h & 31u));
          }
          wave_lds_sync();
          const bool lds_cnt = a - c <= KTB_COUNTERS;
          for (uint32_t cv = 0; cv < a; cv += GC_WAVE) {
            const uint32_t nv = min(static_cast<uint32_t>(GC_WAVE), a - cv);
            uint32_t st = 0, len = 0;
            if (lane < nv) {
              const uint32_t v = A[cv + lane];
              // an entry u->v whose bit is set in midbits is walked by kt_mid_support at v
              const uint32_t e = row_e + cv + lane;
              const bool mid = midbits != nullptr && ((midbits[e >> 5] >> (e & 31u)) & 1u) != 0;
              if (!mid && (whole || v % parts == part)) {
                st = lo[v];
                len = hi[v] - st;
              }
            }
            unsigned long long pending = __ballot(len != 0);
            while (pending != 0) {
              const int jv = __ffsll(static_cast<long long>(pending)) - 1;
              pending &= pending - 1;
              const uint32_t s1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(st, jv));
              const uint32_t lj = static_cast<uint32_t>(__builtin_amdgcn_readlane(len, jv));
              const uint32_t hits = wave_sum(ktb_walk(targets, s1, lj, L, hot_base, A, c, row_e, lds_cnt, support));
              if (lane == 0 && hits != 0) atomicAdd(&support[row_e + cv + static_cast<uint32_t>(jv)], hits);
            }
          }
          wave_lds_sync();
          if (lds_cnt) {
            for (uint32_t r = lane; r < a - c; r += GC_WAVE) {
              const uint32_t x = (L.cnt[r >> 1] >> (16u * (r & 1u))) & 0xFFFFu;
              if (x != 0) atomicAdd(&support[row_e + c + r], x);
            }
          }
          wave_lds_sync();
          if (lds_cnt) {
            for (uint32_t j = lane; j < (a - c + 1) / 2; j += GC_WAVE) L.cnt[j] = 0;
          }
          for (uint32_t j = c + lane; j < a; j += GC_WAVE) L.bits[(A[j] - hot_base) >> 5] = 0;
          for (uint32_t j = lane; j < c; j += GC_WAVE) L.bits[KTB_BLOOM_BASE + (ktb_bloom_hash(A[j]) >> 5)] = 0;
          wave_lds_sync();
        }
      }
    }
  }
}

// ---- the middle side of the supports (kt_mid_support) ----
//
// As in the triangle count (tc_mid16), a DAG edge u->v with v among the hot
// ids can be walked from v: the entries of L(u) above v against the bitmap
// of L(v), whose entries are all hot. Each hit w is the triangle u < v < w:
// u->w is the walked entry itself (a global atomic), v->w a 16-bit LDS counter
// per entry of L(v) (its rank, as kt_bitmap_support ranks L(u)'s entries),
// added once per unit, and u->v takes the walk's hits at once. Which side
// walks an edge is decided once per graph (kt_mid_records): the middle side
// when the walk, plus `ovh` entries, is shorter than L(v), or empty.

// Records of the entries of rows v in [hot_base, n) (rec[i - offsets[hot_base]]):
// {first entry, entries} of the walk of L(u) above v, or {0, 0}; sets
// midbits and walks[v - hot_base] = 1 for the rows with a non-empty walk.
extern "C" __global__ __launch_bounds__(256) void kt_mid_records(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, const uint32_t* __restrict__ fwd,
    uint32_t n, uint32_t hot_base, uint32_t ovh, uint2* __restrict__ rec, uint32_t* __restrict__ midbits,
    uint8_t* __restrict__ walks) {
  const uint32_t rec_base = offsets[hot_base];
  const uint64_t total = static_cast<uint64_t>(offsets[n]) - rec_base;
  for (uint64_t k = global_thread(); k < total; k += grid_threads()) {
    const uint32_t i = rec_base + static_cast<uint32_t>(k);
    uint32_t lo = hot_base, hi = n - 1;  // the row of entry i
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo + 1) >> 1);
      if (offsets[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const uint32_t v = lo;
    uint2 r = make_uint2(0u, 0u);
    const uint32_t u = targets[i];
    if (u < v) {
      const uint32_t ue = offsets[u + 1];
      const uint32_t j = lower_bound_u32(targets, fwd[u], ue, v);
      if (j < ue && targets[j] == v) {
        const uint32_t walk = ue - j - 1;
        if (walk == 0 || walk + ovh < offsets[v + 1] - fwd[v]) {
          r = make_uint2(j + 1, walk);
          atomicOr(&midbits[j >> 5], 1u << (j & 31u));
          if (walk != 0) walks[v - hot_base] = 1;
        }
      }
    }
    rec[k] = r;
  }
}

// Units {v, first entry, end entry, 0} of the hot rows' in-lists, one per
// atomic on `counter`; per unit the wave sets L(v)'s bits and word ranks,
// walks the records' walks a quarter wave each (four entries per lane and
// step), then adds the v->w counters and clears its sets.
extern "C" __global__ __launch_bounds__(GC_WAVE) void kt_mid_support(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, const uint32_t* __restrict__ fwd,
    uint32_t hot_base, const uint2* __restrict__ rec, uint32_t rec_base, const uint4* __restrict__ units,
    uint32_t num_units, uint32_t* __restrict__ counter, uint32_t* __restrict__ support) {
  __shared__ uint32_t s_bits[KTB_HOT_WORDS + 1];
  __shared__ uint16_t s_wrank[KTB_HOT_WORDS];
  __shared__ uint32_t s_cnt[KTB_COUNTERS / 2];
  const uint32_t lane = lane_id();
  const uint32_t quarter = lane >> 4, ql = lane & 15u;
  for (uint32_t j = lane; j < KTB_HOT_WORDS + 1; j += GC_WAVE) s_bits[j] = 0;
  for (uint32_t j = lane; j < KTB_COUNTERS / 2; j += GC_WAVE) s_cnt[j] = 0;
  wave_lds_sync();
  for (;;) {
    uint32_t k = 0;
    if (lane == 0) k = atomicAdd(counter, 1u);
    k = __builtin_amdgcn_readfirstlane(__shfl(k, 0));
    if (k >= num_units) break;
    const uint4 unit = units[k];
    const uint32_t v = __builtin_amdgcn_readfirstlane(unit.x);
    const uint32_t b = __builtin_amdgcn_readfirstlane(unit.y), e = __builtin_amdgcn_readfirstlane(unit.z);
    const uint32_t lb = fwd[v], le = offsets[v + 1];
    const uint32_t a = le - lb;  // L(v): all hot (v >= hot_base)
    const bool lds_cnt = a <= KTB_COUNTERS;
    for (uint32_t j = lb + lane; j < le; j += GC_WAVE) {
      const uint32_t y = targets[j] - hot_base;
      atomicOr(&s_bits[y >> 5], 1u << (y & 31u));
      if (j == lb || ((targets[j - 1] - hot_base) >> 5) != (y >> 5)) s_wrank[y >> 5] = static_cast<uint16_t>(j - lb);
    }
    wave_lds_sync();
    for (uint32_t i0 = b; i0 < e; i0 += GC_WAVE) {
      const uint32_t i = i0 + lane;
      const uint2 r = i < e ? rec[i - rec_base] : make_uint2(0u, 0u);
      unsigned long long lists = __ballot(r.y != 0);
      uint32_t cur = 0, end = 0, edge = 0, hits = 0;  // the quarter's walk, its edge u->v, its hits
      for (;;) {
        const unsigned long long idle = __ballot(cur >= end);
        uint32_t free4 = static_cast<uint32_t>((idle & 1ull) | ((idle >> 15) & 2ull) | ((idle >> 30) & 4ull) |
                                               ((idle >> 45) & 8ull));
        while (free4 != 0 && lists != 0) {
          const uint32_t fq = static_cast<uint32_t>(__ffs(free4) - 1);
          free4 &= free4 - 1;
          const int jv = __ffsll(static_cast<long long>(lists)) - 1;
          lists &= lists - 1;
          const uint32_t s1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(r.x, jv));
          const uint32_t l1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(r.y, jv));
          if (quarter == fq) {
            cur = s1;
            end = s1 + l1;
            edge = s1 - 1;
            hits = 0;
          }
        }
        if (__ballot(cur < end) == 0) break;
        uint32_t w[4];
        bool act[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          const uint32_t idx = cur + 16u * d + ql;
          act[d] = idx < end;
          w[d] = targets[act[d] ? idx : (end > 0 ? end - 1 : 0u)];
        }
        uint32_t h = 0;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          const uint32_t y = w[d] - hot_base;  // w > v >= hot_base for a walked entry
          const uint32_t word = s_bits[min(y >> 5, static_cast<uint32_t>(KTB_HOT_WORDS))];
          const uint32_t bit = y & 31u;
          if (act[d] && ((word >> bit) & 1u)) {
            atomicAdd(&support[cur + 16u * d + ql], 1u);  // u->w
            const uint32_t rk = s_wrank[y >> 5] + static_cast<uint32_t>(__popc(word & ((1u << bit) - 1u)));
            if (lds_cnt) atomicAdd(&s_cnt[rk >> 1], 1u << (16u * (rk & 1u)));  // v->w
            else atomicAdd(&support[lb + rk], 1u);
            ++h;
          }
        }
        // the quarter's hits this step (lanes of a quarter hold the same walk)
        h += __shfl_xor(h, 8);
        h += __shfl_xor(h, 4);
        h += __shfl_xor(h, 2);
        h += __shfl_xor(h, 1);
        const bool had = cur < end;
        hits += had ? h : 0u;
        cur += 64u;
        if (had && cur >= end && ql == 0 && hits != 0) atomicAdd(&support[edge], hits);  // u->v
      }
    }
    wave_lds_sync();
    if (lds_cnt) {
      for (uint32_t rk = lane; rk < a; rk += GC_WAVE) {
        const uint32_t x = (s_cnt[rk >> 1] >> (16u * (rk & 1u))) & 0xFFFFu;
        if (x != 0) atomicAdd(&support[lb + rk], x);
      }
    }
    wave_lds_sync();
    if (lds_cnt) {
      for (uint32_t j = lane; j < (a + 1) / 2; j += GC_WAVE) s_cnt[j] = 0;
    }
    for (uint32_t j = lb + lane; j < le; j += GC_WAVE) s_bits[(targets[j] - hot_base) >> 5] = 0;
    wave_lds_sync();
  }
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
  for (uint64_t uu = kt_wave(); uu < n; uu += kt_waves()) {
    const uint32_t u = static_cast<uint32_t>(uu);
    const uint32_t f = fwd[u], e = offsets[u + 1];
    for (uint32_t i = offsets[u] + lane_id(); i < e; i += GC_WAVE) {
      canon[i] = i;  // entries below f: overwritten by kt_canonical_reverse for non-loops
      state[i] = static_cast<uint8_t>(i < f ? KT_NONE : KT_ALIVE);
    }
  }
}

// Second pass (after kt_canonical_index has finished everywhere): the
// reverse entry of each forward entry points at it.
extern "C" __global__ __launch_bounds__(256) void kt_canonical_reverse(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ fwd,
    const uint32_t* __restrict__ rev, uint32_t n, uint32_t* __restrict__ canon) {
  for (uint64_t uu = kt_wave(); uu < n; uu += kt_waves()) {
    const uint32_t u = static_cast<uint32_t>(uu);
    const uint32_t e = offsets[u + 1];
    for (uint32_t i = fwd[u] + lane_id(); i < e; i += GC_WAVE) canon[rev[i]] = i;
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

// One batch: `width` lanes (16 or 64) per frontier edge (u, v) walk the
// shorter of the two full rows and binary-search each neighbour w in the
// longer one. Large batches take 16 lanes per edge: the shorter row is short
// for most of their edges, which a whole wave per edge left mostly idle (32
// against 14 ms for the 8 rounds of the 4-truss at R-MAT scale 22); the few
// edges of the late trussness levels, between hubs, take a whole wave.
extern "C" __global__ __launch_bounds__(256) void kt_truss_peel(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ canon, uint32_t n, const uint32_t* __restrict__ queue,
    uint32_t queue_size, uint8_t* __restrict__ state, uint32_t* __restrict__ support, uint32_t t,
    uint32_t* __restrict__ next_queue, uint32_t* __restrict__ next_count, uint32_t width) {
  const uint32_t ql = lane_id() & (width - 1u);
  const uint64_t group = global_thread() / width;
  const uint64_t groups = grid_threads() / width;
  for (uint64_t q = group; q < queue_size; q += groups) {
    const uint32_t e = queue[q];
    const uint32_t u = row_of_entry(offsets, n, e);
    const uint32_t v = targets[e];
    const uint32_t du = offsets[u + 1] - offsets[u], dv = offsets[v + 1] - offsets[v];
    const uint32_t s = du <= dv ? u : v;
    const uint32_t l = du <= dv ? v : u;
    const uint32_t s_begin = offsets[s], s_end = offsets[s + 1];
    const uint32_t l_begin = offsets[l], l_end = offsets[l + 1];
    for (uint32_t es = s_begin + ql; es < s_end; es += width) {
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
