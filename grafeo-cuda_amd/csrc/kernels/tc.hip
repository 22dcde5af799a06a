// Triangle counting by sorted-list intersection, one 64-lane wave per vertex.
//
// Counts, over every vertex u and every v in L(u), the w in L(v) that are also
// in L(u), where L(x) = targets[lo[x] .. hi[x]). The host code chooses L:
//  - triangle_count (crates/grafeo-cuda-algos/kernels/triangle_counting.cu:246-315
//    counts w > v > u with v, w in N(u) and w in N(v)): L(x) is the suffix of
//    N(x) above x, and `after_only` restricts the search to the part of L(u)
//    above v (w > v holds for every w in L(v)). On a symmetric graph the
//    driver passes the degree-ordered relabelled CSR, where this suffix is
//    the degree-oriented DAG;
//  - triangle_count_oriented (triangle_counting.cu:331-376, the sum over
//    entries u->v of |N(u) intersect N(v)|): L(x) = N(x), `after_only` = 0.
//
// Work layout (CDNA4): a workgroup is one wave. The wave stages L(u) in LDS
// (up to TC_LDS_CAP entries, else it searches global memory), then walks the
// lists L(v) of 64 neighbours v at a time as one flattened item range: a wave
// prefix sum over their lengths spreads the items of short and long lists
// evenly over the 64 lanes, every lane reads consecutive list entries
// (coalesced), and each item is a binary search in the LDS copy of L(u).
// Vertices are handed out dynamically: the first `heavy` positions one at a
// time, the rest in chunks of `light_chunk`, so hubs (processed first when
// `descending`) do not leave a tail. Counts are u64 per wave, added with one
// atomic per wave at the end (integer sums are exact in any order).
//
// tc_count, tc_hash and tc_bitmap take the rows [row_lo, n) only, so that a
// driver can give each degree bucket (a contiguous row range of the
// degree-ordered layout) its own method; tc_bucket_stats sums each bucket's
// work for the scheduler's features.
#include "device.hpp"
#include "wedge.hpp"

using namespace gcd;

#define TC_LDS_CAP 1024

extern "C" __global__ __launch_bounds__(64) void tc_count(
    const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi,
    const uint32_t* __restrict__ targets, uint32_t n, uint32_t row_lo, uint32_t descending,
    uint32_t after_only, uint32_t heavy, uint32_t light_chunk, uint32_t* __restrict__ counters,
    unsigned long long* __restrict__ result) {
  const uint32_t positions = n - row_lo;  // rows [row_lo, n)
  __shared__ uint32_t s_a[TC_LDS_CAP];
  __shared__ uint32_t s_pos[GC_WAVE];
  __shared__ uint32_t s_start[GC_WAVE];
  __shared__ uint32_t s_range[2];

  const uint32_t lane = threadIdx.x;
  unsigned long long count = 0;
  bool heavy_left = heavy > 0;

  for (;;) {
    // ---- fetch the next range of positions [kb, ke) ----
    if (lane == 0) {
      uint32_t kb = positions, ke = positions;
      if (heavy_left) {
        uint32_t k = atomicAdd(&counters[0], 1u);
        if (k < heavy) {
          kb = k;
          ke = k + 1;
        } else {
          heavy_left = false;
        }
      }
      if (!heavy_left && kb == positions) {
        uint32_t c = atomicAdd(&counters[1], light_chunk);
        uint64_t b = static_cast<uint64_t>(heavy) + c;
        if (b < positions) {
          kb = static_cast<uint32_t>(b);
          uint64_t e = b + light_chunk;
          ke = e < positions ? static_cast<uint32_t>(e) : positions;
        }
      }
      s_range[0] = kb;
      s_range[1] = ke;
    }
    __syncthreads();
    const uint32_t kb = s_range[0], ke = s_range[1];
    heavy_left = __shfl(heavy_left ? 1 : 0, 0) != 0;
    __syncthreads();
    if (kb >= positions) break;

    for (uint32_t k = kb; k < ke; ++k) {
      const uint32_t u = descending ? n - 1 - k : row_lo + k;
      const uint32_t a_lo = lo[u];
      const uint32_t a = hi[u] - a_lo;
      if (a == 0) continue;
      const uint32_t* A = targets + a_lo;
      const bool in_lds = a <= TC_LDS_CAP;
      if (in_lds) {
        for (uint32_t j = lane; j < a; j += GC_WAVE) s_a[j] = A[j];
      }
      __syncthreads();

      for (uint32_t c = 0; c < a; c += GC_WAVE) {
        const uint32_t nv = min(static_cast<uint32_t>(GC_WAVE), a - c);
        uint32_t len = 0, st = 0;
        if (lane < nv) {
          const uint32_t v = in_lds ? s_a[c + lane] : A[c + lane];
          st = lo[v];
          len = hi[v] - st;
        }
        const uint32_t incl = wave_inclusive_scan(len);
        s_pos[lane] = incl;
        s_start[lane] = st;
        const uint32_t total = __shfl(incl, GC_WAVE - 1);
        __syncthreads();

        for (uint32_t t = lane; t < total; t += GC_WAVE) {
          // owner j: first lane whose inclusive end exceeds t
          uint32_t l = 0, h = nv - 1;
          while (l < h) {
            uint32_t mid = (l + h) >> 1;
            if (s_pos[mid] > t) h = mid;
            else l = mid + 1;
          }
          const uint32_t excl = l ? s_pos[l - 1] : 0;
          const uint32_t w = targets[s_start[l] + (t - excl)];
          uint32_t from = after_only ? c + l + 1 : 0;
          uint32_t to = a;
          bool found;
          if (in_lds) {
            while (from < to) {
              uint32_t mid = (from + to) >> 1;
              if (s_a[mid] < w) from = mid + 1;
              else to = mid;
            }
            found = from < a && s_a[from] == w;
          } else {
            while (from < to) {
              uint32_t mid = (from + to) >> 1;
              if (A[mid] < w) from = mid + 1;
              else to = mid;
            }
            found = from < a && A[from] == w;
          }
          count += found ? 1ull : 0ull;
        }
        __syncthreads();
      }
    }
  }

  count = wave_sum(count);
  if (lane == 0 && count != 0) atomicAdd(result, count);
}

// ---- hash-probe variant ------------------------------------------------------
//
// Same count as tc_count on the degree-oriented DAG (L = the suffix above the
// row), with each wave of a 256-lane workgroup working on its own vertex u:
// the wave inserts L(u) into a private LDS hash table (open addressing,
// linear probing, at most half full), then streams the lists L(v) of its
// neighbours v one after the other, 64 consecutive entries per step, and
// tests each entry with one hash probe. The work is instruction-bound, so the
// per-item cost is what counts: no owner search and no binary search. The
// lists L(v) of the DAG's high-rank vertices are short and shared by many u,
// so they stay cache-resident. Rows with more than TCH_TABLE/2 entries fall
// back to a binary search in global memory.

#define TCH_WAVES 4
#define TCH_TABLE 1024
#define TCH_ILP 2

extern "C" __global__ __launch_bounds__(TCH_WAVES * GC_WAVE) void tc_hash(
    const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi,
    const uint32_t* __restrict__ targets, uint32_t n, uint32_t row_lo, uint32_t heavy,
    uint32_t light_chunk,
    uint32_t* __restrict__ counters, unsigned long long* __restrict__ result) {
  const uint32_t positions = n - row_lo;  // rows [row_lo, n)
  __shared__ uint32_t s_table_all[TCH_WAVES][TCH_TABLE];
  const uint32_t wave = threadIdx.x / GC_WAVE;
  const uint32_t lane = lane_id();
  uint32_t* s_table = s_table_all[wave];
  unsigned long long count = 0;
  bool heavy_left = heavy > 0;

  for (;;) {
    uint32_t kb = positions, ke = positions;
    if (lane == 0) {
      if (heavy_left) {
        const uint32_t k = atomicAdd(&counters[0], 1u);
        if (k < heavy) {
          kb = k;
          ke = k + 1;
        } else {
          heavy_left = false;
        }
      }
      if (!heavy_left && kb == positions) {
        const uint32_t c = atomicAdd(&counters[1], light_chunk);
        const uint64_t b = static_cast<uint64_t>(heavy) + c;
        if (b < positions) {
          kb = static_cast<uint32_t>(b);
          const uint64_t e = b + light_chunk;
          ke = e < positions ? static_cast<uint32_t>(e) : positions;
        }
      }
    }
    kb = __shfl(kb, 0);
    ke = __shfl(ke, 0);
    if (kb >= positions) break;

    for (uint32_t k = kb; k < ke; ++k) {
      const uint32_t u = n - 1 - k;
      const uint32_t a_lo = lo[u];
      const uint32_t a = hi[u] - a_lo;
      if (a == 0) continue;
      const uint32_t* A = targets + a_lo;
      const bool in_table = 2 * a <= TCH_TABLE;
      uint32_t bits = 6;
      while ((1u << bits) < 2 * a && bits < 10) ++bits;
      const uint32_t shift = 32u - bits, mask = (1u << bits) - 1;
      if (in_table) {
        for (uint32_t j = lane; j <= mask; j += GC_WAVE) s_table[j] = 0xFFFFFFFFu;
        wave_lds_sync();
        for (uint32_t j = lane; j < a; j += GC_WAVE) {
          const uint32_t w = A[j];
          uint32_t slot = (w * 0x9E3779B1u) >> shift;
          while (atomicCAS(&s_table[slot], 0xFFFFFFFFu, w) != 0xFFFFFFFFu) slot = (slot + 1) & mask;
        }
        wave_lds_sync();
      }
      for (uint32_t c = 0; c < a; c += GC_WAVE) {
        const uint32_t nv = min(static_cast<uint32_t>(GC_WAVE), a - c);
        uint32_t len = 0, st = 0;
        if (lane < nv) {
          const uint32_t v = A[c + lane];
          st = lo[v];
          len = hi[v] - st;
        }
        // One list L(v) at a time, all 64 lanes on consecutive entries: no
        // per-item owner search (the bounds of the 64 lists were prefetched
        // above and are broadcast from their lane), so an item costs one
        // coalesced load and one hash probe.
        for (uint32_t j = 0; j < nv; ++j) {
          const uint32_t sj = __shfl(st, static_cast<int>(j));
          const uint32_t lj = __shfl(len, static_cast<int>(j));
          for (uint32_t k = lane; k < lj; k += TCH_ILP * GC_WAVE) {
            uint32_t w[TCH_ILP];
#pragma unroll
            for (int q = 0; q < TCH_ILP; ++q) {
              const uint32_t kk = k + q * GC_WAVE;
              w[q] = kk < lj ? targets[sj + kk] : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int q = 0; q < TCH_ILP; ++q) {
              if (w[q] == 0xFFFFFFFFu) continue;
              bool found;
              if (in_table) {
                uint32_t slot = (w[q] * 0x9E3779B1u) >> shift;
                uint32_t key = s_table[slot];
                while (key != w[q] && key != 0xFFFFFFFFu) {
                  slot = (slot + 1) & mask;
                  key = s_table[slot];
                }
                found = key == w[q];
              } else {
                const uint32_t p = lower_bound_u32(A, 0, a, w[q]);
                found = p < a && A[p] == w[q];
              }
              count += found ? 1ull : 0ull;
            }
          }
        }
      }
    }
  }
  count = wave_sum(count);
  if (lane == 0 && count != 0) atomicAdd(result, count);
}

// ---- hub-bitmap variant (default for symmetric graphs) ------------------------
//
// Same count as tc_hash. On the degree-ordered ids the lists L(v) point to
// higher-degree vertices, so almost every list entry w lies among the top ids:
// at R-MAT scale 22, 98.8% of the 2.9e10 entries probed fall in the top 32768
// ids. Each wave keeps L(u) ∩ [hot_base, n) as a 32768-bit LDS bitmap (4 KB)
// and tests such w with one LDS read and no branch; the few entries below
// hot_base are searched in the (sorted) cold prefix of L(u). Bits are set
// before and cleared after each u by walking L(u)'s hot suffix, so the
// bitmap is never cleared whole.

#define TCB_WAVES 4
#define TCB_HOT_BITS 32768
#define TCB_HOT_WORDS (TCB_HOT_BITS / 32)
#define TCB_COUNTER_STRIDE 64  // words between two parts' counters (own 256-byte line)
// A one-hash Bloom filter of the row's cold prefix (4096 bits) follows the
// bitmap's zero word: a cold id whose bit is clear is not in the prefix, and
// only the others are binary-searched. At R-MAT scale 22, 0.36 G probes are
// cold, and the rows with a cold prefix stream 68% of all entries: without
// the filter a 256-entry step there ran a divergent global binary search for
// its three or so cold entries.
#define TCB_BLOOM_LOG2 12
#define TCB_BLOOM_WORDS ((1 << TCB_BLOOM_LOG2) / 32)
#define TCB_BLOOM_BASE (TCB_HOT_WORDS + 1)

__device__ __forceinline__ uint32_t bloom_hash(uint32_t w) {
  return (w * 0x9E3779B1u) >> (32 - TCB_BLOOM_LOG2);
}

// hot_base is a multiple of 32, so w - hot_base and w share their low five
// bits: the bit is extracted with w itself (v_bfe_u32 reads only the offset's
// low five bits).
__device__ __forceinline__ uint32_t word_bit(uint32_t word, uint32_t w) {
  return __builtin_amdgcn_ubfe(word, w, 1u);
}

// 1 when w is a hot id whose bit is set in the row's bitmap. Branch-free: a
// cold id (w - hot_base >= TCB_HOT_BITS, wrapped) reads the bitmap's extra
// word TCB_HOT_WORDS, which stays zero, so the four lookups of a step issue
// back to back.
__device__ __forceinline__ uint32_t hot_bit(const uint32_t* s_bits, uint32_t w, uint32_t hot_base) {
  return word_bit(s_bits[min((w - hot_base) >> 5, static_cast<uint32_t>(TCB_HOT_WORDS))], w);
}

// 1 when w is a cold id (below hot_base) found in the sorted cold prefix
// A[0, c) of the row (searched only when the row's Bloom filter holds w).
__device__ __forceinline__ uint32_t cold_hit(const uint32_t* s_bits, const uint32_t* A, uint32_t c, uint32_t w,
                                             uint32_t hot_base) {
  if (w >= hot_base) return 0u;
  const uint32_t h = bloom_hash(w);
  if (((s_bits[TCB_BLOOM_BASE + (h >> 5)] >> (h & 31)) & 1u) == 0) return 0u;
  const uint32_t p = lower_bound_u32(A, 0, c, w);
  return (p < c && A[p] == w) ? 1u : 0u;
}

// Hit of one entry w of L(v): the clamped bitmap lookup plus, when the row
// has a cold prefix (c != 0, wave-uniform), its search. (A variant that skips
// the clamp and the cold test for the lists of hot v, whose entries are all
// hot, measured no faster.)
__device__ __forceinline__ uint32_t entry_hit(const uint32_t* s_bits, uint32_t w, uint32_t hot_base,
                                              const uint32_t* A, uint32_t c) {
  uint32_t hit = hot_bit(s_bits, w, hot_base);
  if (c != 0) hit += cold_hit(s_bits, A, c, w, hot_base);
  return hit;
}

// Hits of the row's lists among the entries L[0, lj) (lj wave-uniform): whole
// steps of 4 x 64 entries (four loads in flight per lane), then one guarded
// step for the rest (a single load per lane when at most 64 entries are left).
// Measured slower at R-MAT scale 22: nontemporal loads for the lists of low
// ids (to keep the lists of high ids cached), fewer loads in the last step
// (one whole step of 2 x 64, then at most two guarded loads), and steps of
// 8 x 64 or 16 x 64 entries.
__device__ __forceinline__ uint32_t probe_list(const uint32_t* L, uint32_t lj, const uint32_t* s_bits,
                                               uint32_t hot_base, const uint32_t* A, uint32_t c) {
  const uint32_t lane = lane_id();
  uint32_t local = 0;
  uint32_t kk = 0;
  for (; kk + 4 * GC_WAVE <= lj; kk += 4 * GC_WAVE) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = L[kk + q * GC_WAVE + lane];
#pragma unroll
    for (int q = 0; q < 4; ++q) local += entry_hit(s_bits, w[q], hot_base, A, c);
  }
  if (kk < lj) {
    // indexes clamped to the list's last entry (always a valid load), hits
    // masked to the entries left: no branches between the loads and lookups
    const uint32_t left = lj - kk;  // 1 .. 255, wave-uniform
    const uint32_t* T = L + kk;
    if (left <= GC_WAVE) {
      const uint32_t w = T[min(lane, left - 1)];
      local += lane < left ? entry_hit(s_bits, w, hot_base, A, c) : 0u;
    } else {
      uint32_t w[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) w[q] = T[min(q * GC_WAVE + lane, left - 1)];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        local += q * GC_WAVE + lane < left ? entry_hit(s_bits, w[q], hot_base, A, c) : 0u;
    }
  }
  return local;
}

// Two lists walked in lockstep (eight loads in flight per lane): whole steps
// of 4 x 64 entries of each while both have them, then guarded steps (index
// clamped to the list's last entry, hit masked) until the longer one ends.
// Consecutive pending lists belong to neighbouring ids, which the degree
// order gives similar lengths, so little of the guarded work is wasted. At
// R-MAT scale 22 pairs took the kernel from 17.7 to 15.8 ms (same box);
// groups of three or four lists measured slower than pairs.
__device__ __forceinline__ uint32_t probe_pair(const uint32_t* L1, uint32_t l1, const uint32_t* L2, uint32_t l2,
                                               const uint32_t* s_bits, uint32_t hot_base, const uint32_t* A,
                                               uint32_t c) {
  const uint32_t lane = lane_id();
  uint32_t local = 0;
  uint32_t kk = 0;
  const uint32_t lmin = min(l1, l2), lmax = max(l1, l2);
  for (; kk + 4 * GC_WAVE <= lmin; kk += 4 * GC_WAVE) {
    uint32_t w[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      w[q] = L1[kk + q * GC_WAVE + lane];
      w[4 + q] = L2[kk + q * GC_WAVE + lane];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) local += entry_hit(s_bits, w[q], hot_base, A, c);
  }
  for (; kk < lmax; kk += 4 * GC_WAVE) {
    uint32_t w[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t i = kk + q * GC_WAVE + lane;
      w[q] = L1[min(i, l1 - 1)];
      w[4 + q] = L2[min(i, l2 - 1)];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t i = kk + q * GC_WAVE + lane;
      local += i < l1 ? entry_hit(s_bits, w[q], hot_base, A, c) : 0u;
      local += i < l2 ? entry_hit(s_bits, w[4 + q], hot_base, A, c) : 0u;
    }
  }
  return local;
}

// ---- 16-bit hot lists (tc_bitmap16) ----
// The hot part of a list is stored as octs: 16 bytes holding eight offsets
// y = w - hot_base (< TCB_HOT_BITS). A part whose length is not a multiple of
// eight ends in pad offsets TCB_HOT_BITS, whose lookups read the bitmap's zero
// word. So one 16-byte load carries eight entries (the texture path's work
// goes with the lane loads, not their width), and no entry needs a mask or a
// clamp.
#define TCB_PAD16 TCB_HOT_BITS
// Masks of the top ids: a list L(v) whose entries among the top ids (ids from
// top_base on, top_octs * 128 of them) would take more octs as offsets than
// their mask takes keeps only the entries below top_base as offsets and stores
// the mask after them; hw[v].y then carries TC_TOP_FLAG and counts the mask's
// octs too. L(v) holds ids above v only, so the mask starts at the oct of id
// max(v + 1, top_base) (tc_mask_first; the octs below it would be zero) and
// runs to the end of the top ids. The mask's bits and the row's bitmap words
// from top_word0 = (top_base - hot_base) / 32 on cover the same ids, so a
// masked top part adds popcount(mask & bitmap words): at most 512 bytes per
// list at top_octs = 32 against two bytes per top entry.
#define TC_TOP_FLAG 0x80000000u

// The first stored oct of row v's mask among the top_octs octs of the top ids.
__device__ __forceinline__ uint32_t tc_mask_first(uint32_t v, uint32_t top_base) {
  return v + 1 > top_base ? (v + 1 - top_base) >> 7 : 0u;
}

__device__ __forceinline__ uint32_t hot16_hit(const uint32_t* s_bits, uint32_t y) {
  return word_bit(s_bits[y >> 5], y);
}

// Two offsets in one word: the low one's bit is taken with x itself (v_bfe_u32
// reads only the offset's low five bits, the low offset's), its word index
// with one shift and mask; the high one's word index is x >> 21 (no mask).
__device__ __forceinline__ uint32_t hot16x2_hits(const uint32_t* s_bits, uint32_t x) {
  return word_bit(s_bits[(x >> 5) & 0x7FFu], x) + word_bit(s_bits[x >> 21], x >> 16);
}

__device__ __forceinline__ uint32_t hot16x8_hits(const uint32_t* s_bits, uint4 x) {
  return hot16x2_hits(s_bits, x.x) + hot16x2_hits(s_bits, x.y) + hot16x2_hits(s_bits, x.z) +
         hot16x2_hits(s_bits, x.w);
}

// As probe_list over the octs O[0, no) of a hot part: whole steps of 64 octs
// (512 entries), then one guarded step (oct index clamped to the last oct,
// hits masked by oct).
__device__ __forceinline__ uint32_t probe16_list(const uint4* O, uint32_t no, const uint32_t* s_bits) {
  const uint32_t lane = lane_id();
  uint32_t local = 0;
  uint32_t kk = 0;
  for (; kk + GC_WAVE <= no; kk += GC_WAVE) local += hot16x8_hits(s_bits, O[kk + lane]);
  if (kk < no) {
    const uint32_t left = no - kk;  // 1 .. 63, wave-uniform
    const uint4 x = O[kk + min(lane, left - 1)];
    local += lane < left ? hot16x8_hits(s_bits, x) : 0u;
  }
  return local;
}

// As probe_pair over the octs of two hot parts (two 16-byte loads in flight
// per lane, sixteen entries).
__device__ __forceinline__ uint32_t probe16_pair(const uint4* O1, uint32_t n1, const uint4* O2, uint32_t n2,
                                                 const uint32_t* s_bits) {
  const uint32_t lane = lane_id();
  uint32_t local = 0;
  uint32_t kk = 0;
  const uint32_t nmin = min(n1, n2), nmax = max(n1, n2);
  for (; kk + GC_WAVE <= nmin; kk += GC_WAVE) {
    const uint4 x1 = O1[kk + lane];
    const uint4 x2 = O2[kk + lane];
    local += hot16x8_hits(s_bits, x1) + hot16x8_hits(s_bits, x2);
  }
  for (; kk < nmax; kk += GC_WAVE) {
    const uint32_t i = kk + lane;
    const uint4 x1 = O1[min(i, n1 - 1)];
    const uint4 x2 = O2[min(i, n2 - 1)];
    local += i < n1 ? hot16x8_hits(s_bits, x1) : 0u;
    local += i < n2 ? hot16x8_hits(s_bits, x2) : 0u;
  }
  return local;
}

// Hits of the hot parts of the lists of the lanes in `lists` (lane j's list:
// octs [st, st + len) of O), a quarter wave (16 lanes, D octs each per step)
// per list: a quarter whose list has ended takes the next one, so the steps'
// lanes mostly carry entries. (One list per wave step, as probe16_list, fills
// half the lanes at R-MAT scale 22, where 99% of the list visits have fewer
// than 64 octs and most 17 to 64.)
template <int D>
__device__ __forceinline__ uint32_t probe16_quarters(const uint4* O, unsigned long long lists, uint32_t st,
                                                     uint32_t len, const uint32_t* s_bits) {
  const uint32_t lane = lane_id();
  const uint32_t quarter = lane >> 4, ql = lane & 15u;
  uint32_t cur = 0, end = 0;  // the quarter's next oct and end (the same on its 16 lanes)
  uint32_t local = 0;
  for (;;) {
    // the quarters without work take the next lists, in quarter order
    const unsigned long long idle = __ballot(cur >= end);
    uint32_t free4 = static_cast<uint32_t>((idle & 1ull) | ((idle >> 15) & 2ull) | ((idle >> 30) & 4ull) |
                                           ((idle >> 45) & 8ull));
    while (free4 != 0 && lists != 0) {
      const uint32_t fq = static_cast<uint32_t>(__ffs(free4) - 1);
      free4 &= free4 - 1;
      const int jv = __ffsll(static_cast<long long>(lists)) - 1;
      lists &= lists - 1;
      const uint32_t s1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(st, jv));
      const uint32_t l1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(len, jv));
      if (quarter == fq) {
        cur = s1;
        end = s1 + l1;
      }
    }
    if (__ballot(cur < end) == 0) break;
    // D octs per lane: lanes past their list's end read its last oct again (a quarter that never
    // had one, oct 0)
    uint4 x[D];
    bool act[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const uint32_t i = cur + 16u * d + ql;
      act[d] = i < end;
      x[d] = O[act[d] ? i : (end > 0 ? end - 1 : 0u)];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) local += act[d] ? hot16x8_hits(s_bits, x[d]) : 0u;
    cur += 16u * D;
  }
  return local;
}

// Lead lanes of the groups of G lanes (bit g * G for group g).
template <int G>
__device__ __forceinline__ constexpr unsigned long long group_leads() {
  return G == 4 ? 0x1111111111111111ull : G == 8 ? 0x0101010101010101ull : G == 16 ? 0x0001000100010001ull
                                                                                    : 0x0000000100000001ull;
}

// As probe16_quarters with groups of G lanes (4, 8, 16 or 32) per list, D octs
// per lane and step: 64 / G lists in flight per step, so that many short
// lists cost one step together rather than one round of four each.
template <int G, int D>
__device__ __forceinline__ uint32_t probe16_groups(const uint4* __restrict__ O, unsigned long long lists,
                                                   uint32_t st, uint32_t len, const uint32_t* s_bits) {
  const uint32_t lane = lane_id();
  const uint32_t grp = lane / G, gl = lane % G;
  uint32_t cur = 0, end = 0;
  uint32_t local = 0;
  for (;;) {
    unsigned long long idle = __ballot(cur >= end) & group_leads<G>();
    while (idle != 0 && lists != 0) {
      const uint32_t fg = static_cast<uint32_t>(__ffsll(static_cast<long long>(idle)) - 1) / G;
      idle &= idle - 1;
      const int jv = __ffsll(static_cast<long long>(lists)) - 1;
      lists &= lists - 1;
      const uint32_t s1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(st, jv));
      const uint32_t l1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(len, jv));
      if (grp == fg) {
        cur = s1;
        end = s1 + l1;
      }
    }
    if (__ballot(cur < end) == 0) break;
    uint4 x[D];
    bool act[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const uint32_t i = cur + static_cast<uint32_t>(G) * d + gl;
      act[d] = i < end;
      x[d] = O[act[d] ? i : (end > 0 ? end - 1 : 0u)];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) local += act[d] ? hot16x8_hits(s_bits, x[d]) : 0u;
    cur += static_cast<uint32_t>(G) * D;
  }
  return local;
}

// Hits of the cold parts T[st, st + len) of the lanes' lists (every entry
// below hot_base) among the row's cold prefix A[0, c): the Bloom filter, then
// a search. G lanes per part, D entries per lane and step (a group whose part
// has ended takes the next one): the parts are short (26 entries on average
// at R-MAT scale 22), and one per wave step left most lanes idle and put a
// global round trip per part on the wave's path.
template <int G, int D>
__device__ __forceinline__ uint32_t probe_cold_groups(const uint32_t* __restrict__ T, unsigned long long lists,
                                                      uint32_t st, uint32_t len, const uint32_t* s_bits,
                                                      const uint32_t* A, uint32_t c) {
  const uint32_t lane = lane_id();
  const uint32_t grp = lane / G, gl = lane % G;
  uint32_t cur = 0, end = 0;
  uint32_t local = 0;
  for (;;) {
    unsigned long long idle = __ballot(cur >= end) & group_leads<G>();
    while (idle != 0 && lists != 0) {
      const uint32_t fg = static_cast<uint32_t>(__ffsll(static_cast<long long>(idle)) - 1) / G;
      idle &= idle - 1;
      const int jv = __ffsll(static_cast<long long>(lists)) - 1;
      lists &= lists - 1;
      const uint32_t s1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(st, jv));
      const uint32_t l1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(len, jv));
      if (grp == fg) {
        cur = s1;
        end = s1 + l1;
      }
    }
    if (__ballot(cur < end) == 0) break;
    uint32_t w[D];
    bool act[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const uint32_t i = cur + static_cast<uint32_t>(G) * d + gl;
      act[d] = i < end;
      w[d] = T[act[d] ? i : (end > 0 ? end - 1 : 0u)];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (act[d]) {
        const uint32_t h = bloom_hash(w[d]);
        if ((s_bits[TCB_BLOOM_BASE + (h >> 5)] >> (h & 31)) & 1u) {
          const uint32_t p = lower_bound_u32(A, 0, c, w[d]);
          local += (p < c && A[p] == w[d]) ? 1u : 0u;
        }
      }
    }
    cur += static_cast<uint32_t>(G) * D;
  }
  return local;
}

// Partitions by middle vertex (`parts` > 1): a row u with |L(u)| >= split_min
// is processed once per part p, and part p walks only the lists L(v) of the
// v in L(u) with v % parts == p; the parts' sums add up to the same count.
// Workgroup b serves part b % parts first, so with parts dividing 8 (the XCD
// count) and the round-robin placement of workgroups over XCDs, each XCD
// streams parts-th of the lists through its own L2 instead of all of them:
// the lists of the top ids, which carry most of the streaming, then fit the
// L2s together (at R-MAT scale 22 with 8 parts the L2 hit rate goes from 20%
// to 52% and the kernel from 35 to 18 ms). A shorter row is processed whole, once, by part u % parts:
// building its bitmap once per part costs more than its streaming saves (at
// scale 22 the rows with |L(u)| < 64 are 2.1 M of the 2.4 M non-empty rows
// and 22% of the streaming; with the 16-bit lists the best threshold is
// 256). Every part has its own pair of work counters
// over all rows; a wave that finds its part drained takes rows of the others,
// so an uneven part does not leave XCDs idle (where a row runs changes its
// speed, never its count). Part p's counters are
// counters[TCB_COUNTER_STRIDE * p + 0, 1], each on a 256-byte line of its own
// (sharing one line, the parts' atomics serialized: 25 ms instead of 21).
// (Packing lo and hi into one uint2 array, one 8-byte load per list instead
// of two 4-byte ones, measured slower: 14.7 -> 15.4 ms at R-MAT scale 22.)
template <bool kHot16>
__device__ __forceinline__ void tc_bitmap_body(
    const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ split, const uint2* __restrict__ hw, const uint4* __restrict__ quads,
    uint32_t n, uint32_t row_lo, uint32_t hot_base,
    uint32_t heavy, uint32_t light_chunk, uint32_t parts, uint32_t split_min,
    uint32_t* __restrict__ counters, unsigned long long* __restrict__ result, uint32_t* s_bits,
    uint32_t top_word0 = 0, uint32_t top_octs = 0, uint32_t quarters = 0,
    const uint32_t* __restrict__ midbits = nullptr, const uint8_t* __restrict__ rowwork = nullptr) {
  const uint32_t positions = n - row_lo;  // rows [row_lo, n)
  const uint32_t lane = lane_id();
  for (uint32_t j = lane; j < TCB_BLOOM_BASE + TCB_BLOOM_WORDS; j += GC_WAVE) s_bits[j] = 0;
  wave_lds_sync();
  unsigned long long count = 0;
  const uint32_t home = blockIdx.x % parts;

  for (uint32_t q = 0; q < parts; ++q) {
    const uint32_t part = (home + q) % parts;
    uint32_t* part_counters = counters + TCB_COUNTER_STRIDE * part;
    bool heavy_left = heavy > 0;
    for (;;) {
      uint32_t kb = positions, ke = positions;
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
        if (!heavy_left && kb == positions) {
          const uint32_t c = atomicAdd(&part_counters[1], light_chunk);
          const uint64_t b = static_cast<uint64_t>(heavy) + c;
          if (b < positions) {
            kb = static_cast<uint32_t>(b);
            const uint64_t e = b + light_chunk;
            ke = e < positions ? static_cast<uint32_t>(e) : positions;
          }
        }
      }
      kb = __shfl(kb, 0);
      ke = __shfl(ke, 0);
      if (kb >= positions) break;

      // the task's rows, 64 at a time: lane j reads row kb + j's list bounds
      for (uint32_t kc = kb; kc < ke; kc += GC_WAVE) {
        uint32_t row_lo_e = 0, row_a = 0;
        bool take = false;
        if (kc + lane < ke) {
          const uint32_t u = n - 1 - (kc + lane);
          row_lo_e = lo[u];
          row_a = hi[u] - row_lo_e;
          // a row with one DAG entry v closes no triangle (v is not in L(v)); with the middle
          // side on, rowwork[u] = 0 marks a row whose every entry that can close one went there
          take = row_a >= 2 && (rowwork == nullptr || rowwork[u] != 0) && (row_a >= split_min || u % parts == part);
        }
        unsigned long long rows = __ballot(take);
        while (rows != 0) {
          const int jr = __ffsll(static_cast<long long>(rows)) - 1;
          rows &= rows - 1;
          const uint32_t a = static_cast<uint32_t>(__builtin_amdgcn_readlane(row_a, jr));
          const uint32_t a0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(row_lo_e, jr));
          const uint32_t* A = targets + a0;
          const bool whole = a < split_min;  // processed by one part, every v
          // cold prefix [0, c): entries below hot_base; set the hot suffix's bits
          uint32_t c;
          if (a <= GC_WAVE) {
            const uint32_t x = lane < a ? A[lane] : 0xFFFFFFFFu;
            c = static_cast<uint32_t>(__popcll(__ballot(x < hot_base)));  // uniform
            if (lane >= c && lane < a) {
              const uint32_t y = x - hot_base;
              atomicOr(&s_bits[y >> 5], 1u << (y & 31));
            } else if (lane < c) {
              const uint32_t h = bloom_hash(x);
              atomicOr(&s_bits[TCB_BLOOM_BASE + (h >> 5)], 1u << (h & 31));
            }
          } else {
            // the row's first hot entry: one load of split with the 16-bit
            // lists, else a binary search of the row
            if (kHot16) {
              const uint32_t u = n - 1 - (kc + static_cast<uint32_t>(jr));
              c = __builtin_amdgcn_readfirstlane(split[u] -
                                                 static_cast<uint32_t>(__builtin_amdgcn_readlane(row_lo_e, jr)));
            } else {
              c = __builtin_amdgcn_readfirstlane(lower_bound_u32(A, 0, a, hot_base));
            }
            for (uint32_t j = c + lane; j < a; j += GC_WAVE) {
              const uint32_t y = A[j] - hot_base;
              atomicOr(&s_bits[y >> 5], 1u << (y & 31));
            }
            for (uint32_t j = lane; j < c; j += GC_WAVE) {
              const uint32_t h = bloom_hash(A[j]);
              atomicOr(&s_bits[TCB_BLOOM_BASE + (h >> 5)], 1u << (h & 31));
            }
          }
          wave_lds_sync();
          uint32_t local = 0;
          for (uint32_t cv = 0; cv < a; cv += GC_WAVE) {
            const uint32_t nv = min(static_cast<uint32_t>(GC_WAVE), a - cv);
            // kHot16: [st, st + len) are the octs of L(v)'s hot part and
            // [cst, cst + clen) its cold part in targets (walked only when
            // the row has a cold prefix); otherwise all of L(v) in targets
            uint32_t st = 0, len = 0, cst = 0, clen = 0, mst = 0, mfirst = 0;
            bool masked = false;  // L(v)'s top part is a mask (TC_TOP_FLAG)
            if (lane < nv) {
              const uint32_t v = A[cv + lane];
              // an entry u->v whose bit is set in midbits is counted by tc_mid16 at v
              const uint32_t e = a0 + cv + lane;
              const bool mid = midbits != nullptr && ((midbits[e >> 5] >> (e & 31u)) & 1u) != 0;
              if (!mid && (whole || v % parts == part)) {
                if (kHot16) {
                  const uint2 h = hw[v];
                  st = h.x;
                  len = h.y & ~TC_TOP_FLAG;
                  if (h.y & TC_TOP_FLAG) {
                    masked = true;
                    mfirst = tc_mask_first(v, hot_base + 32u * top_word0);
                    len -= top_octs - mfirst;
                    mst = st + len;
                  }
                  if (c != 0) {
                    cst = lo[v];
                    clen = split[v] - cst;
                  }
                } else {
                  st = lo[v];
                  len = hi[v] - st;
                }
              }
            }
            // the lanes holding a non-empty list to walk; with the 16-bit lists, those of at most
            // quarters >> 8 octs go a quarter wave each, (quarters & 0xFF) octs per lane and step
            // (0: every list a wave step)
            unsigned long long pending = __ballot(len != 0);
            if (kHot16 && (quarters & 0xFFu) != 0) {
              const unsigned long long small = __ballot(len != 0 && len <= ((quarters >> 8) & 0xFFFFu));
              if (small != 0) {
                switch (quarters & 0xFFu) {
                  case 1: local += probe16_quarters<1>(quads, small, st, len, s_bits); break;
                  case 3: local += probe16_quarters<3>(quads, small, st, len, s_bits); break;
                  case 4: local += probe16_quarters<4>(quads, small, st, len, s_bits); break;
                  default: local += probe16_quarters<2>(quads, small, st, len, s_bits); break;
                }
              }
              pending &= ~small;
            }
            while (pending != 0) {
              const int jv = __ffsll(static_cast<long long>(pending)) - 1;
              pending &= pending - 1;
              // jv is wave-uniform: read the lane's values into scalars
              const uint32_t s1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(st, jv));
              const uint32_t lj = static_cast<uint32_t>(__builtin_amdgcn_readlane(len, jv));
              if (pending != 0) {  // a second list: walk the two in lockstep
                const int jv2 = __ffsll(static_cast<long long>(pending)) - 1;
                pending &= pending - 1;
                const uint32_t s2 = static_cast<uint32_t>(__builtin_amdgcn_readlane(st, jv2));
                const uint32_t l2 = static_cast<uint32_t>(__builtin_amdgcn_readlane(len, jv2));
                if (kHot16) local += probe16_pair(quads + s1, lj, quads + s2, l2, s_bits);
                else local += probe_pair(targets + s1, lj, targets + s2, l2, s_bits, hot_base, A, c);
              } else {
                if (kHot16) local += probe16_list(quads + s1, lj, s_bits);
                else local += probe_list(targets + s1, lj, s_bits, hot_base, A, c);
              }
            }
            if (kHot16 && top_octs != 0) {
              // the masks of the top ids: top_octs lanes per mask (one oct,
              // four words, each), 64 / top_octs masks per step, against the
              // same words of the row's bitmap
              unsigned long long mrows = __ballot(masked);
              const uint32_t per = GC_WAVE / top_octs;
              const uint32_t grp = lane / top_octs, chunk = lane % top_octs;
              const uint32_t* ubw = s_bits + top_word0 + 4 * chunk;  // (top_word0 need not be a multiple of 4)
              const uint4 ub = make_uint4(ubw[0], ubw[1], ubw[2], ubw[3]);
              while (mrows != 0) {
                uint32_t mine = 0;
                bool take = false;
                for (uint32_t g = 0; g < per && mrows != 0; ++g) {
                  const uint32_t j = static_cast<uint32_t>(__ffsll(static_cast<long long>(mrows)) - 1);
                  mrows &= mrows - 1;
                  if (grp == g) {
                    mine = j;
                    take = true;
                  }
                }
                const uint32_t base = static_cast<uint32_t>(__shfl(static_cast<int>(mst), static_cast<int>(mine)));
                const uint32_t first = static_cast<uint32_t>(__shfl(static_cast<int>(mfirst), static_cast<int>(mine)));
                if (take && chunk >= first) {
                  const uint4 mk = quads[base + chunk - first];
                  local += __popc(mk.x & ub.x) + __popc(mk.y & ub.y) + __popc(mk.z & ub.z) + __popc(mk.w & ub.w);
                }
              }
            }
            if (kHot16 && c != 0) {
              const unsigned long long cold = __ballot(clen != 0);
              if (cold != 0) {
                // quarters >> 24 lanes per part (4 or 8; else 16)
                switch (quarters >> 24) {
                  case 4: local += probe_cold_groups<4, 4>(targets, cold, cst, clen, s_bits, A, c); break;
                  case 8: local += probe_cold_groups<8, 2>(targets, cold, cst, clen, s_bits, A, c); break;
                  default: local += probe_cold_groups<16, 2>(targets, cold, cst, clen, s_bits, A, c); break;
                }
              }
            }
          }
          count += local;
          wave_lds_sync();
          for (uint32_t j = c + lane; j < a; j += GC_WAVE) s_bits[(A[j] - hot_base) >> 5] = 0;
          for (uint32_t j = lane; j < c; j += GC_WAVE) s_bits[TCB_BLOOM_BASE + (bloom_hash(A[j]) >> 5)] = 0;
          wave_lds_sync();
        }
      }
    }
  }
  count = wave_sum(count);
  if (lane == 0 && count != 0) atomicAdd(result, count);
}

extern "C" __global__ __launch_bounds__(TCB_WAVES * GC_WAVE) void tc_bitmap(
    const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi, const uint32_t* __restrict__ targets,
    uint32_t n, uint32_t row_lo, uint32_t hot_base,
    uint32_t heavy, uint32_t light_chunk, uint32_t parts, uint32_t split_min,
    uint32_t* __restrict__ counters, unsigned long long* __restrict__ result) {
  // per wave: the row's hot bitmap, one word that stays zero (for cold ids)
  // and the Bloom filter of the row's cold prefix
  __shared__ uint32_t s_bits_all[TCB_WAVES][TCB_BLOOM_BASE + TCB_BLOOM_WORDS];
  tc_bitmap_body<false>(lo, hi, targets, nullptr, nullptr, nullptr, n, row_lo, hot_base, heavy, light_chunk, parts,
                        split_min, counters, result, s_bits_all[threadIdx.x / GC_WAVE]);
}

// As tc_bitmap, with the hot part of every list read as 16-bit offsets from
// hot_base, eight to a 16-byte oct (quads; hw[v] = {first oct, octs} of
// L(v)'s hot part, split[v] = the first entry of L(v) at or above hot_base):
// half the bytes and an eighth of the loads for the entries that carry almost
// all of the streaming, no clamp and no cold test per hot entry, and the cold
// parts walked apart, only for the rows that have a cold prefix.
extern "C" __global__ __launch_bounds__(TCB_WAVES * GC_WAVE) void tc_bitmap16_w4(
    const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ split, const uint2* __restrict__ hw, const uint4* __restrict__ quads,
    uint32_t n, uint32_t row_lo, uint32_t hot_base,
    uint32_t heavy, uint32_t light_chunk, uint32_t parts, uint32_t split_min,
    uint32_t* __restrict__ counters, unsigned long long* __restrict__ result, uint32_t top_word0,
    uint32_t top_octs, uint32_t quarters, const uint32_t* __restrict__ midbits,
    const uint8_t* __restrict__ rowwork) {
  __shared__ uint32_t s_bits_all[TCB_WAVES][TCB_BLOOM_BASE + TCB_BLOOM_WORDS];
  tc_bitmap_body<true>(lo, hi, targets, split, hw, quads, n, row_lo, hot_base, heavy, light_chunk, parts,
                       split_min, counters, result, s_bits_all[threadIdx.x / GC_WAVE], top_word0, top_octs,
                       quarters, midbits, rowwork);
}

// The same with one-wave workgroups (the default; tc_bitmap16_w4 above with
// four): the wave's bitmap is the workgroup's only LDS, at address 0, so a
// lookup's address needs no per-wave base (eight adds per oct fewer). The
// kernel waits on its list loads two thirds of the time, so it is built for
// eight waves per SIMD (its SGPRs held it to seven; the compiler now keeps 24
// of them in VGPR lanes). At R-MAT scale 22 on one box: 5.92 ms, against 6.79
// for tc_bitmap16_w4 (6.75 against 6.82 at seven waves).
extern "C" __global__ __launch_bounds__(GC_WAVE) __attribute__((amdgpu_waves_per_eu(8, 8))) void tc_bitmap16(
    const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ split, const uint2* __restrict__ hw, const uint4* __restrict__ quads,
    uint32_t n, uint32_t row_lo, uint32_t hot_base,
    uint32_t heavy, uint32_t light_chunk, uint32_t parts, uint32_t split_min,
    uint32_t* __restrict__ counters, unsigned long long* __restrict__ result, uint32_t top_word0,
    uint32_t top_octs, uint32_t quarters, const uint32_t* __restrict__ midbits,
    const uint8_t* __restrict__ rowwork) {
  __shared__ uint32_t s_bits[TCB_BLOOM_BASE + TCB_BLOOM_WORDS];
  tc_bitmap_body<true>(lo, hi, targets, split, hw, quads, n, row_lo, hot_base, heavy, light_chunk, parts,
                       split_min, counters, result, s_bits, top_word0, top_octs, quarters, midbits, rowwork);
}

// The hot lists of tc_bitmap16, in two passes around an exclusive scan of
// nquads: tc_hot16_rows sets split[v] (the first index of [lo[v], hi[v]) whose
// target is at or above hot_base) and the octs of the hot part; tc_hot16_fill
// writes every hot entry's offset into its row's octs (from the scanned oct
// offset), the pads of the last oct, and hw[v]. The rows are laid out part
// by part (part v % parts, then v), in the middle-vertex parts of tc_bitmap16:
// each part's lists are then contiguous, and an XCD's L2 fetches no line
// shared with another part's lists. part_slot gives row v's place there.
__device__ __forceinline__ uint32_t part_slot(uint32_t v, uint32_t n, uint32_t parts) {
  const uint32_t p = v % parts;
  return p * (n / parts) + min(p, n % parts) + v / parts;
}

extern "C" __global__ __launch_bounds__(256) void tc_hot16_rows(
    const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi, const uint32_t* __restrict__ targets,
    uint32_t n, uint32_t hot_base, uint32_t parts, uint32_t* __restrict__ split, uint32_t* __restrict__ nquads,
    uint32_t top_base, uint32_t top_octs, uint32_t* __restrict__ tsplit) {
  for (uint64_t v = global_thread(); v < n; v += grid_threads()) {
    const uint32_t b = hi[v];
    const uint32_t sp = lower_bound_u32(targets, lo[v], b, hot_base);
    split[v] = sp;
    uint32_t end = b, extra = 0;
    if (top_octs != 0) {
      // a mask (its octs from tc_mask_first on) when the top part would take more octs as offsets
      const uint32_t ts = lower_bound_u32(targets, sp, b, top_base);
      const uint32_t mask_octs = top_octs - tc_mask_first(static_cast<uint32_t>(v), top_base);
      if (((b - ts) >> 3) > mask_octs) {
        end = ts;
        extra = mask_octs;
      }
      tsplit[v] = end;
    }
    nquads[part_slot(static_cast<uint32_t>(v), n, parts)] = ((end - sp + 7) >> 3) + extra;
  }
}

// Within a row, entries are placed by blocks of 512 (64 octs; the last block
// holds Q = ceil(rest / 8) octs): entry r of a block goes to oct r % Q, slot
// r / Q. A probe's lane reads oct q of a block and looks up its eight
// offsets in turn, so each LDS lookup instruction of the wave covers Q
// consecutive entries of the list (close ids, few bitmap words and bank
// conflicts) rather than entries eight apart. The count does not depend on
// the order of a list's entries.
__device__ __forceinline__ uint64_t quad_slot(uint32_t e, uint32_t len) {
  const uint32_t b = e >> 9, r = e & 511u;
  const uint32_t rest = len - (b << 9);
  const uint32_t q_count = rest >= 512u ? 64u : (rest + 7u) >> 3;
  return 8ull * (64ull * b + r % q_count) + r / q_count;
}

extern "C" __global__ __launch_bounds__(256) void tc_hot16_fill(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t n, uint64_t m,
    uint32_t hot_base, uint32_t parts, const uint32_t* __restrict__ split, const uint32_t* __restrict__ qoff,
    uint2* __restrict__ hw, uint16_t* __restrict__ half, uint32_t top_base, uint32_t top_octs,
    const uint32_t* __restrict__ tsplit) {
  const uint64_t total = m > n ? m : static_cast<uint64_t>(n);
  for (uint64_t i = global_thread(); i < total; i += grid_threads()) {
    if (i < m) {
      const uint32_t v = row_of_entry(offsets, n, static_cast<uint32_t>(i));
      const uint32_t sp = split[v];
      const uint32_t row_end = offsets[v + 1];
      const uint32_t end = top_octs != 0 ? tsplit[v] : row_end;  // the offsets' part
      const uint64_t q8 = 8ull * qoff[part_slot(v, n, parts)];
      if (i >= sp && i < end) {
        half[q8 + quad_slot(static_cast<uint32_t>(i) - sp, end - sp)] = static_cast<uint16_t>(targets[i] - hot_base);
      } else if (i >= end && i < row_end) {
        // a masked top entry: its bit in the mask after the offsets' octs (zeroed beforehand)
        const uint32_t y = targets[i] - top_base;
        uint32_t* mask = reinterpret_cast<uint32_t*>(half + q8 + 8ull * ((end - sp + 7) >> 3));
        atomicOr(mask + ((y >> 5) - 4u * tc_mask_first(v, top_base)), 1u << (y & 31));
      }
    }
    if (i < n) {
      const uint32_t v = static_cast<uint32_t>(i);
      const uint32_t q = qoff[part_slot(v, n, parts)];
      const uint32_t row_end = offsets[v + 1];
      const uint32_t end = top_octs != 0 ? tsplit[v] : row_end;
      const uint32_t len = end - split[v];
      const uint32_t nq = (len + 7) >> 3;
      hw[v] = end < row_end ? make_uint2(q, (nq + top_octs - tc_mask_first(v, top_base)) | TC_TOP_FLAG)
                            : make_uint2(q, nq);
      // the last block's free slots (at most seven) hold the pad
      if (len & 7u) {
        const uint32_t last = (len - 1) >> 9;
        const uint32_t rest = len - (last << 9);
        const uint32_t q_count = (rest + 7u) >> 3;
        for (uint32_t r = rest; r < 8 * q_count; ++r)
          half[8ull * (q + 64ull * last + r % q_count) + r / q_count] = static_cast<uint16_t>(TCB_PAD16);
      }
    }
  }
}

// ---- middle-vertex walk for the heavy lists (tc_mid16) ----
//
// A triangle u < v < w of the DAG (u->v, u->w, v->w) can be found from either
// side of its edge u->v: tc_bitmap16 streams all of L(v) against the bitmap of
// L(u); tc_mid16 streams the part of L(u) above v against the bitmap of L(v).
// Both give |L(u) ∩ L(v)| (every entry of L(v) is above v). The second walk is
// short wherever v sits near the end of L(u) and L(v) is long, which at R-MAT
// scale 22 holds for 71% of the DAG edges: their share of tc_bitmap16's
// streaming falls from 2.9e10 entries to 4.7e9 walked by tc_mid16. Each edge
// u->v with v among the hot ids is given to the cheaper side once per graph
// (tc_mid_records): the middle side when the octs of its walk, plus a fixed
// cost per edge, are fewer than the octs tc_bitmap16 streams for L(v). Its
// bit is then set in midbits (by DAG entry), which tc_bitmap16 skips.
//
// tc_mid16's lists are the DAG rows' hot parts in natural order, 16-bit
// offsets from hot_base, each row from an oct boundary and padded to the next
// with TCB_PAD16 (tc_mid_rows, then tc_mid_fill around an exclusive scan of
// the octs). A walk starts at the oct holding the first entry above v: the
// entries before it in that oct are ids of L(u) up to v, never in L(v) (whose
// ids are all above v), and the pads read the bitmap's zero word, so no entry
// needs a mask.

// Octs of row v's natural-order hot list.
extern "C" __global__ __launch_bounds__(256) void tc_mid_rows(const uint32_t* __restrict__ split,
                                                              const uint32_t* __restrict__ hi, uint32_t n,
                                                              uint32_t* __restrict__ nocts) {
  for (uint64_t v = global_thread(); v < n; v += grid_threads()) nocts[v] = (hi[v] - split[v] + 7u) >> 3;
}

// Every hot DAG entry's offset at its place (hq[v] = the row's first oct) and
// the pads of each row's last oct.
extern "C" __global__ __launch_bounds__(256) void tc_mid_fill(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, uint32_t n, uint64_t m,
    uint32_t hot_base, const uint32_t* __restrict__ split, const uint32_t* __restrict__ hq,
    uint16_t* __restrict__ half) {
  const uint64_t total = m > n ? m : static_cast<uint64_t>(n);
  for (uint64_t i = global_thread(); i < total; i += grid_threads()) {
    if (i < m) {
      const uint32_t v = row_of_entry(offsets, n, static_cast<uint32_t>(i));
      const uint32_t sp = split[v];
      if (i >= sp) half[8ull * hq[v] + (i - sp)] = static_cast<uint16_t>(targets[i] - hot_base);
    }
    if (i < n) {
      const uint32_t v = static_cast<uint32_t>(i);
      const uint32_t len = offsets[v + 1] - split[v];
      for (uint32_t r = len; r < ((len + 7u) & ~7u); ++r) half[8ull * hq[v] + r] = static_cast<uint16_t>(TCB_PAD16);
    }
  }
}

// One record per entry i of the hot rows v in [hot_base, n) (rec[i - rec_base]):
// for an entry u of v's in-list (u < v), the walk {first oct, octs} of the part of
// L(u) above v when the edge u->v goes to the middle side, else {0, 0}; entries
// of v's own DAG part and self-loops get {0, 0}. row_octs(v) = hw[v]'s oct count
// (the offsets' octs and, for a masked list, the mask's): what tc_bitmap16 streams
// for L(v). An edge whose walk is empty (v the last entry of L(u)) closes no
// triangle and always goes to the middle side, where it costs nothing.
// walks[v - hot_base] = 1 marks the rows with a non-empty walk (zeroed first):
// only they get work units.
extern "C" __global__ __launch_bounds__(256) void tc_mid_records(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, const uint32_t* __restrict__ fwd,
    uint32_t n, uint32_t hot_base, const uint32_t* __restrict__ split, const uint2* __restrict__ hw,
    const uint32_t* __restrict__ hq, uint32_t ovh, uint2* __restrict__ rec, uint32_t* __restrict__ midbits,
    uint8_t* __restrict__ walks) {
  const uint32_t rec_base = offsets[hot_base];
  const uint64_t total = static_cast<uint64_t>(offsets[n]) - rec_base;
  for (uint64_t k = global_thread(); k < total; k += grid_threads()) {
    const uint32_t i = rec_base + static_cast<uint32_t>(k);
    uint32_t lo = hot_base, hi = n - 1;  // the row of entry i: the largest v with offsets[v] <= i
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
        const uint32_t p = j - split[u];  // v's place in u's hot list (v >= hot_base)
        const uint32_t hot_octs = (ue - split[u] + 7u) >> 3;
        const uint32_t walk = j + 1 == ue ? 0u : hot_octs - ((p + 1) >> 3);
        const uint32_t row_octs = hw[v].y & ~TC_TOP_FLAG;
        if (walk == 0 || walk + ovh < row_octs) {
          r = make_uint2(hq[u] + ((p + 1) >> 3), walk);
          atomicOr(&midbits[j >> 5], 1u << (j & 31u));
          if (walk != 0) walks[v - hot_base] = 1;  // the row has a walk: it needs a unit
        }
      }
    }
    rec[k] = r;
  }
}

// rowwork[u] = 1 when row u has a DAG entry u->v that tc_bitmap16 still walks
// (its bit clear in midbits) and at least two DAG entries, else 0.
extern "C" __global__ __launch_bounds__(256) void tc_mid_rowwork(const uint32_t* __restrict__ lo,
                                                                 const uint32_t* __restrict__ hi, uint32_t n,
                                                                 const uint32_t* __restrict__ midbits,
                                                                 uint8_t* __restrict__ rowwork) {
  for (uint64_t u = global_thread(); u < n; u += grid_threads()) {
    const uint32_t b = lo[u], e = hi[u];
    uint8_t work = 0;
    if (e - b >= 2) {
      for (uint32_t j = b; j < e && work == 0; ++j) work = ((midbits[j >> 5] >> (j & 31u)) & 1u) == 0 ? 1 : 0;
    }
    rowwork[u] = work;
  }
}

// The middle side's count: units {v, first entry, end entry, 0} of the hot rows'
// in-lists, heaviest first from one work counter, `grab` units per atomic (the
// atomics of all waves on one counter serialize across the XCDs); per unit the wave sets L(v)'s
// bits in its LDS bitmap (L(v) is all hot: v >= hot_base), walks the records'
// lists (those of at most quarters >> 8 octs a quarter wave each, the others two
// at a time a wave step each, as tc_bitmap16 walks L(v)), then clears the bits.
extern "C" __global__ __launch_bounds__(GC_WAVE) __attribute__((amdgpu_waves_per_eu(8, 8))) void tc_mid16(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, const uint32_t* __restrict__ fwd,
    uint32_t hot_base, const uint2* __restrict__ rec, uint32_t rec_base, const uint4* __restrict__ octs,
    const uint4* __restrict__ units, uint32_t num_units, uint32_t quarters, uint32_t grab, uint32_t group,
    uint32_t* __restrict__ counter, unsigned long long* __restrict__ result) {
  __shared__ uint32_t s_bits[TCB_HOT_WORDS + 1];
  const uint32_t lane = lane_id();
  for (uint32_t j = lane; j < TCB_HOT_WORDS + 1; j += GC_WAVE) s_bits[j] = 0;
  wave_lds_sync();
  unsigned long long count = 0;
  uint32_t k = 0, k_end = 0;
  for (;; ++k) {
    if (k >= k_end) {  // the next `grab` units, one atomic on the shared counter for them all
      uint32_t g0 = 0;
      if (lane == 0) g0 = atomicAdd(counter, grab);
      k = __builtin_amdgcn_readfirstlane(__shfl(g0, 0));
      k_end = min(k + grab, num_units);
    }
    if (k >= num_units) break;
    const uint4 unit = units[k];
    const uint32_t v = __builtin_amdgcn_readfirstlane(unit.x);
    const uint32_t b = __builtin_amdgcn_readfirstlane(unit.y), e = __builtin_amdgcn_readfirstlane(unit.z);
    const uint32_t lb = fwd[v], le = offsets[v + 1];
    for (uint32_t j = lb + lane; j < le; j += GC_WAVE) {
      const uint32_t y = targets[j] - hot_base;
      atomicOr(&s_bits[y >> 5], 1u << (y & 31u));
    }
    wave_lds_sync();
    uint32_t local = 0;
    for (uint32_t i0 = b; i0 < e; i0 += GC_WAVE) {
      const uint32_t i = i0 + lane;
      const uint2 r = i < e ? rec[i - rec_base] : make_uint2(0u, 0u);
      const uint32_t st = r.x, len = r.y;
      unsigned long long pending = __ballot(len != 0);
      if ((quarters & 0xFFu) != 0) {
        const unsigned long long small = __ballot(len != 0 && len <= ((quarters >> 8) & 0xFFFFu));
        if (small != 0) {
          // `group` lanes per walk (4, 8 or 16), two octs per lane and step; 0: quarters as tc_bitmap16
          switch (group) {
            case 4: local += probe16_groups<4, 2>(octs, small, st, len, s_bits); break;
            case 8: local += probe16_groups<8, 2>(octs, small, st, len, s_bits); break;
            case 16: local += probe16_groups<16, 2>(octs, small, st, len, s_bits); break;
            default:
              switch (quarters & 0xFFu) {
                case 1: local += probe16_quarters<1>(octs, small, st, len, s_bits); break;
                case 3: local += probe16_quarters<3>(octs, small, st, len, s_bits); break;
                case 4: local += probe16_quarters<4>(octs, small, st, len, s_bits); break;
                default: local += probe16_quarters<2>(octs, small, st, len, s_bits); break;
              }
          }
        }
        pending &= ~small;
      }
      while (pending != 0) {
        const int jv = __ffsll(static_cast<long long>(pending)) - 1;
        pending &= pending - 1;
        const uint32_t s1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(st, jv));
        const uint32_t l1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(len, jv));
        if (pending != 0) {
          const int jv2 = __ffsll(static_cast<long long>(pending)) - 1;
          pending &= pending - 1;
          const uint32_t s2 = static_cast<uint32_t>(__builtin_amdgcn_readlane(st, jv2));
          const uint32_t l2 = static_cast<uint32_t>(__builtin_amdgcn_readlane(len, jv2));
          local += probe16_pair(octs + s1, l1, octs + s2, l2, s_bits);
        } else {
          local += probe16_list(octs + s1, l1, s_bits);
        }
      }
    }
    count += local;
    wave_lds_sync();
    for (uint32_t j = lb + lane; j < le; j += GC_WAVE) s_bits[(targets[j] - hot_base) >> 5] = 0;
    wave_lds_sync();
  }
  count = wave_sum(count);
  if (lane == 0 && count != 0) atomicAdd(result, count);
}

// ---- the middle side below the hot ids (tc_mid32) ----
//
// The edges u->v whose v lies just below the hot ids, in [mid_base, hot_base)
// (the host's choice of mid_base: by default the top 131072 ids), go to the middle side by the
// same rule, with the costs in 16-byte units: the walk at 4 bytes an entry
// against L(v)'s hot octs plus its cold entries at 4 bytes. Their walks are the
// 32-bit entries of L(u) above v, read in place from targets. L(v) is held as
// tc_bitmap16 holds L(u): its hot part in the wave's LDS bitmap, its cold
// prefix (ids in (v, hot_base)) behind a Bloom filter and searched in place.
// Records of the entries of rows v in [mid_base, hot_base) (rec[i - rec_base]):
// {first entry, entries} of the walk of L(u) above v, or {0, 0}; sets midbits,
// and walks[v - mid_base] = 1 for the rows with a non-empty walk.
extern "C" __global__ __launch_bounds__(256) void tc_mid32_records(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, const uint32_t* __restrict__ fwd,
    uint32_t mid_base, uint32_t hot_base, const uint32_t* __restrict__ split, const uint2* __restrict__ hw,
    uint32_t ovh, uint2* __restrict__ rec, uint32_t* __restrict__ midbits, uint8_t* __restrict__ walks) {
  const uint32_t rec_base = offsets[mid_base];
  const uint64_t total = static_cast<uint64_t>(offsets[hot_base]) - rec_base;
  for (uint64_t k = global_thread(); k < total; k += grid_threads()) {
    const uint32_t i = rec_base + static_cast<uint32_t>(k);
    uint32_t lo = mid_base, hi = hot_base - 1;  // the row of entry i
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
        const uint32_t row_cost = (hw[v].y & ~TC_TOP_FLAG) + ((split[v] - fwd[v] + 3u) >> 2);
        if (walk == 0 || ((walk + 3u) >> 2) + ovh < row_cost) {
          r = make_uint2(j + 1, walk);
          atomicOr(&midbits[j >> 5], 1u << (j & 31u));
          if (walk != 0) walks[v - mid_base] = 1;
        }
      }
    }
    rec[k] = r;
  }
}

// Hits of the walks T[st, st + len) of the lanes in `lists`, a quarter wave
// per walk, four entries per lane and step (the quarter's lanes on consecutive
// entries; a quarter whose walk has ended takes the next one), each entry
// tested as tc_bitmap tests one (entry_hit: the clamped bitmap lookup, and
// for a cold id, when L(v) has a cold prefix A[0, c), the filter and search).
__device__ __forceinline__ uint32_t probe32_quarters(const uint32_t* __restrict__ T, unsigned long long lists,
                                                    uint32_t st, uint32_t len, const uint32_t* s_bits,
                                                    uint32_t hot_base, const uint32_t* A, uint32_t c) {
  const uint32_t lane = lane_id();
  const uint32_t quarter = lane >> 4, ql = lane & 15u;
  uint32_t cur = 0, end = 0;
  uint32_t local = 0;
  for (;;) {
    const unsigned long long idle = __ballot(cur >= end);
    uint32_t free4 = static_cast<uint32_t>((idle & 1ull) | ((idle >> 15) & 2ull) | ((idle >> 30) & 4ull) |
                                           ((idle >> 45) & 8ull));
    while (free4 != 0 && lists != 0) {
      const uint32_t fq = static_cast<uint32_t>(__ffs(free4) - 1);
      free4 &= free4 - 1;
      const int jv = __ffsll(static_cast<long long>(lists)) - 1;
      lists &= lists - 1;
      const uint32_t s1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(st, jv));
      const uint32_t l1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(len, jv));
      if (quarter == fq) {
        cur = s1;
        end = s1 + l1;
      }
    }
    if (__ballot(cur < end) == 0) break;
    uint32_t w[4];
    bool act[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const uint32_t i = cur + 16u * d + ql;
      act[d] = i < end;
      w[d] = T[act[d] ? i : (end > 0 ? end - 1 : 0u)];
    }
#pragma unroll
    for (int d = 0; d < 4; ++d) local += act[d] ? entry_hit(s_bits, w[d], hot_base, A, c) : 0u;
    cur += 64u;
  }
  return local;
}

// The count of the edges given to this side: units {v, first entry, end entry,
// 0} of the rows [mid_base, hot_base) from one work counter, `grab` units per
// atomic; per unit the wave
// sets L(v)'s bits (hot part) and filter bits (cold prefix), walks the unit's
// records 64 at a time, then clears both.
extern "C" __global__ __launch_bounds__(GC_WAVE) __attribute__((amdgpu_waves_per_eu(8, 8))) void tc_mid32(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets, const uint32_t* __restrict__ fwd,
    const uint32_t* __restrict__ split, uint32_t hot_base, const uint2* __restrict__ rec, uint32_t rec_base,
    const uint4* __restrict__ units, uint32_t num_units, uint32_t grab, uint32_t* __restrict__ counter,
    unsigned long long* __restrict__ result) {
  __shared__ uint32_t s_bits[TCB_BLOOM_BASE + TCB_BLOOM_WORDS];
  const uint32_t lane = lane_id();
  for (uint32_t j = lane; j < TCB_BLOOM_BASE + TCB_BLOOM_WORDS; j += GC_WAVE) s_bits[j] = 0;
  wave_lds_sync();
  unsigned long long count = 0;
  uint32_t k = 0, k_end = 0;
  for (;; ++k) {
    if (k >= k_end) {  // the next `grab` units, one atomic on the shared counter for them all
      uint32_t g0 = 0;
      if (lane == 0) g0 = atomicAdd(counter, grab);
      k = __builtin_amdgcn_readfirstlane(__shfl(g0, 0));
      k_end = min(k + grab, num_units);
    }
    if (k >= num_units) break;
    const uint4 unit = units[k];
    const uint32_t v = __builtin_amdgcn_readfirstlane(unit.x);
    const uint32_t b = __builtin_amdgcn_readfirstlane(unit.y), e = __builtin_amdgcn_readfirstlane(unit.z);
    const uint32_t lb = fwd[v], le = offsets[v + 1];
    const uint32_t c = __builtin_amdgcn_readfirstlane(split[v] - lb);  // L(v)'s cold prefix A[0, c)
    const uint32_t* A = targets + lb;
    for (uint32_t j = lb + c + lane; j < le; j += GC_WAVE) {
      const uint32_t y = targets[j] - hot_base;
      atomicOr(&s_bits[y >> 5], 1u << (y & 31u));
    }
    for (uint32_t j = lane; j < c; j += GC_WAVE) {
      const uint32_t h = bloom_hash(A[j]);
      atomicOr(&s_bits[TCB_BLOOM_BASE + (h >> 5)], 1u << (h & 31u));
    }
    wave_lds_sync();
    uint32_t local = 0;
    for (uint32_t i0 = b; i0 < e; i0 += GC_WAVE) {
      const uint32_t i = i0 + lane;
      const uint2 r = i < e ? rec[i - rec_base] : make_uint2(0u, 0u);
      const unsigned long long lists = __ballot(r.y != 0);
      if (lists != 0) local += probe32_quarters(targets, lists, r.x, r.y, s_bits, hot_base, A, c);
    }
    count += local;
    wave_lds_sync();
    for (uint32_t j = lb + c + lane; j < le; j += GC_WAVE) s_bits[(targets[j] - hot_base) >> 5] = 0;
    for (uint32_t j = lane; j < c; j += GC_WAVE) s_bits[TCB_BLOOM_BASE + (bloom_hash(A[j]) >> 5)] = 0;
    wave_lds_sync();
  }
  count = wave_sum(count);
  if (lane == 0 && count != 0) atomicAdd(result, count);
}

// Triangle count of a symmetric graph on its degree-ordered relabelling
// (the count of crates/grafeo-cuda-algos/kernels/triangle_counting.cu:246-315),
// by enumeration at the middle vertex; see wedge.hpp.
extern "C" __global__ __launch_bounds__(64) void tc_wedge(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ fwd, uint32_t n, const uint32_t* __restrict__ hub_units,
    uint32_t num_hub_units, uint32_t hub_chunk, uint32_t table_bits, uint32_t* __restrict__ counters,
    unsigned long long* __restrict__ result) {
  wedge_body<false>(offsets, targets, fwd, nullptr, n, hub_units, num_hub_units, hub_chunk, table_bits,
                    counters, result, nullptr);
}

// Per degree bucket b (rows [bucket_lo[b], bucket_lo[b+1]) of the relabelled
// DAG): stats[4b + 0] rows, + 1 DAG entries, + 2 probe items (sum over the
// rows u of sum over v in L(u) of |L(v)|), + 3 largest |L(u)|. One atomic per
// wave when the wave's rows share a bucket.
extern "C" __global__ __launch_bounds__(256) void tc_bucket_stats(
    const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi,
    const uint32_t* __restrict__ targets, const uint32_t* __restrict__ bucket_lo, uint32_t buckets,
    uint32_t hot_base, unsigned long long* __restrict__ stats) {
  const uint32_t first = bucket_lo[0], last = bucket_lo[buckets];
  const uint64_t stride = grid_threads();
  const uint64_t steps = (static_cast<uint64_t>(last - first) + stride - 1) / stride;
  for (uint64_t k = 0; k < steps; ++k) {
    const uint64_t uu = first + global_thread() + k * stride;
    const bool valid = uu < last;
    uint32_t b = 0;
    unsigned long long d = 0, probes = 0, hot = 0;
    if (valid) {
      const uint32_t u = static_cast<uint32_t>(uu);
      uint32_t l = 0, r = buckets;  // last bucket whose first row is <= u
      while (r - l > 1) {
        const uint32_t mid = (l + r) / 2;
        if (bucket_lo[mid] <= u) l = mid;
        else r = mid;
      }
      b = l;
      const uint32_t a = lo[u], e = hi[u];
      d = e - a;
      for (uint32_t i = a; i < e; ++i) {
        const uint32_t v = targets[i];
        probes += hi[v] - lo[v];
        hot += hi[v] - lower_bound_u32(targets, lo[v], hi[v], hot_base);  // L(v) among the hot ids
      }
    }
    const uint32_t b0 = __shfl(b, 0);
    const bool uniform = __ballot(!valid || b == b0) == ~0ull && __shfl(valid ? 1u : 0u, 0) != 0u;
    if (uniform) {
      const unsigned long long rows = wave_sum(valid ? 1ull : 0ull);
      const unsigned long long entries = wave_sum(d);
      const unsigned long long items = wave_sum(probes);
      const unsigned long long hot_items = wave_sum(hot);
      unsigned long long mx = d;
#pragma unroll
      for (int o = GC_WAVE / 2; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(mx, o);
        mx = other > mx ? other : mx;
      }
      if (lane_id() == 0) {
        atomicAdd(&stats[5 * b0 + 0], rows);
        atomicAdd(&stats[5 * b0 + 1], entries);
        atomicAdd(&stats[5 * b0 + 2], items);
        atomicMax(&stats[5 * b0 + 3], mx);
        atomicAdd(&stats[5 * b0 + 4], hot_items);
      }
    } else if (valid) {
      atomicAdd(&stats[5 * b + 0], 1ull);
      atomicAdd(&stats[5 * b + 1], d);
      atomicAdd(&stats[5 * b + 2], probes);
      atomicMax(&stats[5 * b + 3], d);
      atomicAdd(&stats[5 * b + 4], hot);
    }
  }
}
