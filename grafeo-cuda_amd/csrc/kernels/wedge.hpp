// Triangle enumeration at the middle vertex of the degree-oriented DAG.
//
// On the degree-ordered relabelling of a symmetric graph, every triangle has
// vertices a < b < c and is found exactly once, at b: with N+(x) the entries
// of row x above x and N-(x) the entries below x,
//     triangles = sum over v, over u in N-(v), of |(suffix of N+(u) after v) ∩ N+(v)|.
// The work is the number of DAG wedges, sum_u C(d+(u), 2) (about 7.2e9 at
// R-MAT scale 22), against sum over entries u->v of d+(v) (2.9e10) for the
// neighbour-list form of tc_count.
//
// Layout: one wave per work unit (a vertex v and a slice of its in-list). The
// wave inserts N+(v) into an LDS hash table (open addressing, linear probing,
// at most half full; keys are unique), then walks 64 in-neighbours u at a
// time: each lane finds v in row u (binary search), the wave prefix-sums the
// suffix lengths and spreads the suffix entries w evenly over its lanes; each
// w is one hash probe. Hubs, whose in-lists run to 1e5 entries, are split into
// several units so no single wave holds the tail. Units: the hub slices listed
// by the host first (one per fetch), then every other vertex in descending id
// order (heaviest first) in chunks.
//
// SUPPORT mode (k-truss): only active entries take part (all of them when
// `removed` is null: the byte loads of the removal flags, one per wedge,
// cost more than the enumeration itself), and each triangle
// adds 1 to the support of its three forward entries u->v, u->w and v->w.
// Only u->w takes a global atomic per triangle: v->w is counted in the LDS
// table next to w (added to the support once per unit, after a binary search
// for its entry), and u->v in an LDS counter per in-neighbour lane (added
// once per chunk of 64 in-neighbours). Three global atomics per triangle made
// the pass at R-MAT scale 22 take 420 ms.
#pragma once

#include "device.hpp"

#define WEDGE_EMPTY 0xFFFFFFFFu
#define WEDGE_LIGHT_CHUNK 32u

namespace gcd {

__device__ __forceinline__ uint32_t wedge_hash(uint32_t key, uint32_t shift) {
  return (key * 0x9E3779B1u) >> shift;
}

// Entries of row v below v (its in-neighbours in the DAG), i.e. [offsets[v], end).
__device__ __forceinline__ uint32_t in_list_end(const uint32_t* __restrict__ offsets,
                                                const uint32_t* __restrict__ targets,
                                                const uint32_t* __restrict__ fwd, uint32_t v) {
  uint32_t end = fwd[v];
  if (end > offsets[v] && targets[end - 1] == v) --end;  // self-loop sits just below fwd
  return end;
}

template <bool SUPPORT>
__device__ __forceinline__ void wedge_body(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ fwd, const uint8_t* __restrict__ removed, uint32_t n,
    const uint32_t* __restrict__ hub_units, uint32_t num_hub_units, uint32_t hub_chunk,
    uint32_t table_bits, uint32_t* __restrict__ counters, unsigned long long* __restrict__ result,
    uint32_t* __restrict__ support) {
  extern __shared__ uint32_t s_dyn[];
  const uint32_t table_size = 1u << table_bits;
  uint32_t* s_key = s_dyn;
  uint32_t* s_val = s_dyn + table_size;  // used in SUPPORT mode only
  __shared__ uint32_t s_pos[GC_WAVE];
  __shared__ uint32_t s_start[GC_WAVE];
  __shared__ uint32_t s_edge[GC_WAVE];
  __shared__ uint32_t s_range[3];
  __shared__ uint32_t s_ucnt[GC_WAVE];  // SUPPORT: triangles of in-neighbour lane l's entry u->v

  const uint32_t lane = threadIdx.x;
  if (SUPPORT) s_ucnt[lane] = 0;
  unsigned long long count = 0;
  bool hubs_left = num_hub_units > 0;

  for (;;) {
    // ---- next unit(s): one hub slice, or a chunk of light vertices ----
    if (lane == 0) {
      uint32_t kind = 2, a = 0, b = 0;  // 0: hub slice a, 1: light positions [a, b), 2: done
      if (hubs_left) {
        const uint32_t k = atomicAdd(&counters[0], 1u);
        if (k < num_hub_units) {
          kind = 0;
          a = k;
        } else {
          hubs_left = false;
        }
      }
      if (kind == 2) {
        const uint32_t c = atomicAdd(&counters[1], WEDGE_LIGHT_CHUNK);
        if (c < n) {
          kind = 1;
          a = c;
          b = min(n, c + WEDGE_LIGHT_CHUNK);
        }
      }
      s_range[0] = kind;
      s_range[1] = a;
      s_range[2] = b;
    }
    __syncthreads();
    const uint32_t kind = s_range[0], ra = s_range[1], rb = s_range[2];
    __syncthreads();
    if (kind == 2) break;

    const uint32_t units = kind == 0 ? 1u : rb - ra;
    for (uint32_t x = 0; x < units; ++x) {
      uint32_t v, ib, ie;
      if (kind == 0) {
        v = hub_units[3 * ra];
        ib = hub_units[3 * ra + 1];
        ie = hub_units[3 * ra + 2];
      } else {
        v = n - 1 - (ra + x);
        ib = offsets[v];
        ie = in_list_end(offsets, targets, fwd, v);
        if (ie - ib > hub_chunk) continue;  // covered by hub slices
      }
      const uint32_t a_lo = fwd[v];
      const uint32_t a = offsets[v + 1] - a_lo;
      if (a == 0 || ib >= ie) continue;
      const bool in_table = 2 * a <= table_size;
      // per-vertex table: the smallest power of two >= 2a (at least one slot per lane)
      uint32_t bits_v = 6;
      while (bits_v < table_bits && (1u << bits_v) < 2 * a) ++bits_v;
      const uint32_t size_v = 1u << bits_v;
      const uint32_t shift_v = 32u - bits_v;
      const uint32_t mask_v = size_v - 1;
      if (in_table) {
        for (uint32_t j = lane; j < size_v; j += GC_WAVE) s_key[j] = WEDGE_EMPTY;
        __syncthreads();
        for (uint32_t j = lane; j < a; j += GC_WAVE) {
          const uint32_t e = a_lo + j;
          if (SUPPORT && removed != nullptr && removed[e]) continue;
          const uint32_t w = targets[e];
          uint32_t slot = wedge_hash(w, shift_v);
          for (;;) {
            const uint32_t prev = atomicCAS(&s_key[slot], WEDGE_EMPTY, w);
            if (prev == WEDGE_EMPTY) {
              if (SUPPORT) s_val[slot] = 0;  // triangles of v -> w in this unit
              break;
            }
            slot = (slot + 1) & mask_v;
          }
        }
        __syncthreads();
      }

      for (uint32_t c = ib; c < ie; c += GC_WAVE) {
        const uint32_t nv = min(static_cast<uint32_t>(GC_WAVE), ie - c);
        uint32_t len = 0, st = 0, e1 = 0;
        if (lane < nv) {
          const uint32_t u = targets[c + lane];
          const uint32_t hi = offsets[u + 1];
          const uint32_t q = lower_bound_u32(targets, fwd[u], hi, v);
          if (!SUPPORT || removed == nullptr || !removed[q]) {
            st = q + 1;
            len = hi - st;
            e1 = q;
          }
        }
        const uint32_t incl = wave_inclusive_scan(len);
        s_pos[lane] = incl;
        s_start[lane] = st;
        if (SUPPORT) s_edge[lane] = e1;
        const uint32_t total = __shfl(incl, GC_WAVE - 1);
        __syncthreads();
        for (uint32_t t = lane; t < total; t += GC_WAVE) {
          uint32_t l = 0, h = nv - 1;
          while (l < h) {
            const uint32_t mid = (l + h) >> 1;
            if (s_pos[mid] > t) h = mid;
            else l = mid + 1;
          }
          const uint32_t excl = l ? s_pos[l - 1] : 0;
          const uint32_t e2 = s_start[l] + (t - excl);
          if (SUPPORT && removed != nullptr && removed[e2]) continue;
          const uint32_t w = targets[e2];
          uint32_t e3 = WEDGE_EMPTY;
          if (in_table) {
            uint32_t slot = wedge_hash(w, shift_v);
            for (;;) {
              const uint32_t key = s_key[slot];
              if (key == w) {
                e3 = slot;
                break;
              }
              if (key == WEDGE_EMPTY) break;
              slot = (slot + 1) & mask_v;
            }
          } else {
            const uint32_t p = lower_bound_u32(targets, a_lo, a_lo + a, w);
            if (p < a_lo + a && targets[p] == w && !(SUPPORT && removed != nullptr && removed[p])) e3 = p;
          }
          if (e3 == WEDGE_EMPTY) continue;
          if (SUPPORT) {
            atomicAdd(&s_ucnt[l], 1u);
            atomicAdd(&support[e2], 1u);
            if (in_table) atomicAdd(&s_val[e3], 1u);
            else atomicAdd(&support[e3], 1u);
          } else {
            ++count;
          }
        }
        __syncthreads();
        if (SUPPORT) {  // the chunk's u -> v counts
          if (lane < nv && s_ucnt[lane] != 0) {
            atomicAdd(&support[s_edge[lane]], s_ucnt[lane]);
            s_ucnt[lane] = 0;
          }
          __syncthreads();
        }
      }
      if (SUPPORT && in_table) {  // the unit's v -> w counts
        for (uint32_t j = lane; j < size_v; j += GC_WAVE) {
          const uint32_t key = s_key[j];
          if (key != WEDGE_EMPTY && s_val[j] != 0)
            atomicAdd(&support[lower_bound_u32(targets, a_lo, a_lo + a, key)], s_val[j]);
        }
        __syncthreads();
      }
    }
  }
  if (!SUPPORT) {
    count = wave_sum(count);
    if (lane == 0 && count != 0) atomicAdd(result, count);
  }
}

}  // namespace gcd
