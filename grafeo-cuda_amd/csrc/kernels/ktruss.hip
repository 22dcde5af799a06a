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
//    once from its lowest vertex with the wave intersection of tc.hip and adds
//    1 to the support of its three (forward) edges. On a symmetric graph this
//    equals the reference's support except in round 1, where the reference's
//    merge also counts w = u when u has an active self-loop and w = v when v
//    has one (CHANGELOG.md:76-77); kt_dag_finish adds those two terms in round
//    1. Self-loops get support 0 and are removed in round 1, so later rounds
//    need no correction.
#include "device.hpp"

using namespace gcd;

#define KT_LDS_CAP 1024

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

// Adds 1 to the support of each edge of every triangle whose three DAG
// entries are active. Same wave layout and scheduling as tc_count.
extern "C" __global__ __launch_bounds__(64) void kt_dag_support(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint32_t* __restrict__ fwd, const uint8_t* __restrict__ removed, uint32_t n,
    uint32_t heavy, uint32_t light_chunk, uint32_t* __restrict__ counters,
    uint32_t* __restrict__ support) {
  __shared__ uint32_t s_a[KT_LDS_CAP];
  __shared__ uint32_t s_pos[GC_WAVE];
  __shared__ uint32_t s_start[GC_WAVE];
  __shared__ uint32_t s_range[2];
  const uint32_t lane = threadIdx.x;
  bool heavy_left = heavy > 0;

  for (;;) {
    if (lane == 0) {
      uint32_t kb = n, ke = n;
      if (heavy_left) {
        uint32_t k = atomicAdd(&counters[0], 1u);
        if (k < heavy) {
          kb = k;
          ke = k + 1;
        } else {
          heavy_left = false;
        }
      }
      if (!heavy_left && kb == n) {
        uint32_t c = atomicAdd(&counters[1], light_chunk);
        uint64_t b = static_cast<uint64_t>(heavy) + c;
        if (b < n) {
          kb = static_cast<uint32_t>(b);
          uint64_t e = b + light_chunk;
          ke = e < n ? static_cast<uint32_t>(e) : n;
        }
      }
      s_range[0] = kb;
      s_range[1] = ke;
    }
    __syncthreads();
    const uint32_t kb = s_range[0], ke = s_range[1];
    __syncthreads();
    if (kb >= n) break;

    for (uint32_t k = kb; k < ke; ++k) {
      const uint32_t u = n - 1 - k;
      const uint32_t a_lo = fwd[u];
      const uint32_t a = offsets[u + 1] - a_lo;
      if (a == 0) continue;
      const uint32_t* A = targets + a_lo;
      const bool in_lds = a <= KT_LDS_CAP;
      if (in_lds) {
        for (uint32_t j = lane; j < a; j += GC_WAVE) s_a[j] = A[j];
      }
      __syncthreads();
      for (uint32_t c = 0; c < a; c += GC_WAVE) {
        const uint32_t nv = min(static_cast<uint32_t>(GC_WAVE), a - c);
        uint32_t len = 0, st = 0;
        if (lane < nv && !removed[a_lo + c + lane]) {
          const uint32_t v = in_lds ? s_a[c + lane] : A[c + lane];
          st = fwd[v];
          len = offsets[v + 1] - st;
        }
        const uint32_t incl = wave_inclusive_scan(len);
        s_pos[lane] = incl;
        s_start[lane] = st;
        const uint32_t total = __shfl(incl, GC_WAVE - 1);
        __syncthreads();
        for (uint32_t t = lane; t < total; t += GC_WAVE) {
          uint32_t l = 0, h = nv - 1;
          while (l < h) {
            uint32_t mid = (l + h) >> 1;
            if (s_pos[mid] > t) h = mid;
            else l = mid + 1;
          }
          const uint32_t excl = l ? s_pos[l - 1] : 0;
          const uint32_t e2 = s_start[l] + (t - excl);
          if (removed[e2]) continue;
          const uint32_t w = targets[e2];
          uint32_t from = c + l + 1, to = a;
          if (in_lds) {
            while (from < to) {
              uint32_t mid = (from + to) >> 1;
              if (s_a[mid] < w) from = mid + 1;
              else to = mid;
            }
            if (!(from < a && s_a[from] == w)) continue;
          } else {
            while (from < to) {
              uint32_t mid = (from + to) >> 1;
              if (A[mid] < w) from = mid + 1;
              else to = mid;
            }
            if (!(from < a && A[from] == w)) continue;
          }
          const uint32_t e3 = a_lo + from;
          if (removed[e3]) continue;
          atomicAdd(&support[a_lo + c + l], 1u);
          atomicAdd(&support[e2], 1u);
          atomicAdd(&support[e3], 1u);
        }
        __syncthreads();
      }
    }
  }
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
