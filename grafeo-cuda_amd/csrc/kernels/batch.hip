// Batched triangle counting: many independent symmetric graphs, one launch.
//
// Each graph of the batch is relabelled by degree on the host and its
// degree-oriented DAG (u -> v for every edge with v later in the order) is
// stored with global vertex ids (graph base + local id) in one concatenated
// CSR. Triangles of a graph are counted as the reference counts them
// (crates/grafeo-cuda-algos/kernels/triangle_counting.cu:246-315): one per
// DAG edge u -> v and common out-neighbour w, i.e. sum over u of
// |{(v, w) : v, w in L(u), w in L(v)}|.
//
// The work is a list of tasks taken by the waves of one persistent launch
// from a global counter (heaviest first). A task is a run of rows of one
// graph, all of one kind:
//   - heavy rows (out-degree at or above the tuned threshold, graphs of at
//     most 65,536 vertices): the wave sets the bits of L(u) in its own 8 KB
//     LDS bitmap of the graph's local id space, streams every L(v), v in
//     L(u), 64 entries per step, and tests bits;
//   - light rows: the wave flattens the DAG edges of up to 64 rows at a time
//     and each lane merges the part of L(u) after v with L(v).
// Counts are summed per graph in u64 (one atomic per task), so the result
// does not depend on the task order, the tile shapes or the grid.
#include "device.hpp"

using namespace gcd;

#define BT_WAVES 4
#define BT_LOCAL_BITS 65536
#define BT_LOCAL_WORDS (BT_LOCAL_BITS / 32)
#define BT_HEAVY 0x80000000u

struct BatchTask {
  uint32_t graph;  // graph index, | BT_HEAVY for bitmap rows
  uint32_t begin;  // first position in `rows`
  uint32_t end;    // one past the last position
  uint32_t base;   // first global vertex id of the graph
};

// Clears the per-graph counts and the task counter (a kernel rather than a
// memset so that it can sit in a captured graph).
extern "C" __global__ __launch_bounds__(256) void bt_reset(unsigned long long* __restrict__ counts,
                                                           uint32_t num_graphs,
                                                           uint32_t* __restrict__ work) {
  for (uint64_t i = global_thread(); i < num_graphs; i += grid_threads()) counts[i] = 0;
  if (global_thread() == 0) *work = 0;
}

extern "C" __global__ __launch_bounds__(BT_WAVES * GC_WAVE) void bt_count(
    const uint32_t* __restrict__ dag_off, const uint32_t* __restrict__ dag_tgt,
    const uint32_t* __restrict__ rows, const BatchTask* __restrict__ tasks, uint32_t num_tasks,
    uint32_t* __restrict__ work, unsigned long long* __restrict__ counts) {
  __shared__ uint32_t s_bits_all[BT_WAVES][BT_LOCAL_WORDS];
  const uint32_t lane = lane_id();
  uint32_t* s_bits = s_bits_all[threadIdx.x / GC_WAVE];
  for (uint32_t j = lane; j < BT_LOCAL_WORDS; j += GC_WAVE) s_bits[j] = 0;
  wave_lds_sync();

  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(work, 1u);
    t = __shfl(t, 0);
    if (t >= num_tasks) break;
    const BatchTask task = tasks[t];
    const uint32_t base = task.base;
    unsigned long long local = 0;
    if (task.graph & BT_HEAVY) {
      for (uint32_t r = task.begin; r < task.end; ++r) {
        const uint32_t u = rows[r];
        const uint32_t a_lo = dag_off[u];
        const uint32_t a = dag_off[u + 1] - a_lo;
        const uint32_t* A = dag_tgt + a_lo;
        for (uint32_t j = lane; j < a; j += GC_WAVE) {
          const uint32_t x = A[j] - base;
          atomicOr(&s_bits[x >> 5], 1u << (x & 31));
        }
        wave_lds_sync();
        for (uint32_t cv = 0; cv < a; cv += GC_WAVE) {
          const uint32_t nv = min(static_cast<uint32_t>(GC_WAVE), a - cv);
          uint32_t st = 0, len = 0;
          if (lane < nv) {
            const uint32_t v = A[cv + lane];
            st = dag_off[v];
            len = dag_off[v + 1] - st;
          }
          for (uint32_t jv = 0; jv < nv; ++jv) {
            const uint32_t sj = __shfl(st, static_cast<int>(jv));
            const uint32_t lj = __shfl(len, static_cast<int>(jv));
            for (uint32_t k = lane; k < lj; k += GC_WAVE) {
              const uint32_t x = dag_tgt[sj + k] - base;
              local += (s_bits[x >> 5] >> (x & 31)) & 1u;
            }
          }
        }
        wave_lds_sync();
        for (uint32_t j = lane; j < a; j += GC_WAVE) s_bits[(A[j] - base) >> 5] = 0;
        wave_lds_sync();
      }
    } else {
      // light rows: the DAG edges of up to 64 rows flattened over the lanes
      for (uint32_t r0 = task.begin; r0 < task.end; r0 += GC_WAVE) {
        const uint32_t nr = min(static_cast<uint32_t>(GC_WAVE), task.end - r0);
        uint32_t u = 0, first = 0, len = 0;
        if (lane < nr) {
          u = rows[r0 + lane];
          first = dag_off[u];
          len = dag_off[u + 1] - first;
        }
        const uint32_t incl = wave_inclusive_scan(len);
        const uint32_t total = __shfl(incl, GC_WAVE - 1);
        // every lane runs every step (the shuffles need the whole wave)
        for (uint32_t k0 = 0; k0 < total; k0 += GC_WAVE) {
          const uint32_t k = k0 + lane;
          const uint32_t j = wave_upper_bound(incl, k < total ? k : total - 1);  // row slot of item k
          const uint32_t fj = __shfl(first, static_cast<int>(j));
          const uint32_t lj = __shfl(len, static_cast<int>(j));
          const uint32_t ij = __shfl(incl, static_cast<int>(j));
          if (k < total) {
            const uint32_t e = fj + (k - (ij - lj));
            const uint32_t row_end = fj + lj;
            const uint32_t v = dag_tgt[e];
            uint32_t p = e + 1, q = dag_off[v];
            const uint32_t q_end = dag_off[v + 1];
            while (p < row_end && q < q_end) {
              const uint32_t x = dag_tgt[p], y = dag_tgt[q];
              local += x == y;
              p += x <= y;
              q += y <= x;
            }
          }
        }
      }
    }
    local = wave_sum(local);
    if (lane == 0 && local) atomicAdd(&counts[task.graph & ~BT_HEAVY], local);
  }
}
