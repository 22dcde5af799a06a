// PageRank as a pull SpMV over the degree-ordered CSR.
//
// One iteration of crates/grafeo-cuda-algos/kernels/pagerank.cu:36-134
// (dangling sum, pull update, L1 difference) fused into one pass:
//   r'[v] = (1-d)/n + d * (sum_{u in N(v)} c[u] + D/n),   c[u] = r[u]/deg u (0 if deg u = 0)
// where D = sum of r over rows of degree 0. The pass reads offsets, targets
// and the gathered contributions c[u] once, writes r' and c' = r'/deg, and
// leaves per-workgroup FP64 partial sums of the next D and of |r' - r|;
// pr_finalize adds the partials in a fixed order (deterministic FP64) and
// updates the iteration state, so no host round trip is needed per iteration.
// Row sums are accumulated in FP64 (free: the pass is bound by HBM traffic),
// ranks and contributions are stored as FP32.
//
// Rows are sorted by degree, so the launch is cut into contiguous segments of
// similar degree and each segment uses a matching group of lanes per row (1
// to 64 lanes, or a whole 256-lane workgroup for rows of degree >= 1024):
// every lane group streams its row's targets with coalesced non-temporal loads
// and the hot contribution vector stays in L2 / Infinity Cache.
#include "device.hpp"

using namespace gcd;

#define PR_THREADS 256
#define PR_MAX_SEGMENTS 24
#define PR_EXACT 0x80000000u  // group flag: every row of the segment has degree (group & 0xFF)

struct PrSegment {
  uint32_t row_begin;
  uint32_t row_end;
  uint32_t group;        // lanes per row: 1, 2, ..., 64, 256 (one workgroup), or PR_EXACT | degree
  uint32_t block_begin;  // first workgroup of the segment
  uint32_t entry_begin;  // offsets[row_begin]
};

struct PrSegments {
  PrSegment seg[PR_MAX_SEGMENTS];
  uint32_t count;
  uint32_t total_blocks;
};

// dst[i] = src[i] for i < count (count >= 1) by the `threads` lanes of a
// workgroup: U loads per lane issued before their LDS stores, clamped rather
// than branched, so a pass costs one memory round trip (a plain loop waits for
// each load before its store: about 26 round trips for the hot sets here).
template <int U>
__device__ __forceinline__ void lds_fill_f32(float* dst, const float* __restrict__ src, uint32_t count,
                                             uint32_t threads) {
  for (uint32_t base = 0; base < count; base += threads * U) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = src[min(base + u * threads + threadIdx.x, count - 1)];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = base + u * threads + threadIdx.x;
      if (i < count) dst[i] = v[u];
    }
  }
}

struct PrState {
  double dangling;     // D of the current input ranks
  double diff;         // L1 change of the last iteration
  uint32_t iterations; // iterations performed
  uint32_t done;       // 1 once diff < tolerance
};

extern "C" __global__ __launch_bounds__(PR_THREADS) void pr_init(
    const uint32_t* __restrict__ offsets, uint32_t n, float r0, float* __restrict__ rank,
    float* __restrict__ contrib) {
  for (uint64_t v = global_thread(); v < n; v += grid_threads()) {
    uint32_t deg = offsets[v + 1] - offsets[v];
    rank[v] = r0;
    contrib[v] = deg ? r0 / static_cast<float>(deg) : 0.0f;
  }
}

// Finishes row `row` whose neighbour sum is `acc`.
__device__ __forceinline__ void pr_finish_row(uint32_t row, uint32_t deg, double acc, double base,
                                              double damping, const float* __restrict__ rank_in,
                                              float* __restrict__ rank_out,
                                              float* __restrict__ contrib_out, double& dangling,
                                              double& diff) {
  const float r = static_cast<float>(base + damping * acc);
  rank_out[row] = r;
  contrib_out[row] = deg ? r / static_cast<float>(deg) : 0.0f;
  if (deg == 0) dangling += static_cast<double>(r);
  diff += fabs(static_cast<double>(r) - static_cast<double>(rank_in[row]));
}

// Sum of contrib[targets[i]] for i = first, first + STRIDE, ... < end. Eight
// independent target loads are issued before their gathers so each lane keeps
// several memory requests in flight (the pass is bound by memory latency
// times requests in flight, not by arithmetic).
template <int STRIDE>
__device__ __forceinline__ double pr_row_sum(const uint32_t* __restrict__ targets,
                                             const float* __restrict__ contrib, uint32_t first,
                                             uint32_t end) {
  double acc = 0.0;
  uint32_t i = first;
  for (; i + 7 * STRIDE < end; i += 8 * STRIDE) {
    uint32_t t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = load_stream(targets + i + k * STRIDE);
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = contrib[t[k]];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += static_cast<double>(x[k]);
  }
  for (; i < end; i += STRIDE) acc += static_cast<double>(contrib[load_stream(targets + i)]);
  return acc;
}

template <int G>
__device__ __forceinline__ void pr_rows_group(const PrSegment& seg, uint32_t local_block,
                                              const uint32_t* __restrict__ offsets,
                                              const uint32_t* __restrict__ targets,
                                              const float* __restrict__ contrib_in,
                                              const float* __restrict__ rank_in,
                                              float* __restrict__ contrib_out,
                                              float* __restrict__ rank_out, double base,
                                              double damping, double& dangling, double& diff) {
  constexpr uint32_t kRowsPerBlock = PR_THREADS / G;
  const uint32_t row = seg.row_begin + local_block * kRowsPerBlock + threadIdx.x / G;
  const uint32_t g = threadIdx.x % G;
  const bool valid = row < seg.row_end;
  double acc = 0.0;
  uint32_t b = 0, e = 0;
  if (valid) {
    b = offsets[row];
    e = offsets[row + 1];
    acc = pr_row_sum<G>(targets, contrib_in, b + g, e);
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (valid && g == 0) {
    pr_finish_row(row, e - b, acc, base, damping, rank_in, rank_out, contrib_out, dangling, diff);
  }
}

// Rows of one exact degree D (rows are degree-sorted, so they are contiguous
// and row r's entries start at entry_begin + (r - row_begin) * D: no offsets
// load). Each thread takes R consecutive rows and issues all R*D target loads,
// then all R*D gathers, before summing.
template <int D, int R>
__device__ __forceinline__ void pr_rows_exact(const PrSegment& seg, uint32_t local,
                                              const uint32_t* __restrict__ targets,
                                              const float* __restrict__ contrib_in,
                                              const float* __restrict__ rank_in,
                                              float* __restrict__ contrib_out,
                                              float* __restrict__ rank_out, double base,
                                              double damping, double& dangling, double& diff) {
  const uint32_t first = seg.row_begin + (local * PR_THREADS + threadIdx.x) * R;
  if (first >= seg.row_end) return;
  const uint32_t rows = min(static_cast<uint32_t>(R), seg.row_end - first);
  const uint32_t e0 = seg.entry_begin + (first - seg.row_begin) * D;
  uint32_t t[R * D > 0 ? R * D : 1];
#pragma unroll
  for (int k = 0; k < R * D; ++k) t[k] = k / (D > 0 ? D : 1) < static_cast<int>(rows) ? load_stream(targets + e0 + k) : 0u;
  float x[R * D > 0 ? R * D : 1];
#pragma unroll
  for (int k = 0; k < R * D; ++k) x[k] = contrib_in[t[k]];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (r >= static_cast<int>(rows)) break;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) acc += static_cast<double>(x[r * D + j]);
    pr_finish_row(first + r, D, acc, base, damping, rank_in, rank_out, contrib_out, dangling, diff);
  }
}

__device__ __forceinline__ void pr_row_block(const PrSegment& seg, uint32_t local_block,
                                             const uint32_t* __restrict__ offsets,
                                             const uint32_t* __restrict__ targets,
                                             const float* __restrict__ contrib_in,
                                             const float* __restrict__ rank_in,
                                             float* __restrict__ contrib_out,
                                             float* __restrict__ rank_out, double base,
                                             double damping, double& dangling, double& diff,
                                             double* s_red) {
  const uint32_t row = seg.row_begin + local_block;
  const uint32_t b = offsets[row], e = offsets[row + 1];
  double acc = pr_row_sum<PR_THREADS>(targets, contrib_in, b + threadIdx.x, e);
  acc = wave_sum(acc);
  if (lane_id() == 0) s_red[threadIdx.x / GC_WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double total = 0.0;
    for (int w = 0; w < PR_THREADS / GC_WAVE; ++w) total += s_red[w];
    pr_finish_row(row, e - b, total, base, damping, rank_in, rank_out, contrib_out, dangling, diff);
  }
  __syncthreads();
}

// Workgroups [0, num_hub_chunks) each sum one chunk of a hub row (rows of
// degree >= the hub limit are cut into chunks of at most PR_HUB_CHUNK entries,
// heaviest rows first) into hub_partials; pr_hub_finish adds each row's chunks
// in order. The remaining workgroups take the segments of the other rows,
// highest-degree segment first, so no single long row sits at the tail.
#define PR_HUB_CHUNK 4096

extern "C" __global__ __launch_bounds__(PR_THREADS) __attribute__((amdgpu_waves_per_eu(8, 8))) void pr_spmv(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const float* __restrict__ contrib_in, const float* __restrict__ rank_in,
    float* __restrict__ contrib_out, float* __restrict__ rank_out, uint32_t n, double damping,
    PrSegments segs, const PrState* __restrict__ state, double* __restrict__ partials,
    const uint32_t* __restrict__ hub_chunks, uint32_t num_hub_chunks,
    double* __restrict__ hub_partials) {
  __shared__ double s_red[PR_THREADS / GC_WAVE];
  __shared__ double s_part[2][PR_THREADS / GC_WAVE];
  if (state->done) return;
  const double inv_n = 1.0 / static_cast<double>(n);
  const double base = (1.0 - damping) * inv_n + damping * state->dangling * inv_n;
  double dangling = 0.0, diff = 0.0;
  // A resident workgroup takes virtual blocks round-robin: hub chunks first,
  // then the segment tiles from the highest-degree segment down, so launch and
  // reduction overheads are paid once per workgroup instead of once per tile.
  const uint32_t num_virtual = num_hub_chunks + segs.total_blocks;
  for (uint32_t vb = blockIdx.x; vb < num_virtual; vb += gridDim.x) {
  if (vb < num_hub_chunks) {
    const uint32_t e0 = hub_chunks[2 * vb], e1 = hub_chunks[2 * vb + 1];
    double acc = pr_row_sum<PR_THREADS>(targets, contrib_in, e0 + threadIdx.x, e1);
    acc = wave_sum(acc);
    if (lane_id() == 0) s_red[threadIdx.x / GC_WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double total = 0.0;
      for (int w = 0; w < PR_THREADS / GC_WAVE; ++w) total += s_red[w];
      hub_partials[vb] = total;
    }
    __syncthreads();
    continue;
  }
  const uint32_t vblk = segs.total_blocks - 1 - (vb - num_hub_chunks);  // heaviest segment first
  uint32_t s = 0;
  while (s + 1 < segs.count && vblk >= segs.seg[s + 1].block_begin) ++s;
  const PrSegment seg = segs.seg[s];
  const uint32_t local = vblk - seg.block_begin;
  if (seg.group & PR_EXACT) {
#define PR_EXACT_CASE(D, R) \
  case D: pr_rows_exact<D, R>(seg, local, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    switch (seg.group & 0xFFu) {
      PR_EXACT_CASE(0, 4) PR_EXACT_CASE(1, 4) PR_EXACT_CASE(2, 4) PR_EXACT_CASE(3, 4)
      PR_EXACT_CASE(4, 2) PR_EXACT_CASE(5, 2) PR_EXACT_CASE(6, 2) PR_EXACT_CASE(7, 2)
      PR_EXACT_CASE(8, 1) PR_EXACT_CASE(9, 1) PR_EXACT_CASE(10, 1) PR_EXACT_CASE(11, 1)
      PR_EXACT_CASE(12, 1) PR_EXACT_CASE(13, 1) PR_EXACT_CASE(14, 1) PR_EXACT_CASE(15, 1)
      default: break;
    }
#undef PR_EXACT_CASE
  } else switch (seg.group) {
    case 1: pr_rows_group<1>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 2: pr_rows_group<2>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 4: pr_rows_group<4>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 8: pr_rows_group<8>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 16: pr_rows_group<16>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 32: pr_rows_group<32>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 64: pr_rows_group<64>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    default: pr_row_block(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff, s_red); break;
  }
  }
  dangling = wave_sum(dangling);
  diff = wave_sum(diff);
  const uint32_t wave = threadIdx.x / GC_WAVE;
  if (lane_id() == 0) {
    s_part[0][wave] = dangling;
    s_part[1][wave] = diff;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double d0 = 0.0, d1 = 0.0;
    for (int w = 0; w < PR_THREADS / GC_WAVE; ++w) {
      d0 += s_part[0][w];
      d1 += s_part[1][w];
    }
    partials[2 * static_cast<uint64_t>(blockIdx.x)] = d0;
    partials[2 * static_cast<uint64_t>(blockIdx.x) + 1] = d1;
  }
}

// Finishes the hub rows [hub_begin, n): row hub_begin + i adds its chunk sums
// hub_partials[chunk_first[i] .. chunk_first[i] + chunk_count[i]) in order.
// Per-workgroup partials of the dangling and L1 sums go to partials from
// `partial_base` on.
extern "C" __global__ __launch_bounds__(PR_THREADS) void pr_hub_finish(
    const uint32_t* __restrict__ offsets, const float* __restrict__ rank_in,
    float* __restrict__ contrib_out, float* __restrict__ rank_out, uint32_t n, double damping,
    const PrState* __restrict__ state, uint32_t hub_begin,
    const uint32_t* __restrict__ chunk_first, const uint32_t* __restrict__ chunk_count,
    const double* __restrict__ hub_partials, double* __restrict__ partials, uint32_t partial_base) {
  __shared__ double s_part[2][PR_THREADS / GC_WAVE];
  if (state->done) return;
  const double inv_n = 1.0 / static_cast<double>(n);
  const double base = (1.0 - damping) * inv_n + damping * state->dangling * inv_n;
  double dangling = 0.0, diff = 0.0;
  const uint32_t i = blockIdx.x * PR_THREADS + threadIdx.x;
  if (i < n - hub_begin) {
    const uint32_t row = hub_begin + i;
    double acc = 0.0;
    for (uint32_t c = 0; c < chunk_count[i]; ++c) acc += hub_partials[chunk_first[i] + c];
    pr_finish_row(row, offsets[row + 1] - offsets[row], acc, base, damping, rank_in, rank_out,
                  contrib_out, dangling, diff);
  }
  dangling = wave_sum(dangling);
  diff = wave_sum(diff);
  const uint32_t wave = threadIdx.x / GC_WAVE;
  if (lane_id() == 0) {
    s_part[0][wave] = dangling;
    s_part[1][wave] = diff;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double d0 = 0.0, d1 = 0.0;
    for (int w = 0; w < PR_THREADS / GC_WAVE; ++w) {
      d0 += s_part[0][w];
      d1 += s_part[1][w];
    }
    const uint64_t slot = partial_base + blockIdx.x;
    partials[2 * slot] = d0;
    partials[2 * slot + 1] = d1;
  }
}

// ---- persistent variant (the default): hot contributions in LDS -------------
//
// The rank mass of a power-law graph is concentrated: on the degree-ordered
// ids, the top 19K vertices receive 42% of all gathers (R-MAT scale 22).
// Each 1024-lane workgroup copies c[n-hot .. n) into LDS once per iteration
// and then takes row tiles round-robin over the workgroups, heaviest first
// (a tile is 1024/G rows of one segment, or one row of degree >= 1024 for the
// whole workgroup). Gathers of hot vertices become LDS reads, which removes
// them from the L1 request stream (one tag lookup per distinct line, the
// texture path 60-70% busy in pr_spmv) that bounds the pass. Two workgroups
// per CU keep 32 waves resident: with one workgroup and twice the hot set the
// pass is latency-bound and no faster than pr_spmv.

#define PRH_THREADS 1024

template <int STRIDE>
__device__ __forceinline__ double prh_row_sum(const uint32_t* __restrict__ targets,
                                              const float* __restrict__ contrib,
                                              const float* __restrict__ s_hot, uint32_t hot_base,
                                              uint32_t first, uint32_t end) {
  double acc = 0.0;
  uint32_t i = first;
  for (; i + 7 * STRIDE < end; i += 8 * STRIDE) {
    uint32_t t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = load_stream(targets + i + k * STRIDE);
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = t[k] >= hot_base ? s_hot[t[k] - hot_base] : contrib[t[k]];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += static_cast<double>(x[k]);
  }
  if (i < end) {
    // fewer than 8 entries per lane left: one guarded step (index clamped to
    // the row's last entry, value masked), so all its loads are in flight at
    // once instead of one target-then-gather round trip per entry
    uint32_t t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = load_stream(targets + min(i + k * STRIDE, end - 1));
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = t[k] >= hot_base ? s_hot[t[k] - hot_base] : contrib[t[k]];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += i + k * STRIDE < end ? static_cast<double>(x[k]) : 0.0;
  }
  return acc;
}

template <int G>
__device__ __forceinline__ void prh_tile_group(const PrSegment& seg, uint32_t local,
                                               const uint32_t* __restrict__ offsets,
                                               const uint32_t* __restrict__ targets,
                                               const float* __restrict__ contrib_in,
                                               const float* __restrict__ s_hot, uint32_t hot_base,
                                               const float* __restrict__ rank_in,
                                               float* __restrict__ contrib_out,
                                               float* __restrict__ rank_out, double base,
                                               double damping, double& dangling, double& diff) {
  constexpr uint32_t kRowsPerTile = PRH_THREADS / G;
  const uint32_t row = seg.row_begin + local * kRowsPerTile + threadIdx.x / G;
  const uint32_t g = threadIdx.x % G;
  const bool valid = row < seg.row_end;
  double acc = 0.0;
  uint32_t b = 0, e = 0;
  if (valid) {
    b = offsets[row];
    e = offsets[row + 1];
    acc = prh_row_sum<G>(targets, contrib_in, s_hot, hot_base, b + g, e);
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (valid && g == 0) {
    pr_finish_row(row, e - b, acc, base, damping, rank_in, rank_out, contrib_out, dangling, diff);
  }
}

extern "C" __global__ __launch_bounds__(PRH_THREADS) void pr_spmv_hot(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const float* __restrict__ contrib_in, const float* __restrict__ rank_in,
    float* __restrict__ contrib_out, float* __restrict__ rank_out, uint32_t n, double damping,
    PrSegments tiles, const PrState* __restrict__ state, uint32_t* __restrict__ work,
    uint32_t hot_count, double* __restrict__ partials) {
  extern __shared__ float s_hot[];
  __shared__ double s_red[PRH_THREADS / GC_WAVE];
  __shared__ double s_part[2][PRH_THREADS / GC_WAVE];
  if (state->done) return;
  const uint32_t hot_base = n - hot_count;
  if (hot_count > 0) lds_fill_f32<16>(s_hot, contrib_in + hot_base, hot_count, PRH_THREADS);
  const double inv_n = 1.0 / static_cast<double>(n);
  const double base = (1.0 - damping) * inv_n + damping * state->dangling * inv_n;
  double dangling = 0.0, diff = 0.0;
  __syncthreads();
  // tiles in round-robin order over the workgroups, heaviest first
  for (uint32_t k = blockIdx.x; k < tiles.total_blocks; k += gridDim.x) {
    const uint32_t t = tiles.total_blocks - 1 - k;
    uint32_t s = 0;
    while (s + 1 < tiles.count && t >= tiles.seg[s + 1].block_begin) ++s;
    const PrSegment seg = tiles.seg[s];
    const uint32_t local = t - seg.block_begin;
    switch (seg.group) {
      case 1: prh_tile_group<1>(seg, local, offsets, targets, contrib_in, s_hot, hot_base, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
      case 2: prh_tile_group<2>(seg, local, offsets, targets, contrib_in, s_hot, hot_base, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
      case 4: prh_tile_group<4>(seg, local, offsets, targets, contrib_in, s_hot, hot_base, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
      case 8: prh_tile_group<8>(seg, local, offsets, targets, contrib_in, s_hot, hot_base, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
      case 16: prh_tile_group<16>(seg, local, offsets, targets, contrib_in, s_hot, hot_base, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
      case 32: prh_tile_group<32>(seg, local, offsets, targets, contrib_in, s_hot, hot_base, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
      case 64: prh_tile_group<64>(seg, local, offsets, targets, contrib_in, s_hot, hot_base, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
      default: {
        // one row for the whole workgroup
        const uint32_t row = seg.row_begin + local;
        const uint32_t b = offsets[row], e = offsets[row + 1];
        double acc = prh_row_sum<PRH_THREADS>(targets, contrib_in, s_hot, hot_base, b + threadIdx.x, e);
        acc = wave_sum(acc);
        if (lane_id() == 0) s_red[threadIdx.x / GC_WAVE] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
          double total = 0.0;
          for (int w = 0; w < PRH_THREADS / GC_WAVE; ++w) total += s_red[w];
          pr_finish_row(row, e - b, total, base, damping, rank_in, rank_out, contrib_out, dangling, diff);
        }
        __syncthreads();
      }
    }
  }
  dangling = wave_sum(dangling);
  diff = wave_sum(diff);
  const uint32_t wave = threadIdx.x / GC_WAVE;
  if (lane_id() == 0) {
    s_part[0][wave] = dangling;
    s_part[1][wave] = diff;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double d0 = 0.0, d1 = 0.0;
    for (int w = 0; w < PRH_THREADS / GC_WAVE; ++w) {
      d0 += s_part[0][w];
      d1 += s_part[1][w];
    }
    partials[2 * static_cast<uint64_t>(blockIdx.x)] = d0;
    partials[2 * static_cast<uint64_t>(blockIdx.x) + 1] = d1;
  }
}

// Adds the per-workgroup partials in a fixed order and advances the state;
// also rearms the work counters for the next iteration. One 1024-lane
// workgroup: a lane adds the partials it strides over, the waves reduce in a
// fixed tree and lane 0 adds the 16 wave sums in order, so the result does
// not depend on timing (64 workgroups with a last-arriver fence took 9 us per
// iteration for the 512 partial pairs).
#define PR_FIN_THREADS 1024

extern "C" __global__ __launch_bounds__(PR_FIN_THREADS) void pr_finalize(
    PrState* __restrict__ state, const double* __restrict__ partials, uint32_t blocks,
    double tolerance, uint32_t* __restrict__ work) {
  __shared__ double s_a[PR_FIN_THREADS / GC_WAVE];
  __shared__ double s_b[PR_FIN_THREADS / GC_WAVE];
  if (state->done) return;
  double d0 = 0.0, d1 = 0.0;
  for (uint32_t i = threadIdx.x; i < blocks; i += PR_FIN_THREADS) {
    d0 += partials[2 * static_cast<uint64_t>(i)];
    d1 += partials[2 * static_cast<uint64_t>(i) + 1];
  }
  d0 = wave_sum(d0);
  d1 = wave_sum(d1);
  if (lane_id() == 0) {
    s_a[threadIdx.x / GC_WAVE] = d0;
    s_b[threadIdx.x / GC_WAVE] = d1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < PR_FIN_THREADS / GC_WAVE; ++w) {
      a += s_a[w];
      b += s_b[w];
    }
    state->dangling = a;
    state->diff = b;
    state->iterations += 1;
    if (b < tolerance) state->done = 1;
    work[0] = 0;
    work[1] = 0;
  }
}

// ---- push-pull variant for symmetric graphs (the default there) ------------
//
// On a symmetric CSR, y[h] = sum_{u in N(h)} c[u] is also the sum of c[u] over
// the rows u that hold an entry h. For the push set P (the top push_count ids,
// the hubs) the pass therefore pushes instead of pulling: every row u adds
// c[u] into an LDS accumulator of each of its entries h in P, and the hubs'
// own lists are read only for their entries inside P (which push c[h] into
// the other hubs). The long in-lists of the hubs, mostly low-degree ids that
// miss every cache, are never gathered, and a row outside P still pulls its
// entries: from the LDS copy of the top hot_count contributions (H, which
// contains P) or from global memory. Accumulators are unsigned 64-bit fixed
// point with 2^-62 resolution (every c <= 1 and every y <= sum of c <= 1, so
// nothing overflows), which makes the per-workgroup sums and their reduction
// exact and independent of the order of the atomics: the result is
// deterministic. Each workgroup writes its accumulators once; pr_push_finish
// adds them and finishes the hub rows.

#define PRP_THREADS 1024
#define PRP_FIX_SCALE 4611686018427387904.0  // 2^62

__device__ __forceinline__ unsigned long long pr_to_fixed(float c) {
  return static_cast<unsigned long long>(__double2ull_rn(static_cast<double>(c) * PRP_FIX_SCALE));
}

// Pull sums of R rows at once: lane entries first[r], first[r] + STRIDE, ...
// < end[r] of each row r, eight per row per step and all R rows' steps in
// flight together (R x 8 target loads, then R x 8 gathers), pushing cfix[r]
// (row r's own contribution) into s_push for every entry in P. A step past a
// row's last entry is masked (its loads read the row's last entry again).
template <int STRIDE, int R>
__device__ __forceinline__ void prp_rows_sum(const uint32_t* __restrict__ targets,
                                             const float* __restrict__ contrib,
                                             const float* __restrict__ s_hot, uint32_t hot_base,
                                             unsigned long long* __restrict__ s_push,
                                             uint32_t push_base, const unsigned long long (&cfix)[R],
                                             const uint32_t (&first)[R], const uint32_t (&end)[R],
                                             double (&acc)[R]) {
  uint32_t steps[R];
  uint32_t most = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    steps[r] = first[r] < end[r] ? (end[r] - first[r] + 8 * STRIDE - 1) / (8 * STRIDE) : 0u;
    most = max(most, steps[r]);
  }
  for (uint32_t s = 0; s < most; ++s) {
    uint32_t t[R][8];
    bool ok[R][8];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint32_t idx = first[r] + (s * 8 + k) * STRIDE;
        ok[r][k] = s < steps[r] && idx < end[r];
        t[r][k] = s < steps[r] ? load_stream(targets + min(idx, end[r] - 1)) : hot_base;
      }
    }
    float x[R][8];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int k = 0; k < 8; ++k) x[r][k] = t[r][k] >= hot_base ? s_hot[t[r][k] - hot_base] : contrib[t[r][k]];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (ok[r][k] && t[r][k] >= push_base) atomicAdd(s_push + (t[r][k] - push_base), cfix[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[r] += ok[r][k] ? static_cast<double>(x[r][k]) : 0.0;
    }
  }
}

// As pr_finish_row, with rank_in[row] already loaded (rin).
__device__ __forceinline__ void pr_finish_row_rin(uint32_t row, uint32_t deg, double acc, double base,
                                                  double damping, float rin, float* __restrict__ rank_out,
                                                  float* __restrict__ contrib_out, double& dangling,
                                                  double& diff) {
  const float r = static_cast<float>(base + damping * acc);
  rank_out[row] = r;
  contrib_out[row] = deg ? r / static_cast<float>(deg) : 0.0f;
  if (deg == 0) dangling += static_cast<double>(r);
  diff += fabs(static_cast<double>(r) - static_cast<double>(rin));
}

// One tile of a segment with G lanes per row: R slices of PRP_THREADS / G
// consecutive rows, lane group i of the workgroup taking row i of every
// slice. All rows' offsets, contributions and input ranks are loaded first.
template <int G, int R>
__device__ __forceinline__ void prp_tile_group(const PrSegment& seg, uint32_t local,
                                               const uint32_t* __restrict__ offsets,
                                               const uint32_t* __restrict__ targets,
                                               const float* __restrict__ contrib_in,
                                               const float* __restrict__ s_hot, uint32_t hot_base,
                                               unsigned long long* __restrict__ s_push,
                                               uint32_t push_base, const float* __restrict__ rank_in,
                                               float* __restrict__ contrib_out,
                                               float* __restrict__ rank_out, double base,
                                               double damping, double& dangling, double& diff) {
  constexpr uint32_t kRowsPerSlice = PRP_THREADS / G;
  const uint32_t g = threadIdx.x % G;
  uint32_t row[R], b[R], e[R], first[R];
  unsigned long long cfix[R];
  float rin[R];
  double acc[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    row[j] = seg.row_begin + (local * R + j) * kRowsPerSlice + threadIdx.x / G;
    b[j] = e[j] = 0;
    cfix[j] = 0ull;
    rin[j] = 0.0f;
    acc[j] = 0.0;
    if (row[j] < seg.row_end) {
      b[j] = offsets[row[j]];
      e[j] = offsets[row[j] + 1];
      cfix[j] = pr_to_fixed(contrib_in[row[j]]);
      if (g == 0) rin[j] = rank_in[row[j]];
    }
    first[j] = b[j] + g;
  }
  prp_rows_sum<G, R>(targets, contrib_in, s_hot, hot_base, s_push, push_base, cfix, first, e, acc);
#pragma unroll
  for (int j = 0; j < R; ++j) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) acc[j] += __shfl_xor(acc[j], o);
  }
#pragma unroll
  for (int j = 0; j < R; ++j) {
    if (row[j] < seg.row_end && g == 0) {
      pr_finish_row_rin(row[j], e[j] - b[j], acc[j], base, damping, rin[j], rank_out, contrib_out, dangling,
                        diff);
    }
  }
}

// Tiles: first `push_tiles` tiles of 16 hub rows each (one wave per row, the
// top rows first), each pushing c[h] into the hubs of its list inside P (from
// psplit[h - push_base], the first such entry); then the segments' tiles of
// the rows below P, heaviest first. A workgroup takes tiles round-robin.
extern "C" __global__ __launch_bounds__(PRP_THREADS) void pr_spmv_push(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const float* __restrict__ contrib_in, const float* __restrict__ rank_in,
    float* __restrict__ contrib_out, float* __restrict__ rank_out, uint32_t n, double damping,
    PrSegments tiles, const PrState* __restrict__ state, uint32_t hot_count, uint32_t push_count,
    const uint32_t* __restrict__ psplit, uint32_t push_tiles, uint32_t rows_per_group,
    unsigned long long* __restrict__ push_partials, double* __restrict__ partials) {
  extern __shared__ unsigned long long s_dyn[];
  unsigned long long* s_push = s_dyn;                          // push_count accumulators
  float* s_hot = reinterpret_cast<float*>(s_dyn + push_count);  // hot_count contributions
  __shared__ double s_red[PRP_THREADS / GC_WAVE];
  __shared__ double s_part[2][PRP_THREADS / GC_WAVE];
  if (state->done) return;
  const uint32_t hot_base = n - hot_count;
  const uint32_t push_base = n - push_count;
  lds_fill_f32<16>(s_hot, contrib_in + hot_base, hot_count, PRP_THREADS);
  for (uint32_t i = threadIdx.x; i < push_count; i += PRP_THREADS) s_push[i] = 0ull;
  const double inv_n = 1.0 / static_cast<double>(n);
  const double base = (1.0 - damping) * inv_n + damping * state->dangling * inv_n;
  double dangling = 0.0, diff = 0.0;
  __syncthreads();
  const uint32_t total = push_tiles + tiles.total_blocks;
  for (uint32_t k = blockIdx.x; k < total; k += gridDim.x) {
    if (k < push_tiles) {
      // hub row h (one wave): push c[h] into its list's hubs
      const uint32_t r = k * (PRP_THREADS / GC_WAVE) + threadIdx.x / GC_WAVE;
      if (r < push_count) {
        const uint32_t h = n - 1 - r;
        const unsigned long long cfix = pr_to_fixed(s_hot[h - hot_base]);
        const uint32_t e = offsets[h + 1];
        uint32_t i = psplit[h - push_base] + lane_id();
        for (; i + 7 * GC_WAVE < e; i += 8 * GC_WAVE) {
          uint32_t t[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) t[j] = load_stream(targets + i + j * GC_WAVE);
#pragma unroll
          for (int j = 0; j < 8; ++j) atomicAdd(s_push + (t[j] - push_base), cfix);
        }
        for (; i < e; i += GC_WAVE) atomicAdd(s_push + (load_stream(targets + i) - push_base), cfix);
      }
      continue;
    }
    const uint32_t t = tiles.total_blocks - 1 - (k - push_tiles);
    uint32_t s = 0;
    while (s + 1 < tiles.count && t >= tiles.seg[s + 1].block_begin) ++s;
    const PrSegment seg = tiles.seg[s];
    const uint32_t local = t - seg.block_begin;
#define PRP_ARGS seg, local, offsets, targets, contrib_in, s_hot, hot_base, s_push, push_base, rank_in, contrib_out, rank_out, base, damping, dangling, diff
#define PRP_GROUP(G)                                                        \
  do {                                                                      \
    /* R = 4 needs more than 128 VGPRs */                                  \
    if (rows_per_group >= 2) prp_tile_group<G, 2>(PRP_ARGS);                \
    else prp_tile_group<G, 1>(PRP_ARGS);                                    \
  } while (0)
    switch (seg.group) {
      case 1: PRP_GROUP(1); break;
      case 2: PRP_GROUP(2); break;
      case 4: PRP_GROUP(4); break;
      case 8: PRP_GROUP(8); break;
      case 16: PRP_GROUP(16); break;
      case 32: PRP_GROUP(32); break;
      case 64: PRP_GROUP(64); break;
      default: {
        // one row for the whole workgroup
        const uint32_t row = seg.row_begin + local;
        const uint32_t b = offsets[row], e = offsets[row + 1];
        const unsigned long long cfix[1] = {pr_to_fixed(contrib_in[row])};
        const uint32_t first[1] = {b + threadIdx.x}, end[1] = {e};
        double accs[1] = {0.0};
        prp_rows_sum<PRP_THREADS, 1>(targets, contrib_in, s_hot, hot_base, s_push, push_base, cfix, first, end, accs);
        double acc = wave_sum(accs[0]);
        if (lane_id() == 0) s_red[threadIdx.x / GC_WAVE] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
          double total_acc = 0.0;
          for (int w = 0; w < PRP_THREADS / GC_WAVE; ++w) total_acc += s_red[w];
          pr_finish_row(row, e - b, total_acc, base, damping, rank_in, rank_out, contrib_out, dangling, diff);
        }
        __syncthreads();
      }
    }
#undef PRP_GROUP
#undef PRP_ARGS
  }
  dangling = wave_sum(dangling);
  diff = wave_sum(diff);
  const uint32_t wave = threadIdx.x / GC_WAVE;
  if (lane_id() == 0) {
    s_part[0][wave] = dangling;
    s_part[1][wave] = diff;
  }
  __syncthreads();  // also: every push of this workgroup has landed
  unsigned long long* out = push_partials + static_cast<uint64_t>(blockIdx.x) * push_count;
  for (uint32_t i = threadIdx.x; i < push_count; i += PRP_THREADS) out[i] = s_push[i];
  if (threadIdx.x == 0) {
    double d0 = 0.0, d1 = 0.0;
    for (int w = 0; w < PRP_THREADS / GC_WAVE; ++w) {
      d0 += s_part[0][w];
      d1 += s_part[1][w];
    }
    partials[2 * static_cast<uint64_t>(blockIdx.x)] = d0;
    partials[2 * static_cast<uint64_t>(blockIdx.x) + 1] = d1;
  }
}

// psplit[i] = first entry of hub row n - push_count + i whose target is in P
// (rows are sorted): one lane per hub row, binary search.
extern "C" __global__ __launch_bounds__(256) void pr_push_split(const uint32_t* __restrict__ offsets,
                                                                const uint32_t* __restrict__ targets,
                                                                uint32_t n, uint32_t push_count,
                                                                uint32_t* __restrict__ psplit) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= push_count) return;
  const uint32_t push_base = n - push_count;
  const uint32_t h = push_base + i;
  psplit[i] = lower_bound_u32(targets, offsets[h], offsets[h + 1], push_base);
}

// Finishes the hub rows: y[h] = sum over the `parts` workgroups' accumulators
// (exact: integer sums), then rank, contribution, dangling and L1 terms as
// pr_finish_row. 16 hubs per 256-lane workgroup, 16 lanes per hub each adding
// every 16th workgroup's value. Per-workgroup FP64 partials of the dangling
// and L1 sums go to partials from `partial_base` on.
#define PRF_HUBS 16
extern "C" __global__ __launch_bounds__(256) void pr_push_finish(
    const uint32_t* __restrict__ offsets, const float* __restrict__ rank_in,
    float* __restrict__ contrib_out, float* __restrict__ rank_out, uint32_t n, double damping,
    const PrState* __restrict__ state, uint32_t push_count,
    const unsigned long long* __restrict__ push_partials, uint32_t parts, double* __restrict__ partials,
    uint32_t partial_base) {
  __shared__ unsigned long long s_sum[256];
  __shared__ double s_part[2][PRF_HUBS];
  if (state->done) return;
  const uint32_t hub = threadIdx.x % PRF_HUBS, j = threadIdx.x / PRF_HUBS;
  const uint32_t i = blockIdx.x * PRF_HUBS + hub;
  unsigned long long sum = 0ull;
  if (i < push_count) {
    for (uint32_t w = j; w < parts; w += 256 / PRF_HUBS) sum += push_partials[static_cast<uint64_t>(w) * push_count + i];
  }
  s_sum[threadIdx.x] = sum;
  __syncthreads();
  if (j == 0) {
    double dangling = 0.0, diff = 0.0;
    if (i < push_count) {
      unsigned long long y = 0ull;
      for (uint32_t q = 0; q < 256 / PRF_HUBS; ++q) y += s_sum[q * PRF_HUBS + hub];
      const double inv_n = 1.0 / static_cast<double>(n);
      const double base = (1.0 - damping) * inv_n + damping * state->dangling * inv_n;
      const uint32_t row = n - push_count + i;
      pr_finish_row(row, offsets[row + 1] - offsets[row], static_cast<double>(y) / PRP_FIX_SCALE, base,
                    damping, rank_in, rank_out, contrib_out, dangling, diff);
    }
    s_part[0][hub] = dangling;
    s_part[1][hub] = diff;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double d0 = 0.0, d1 = 0.0;
    for (int h = 0; h < PRF_HUBS; ++h) {
      d0 += s_part[0][h];
      d1 += s_part[1][h];
    }
    const uint64_t slot = partial_base + blockIdx.x;
    partials[2 * slot] = d0;
    partials[2 * slot + 1] = d1;
  }
}

// ---- push-pull with a nonzero-split pull (GC_PR_KERNEL=merge) --------------
//
// Same push set, LDS sets and hub rows as pr_spmv_push; only the pull over the
// rows below P is cut differently. Rows are in ascending degree order, so the
// rows of degree 0 come first and own no entry, and the rows [z, push_base)
// own the one run of entries [0, E), every row at least one. That run is cut
// into chunks of PRM_CHUNK = 512 entries: one wave per chunk and eight
// consecutive entries per lane (two 16-byte target loads), so every lane has
// eight useful gathers in flight whatever the row lengths (row tiles of G
// lanes per row fill 70% of their lane slots at R-MAT scale 22).
//
// A per-graph bitmap marks the entries that start a row (bit e set iff
// e = offsets[r]); chunk_row[c] is the row of the chunk's first entry minus
// one if that entry starts a row, so the row of entry e of the chunk is
// chunk_row[c] + (starts in [c * 512, e]). A lane sums each row's part of its
// eight entries (in reverse entry order); a segmented suffix scan across the
// wave adds the parts of the lanes that follow up to the next start, so a
// row that starts in the chunk and ends in it is finished by the lane that
// holds its start. The part before the chunk's first start (head) and the
// chunk's last row if it may run on (tail) go to head_partial, tail_partial
// and tail_row (PRM_NONE: the chunk holds no start); pr_merge_finish joins
// them in chunk order. Every sum has a fixed order, so the result is
// deterministic.

#define PRM_CHUNK 512
#define PRM_NONE 0xFFFFFFFFu

// starts (u32 words, zeroed): bit offsets[r] for every row r in [first, last).
extern "C" __global__ __launch_bounds__(256) void pr_merge_starts(const uint32_t* __restrict__ offsets,
                                                                  uint32_t first, uint32_t last,
                                                                  uint32_t* __restrict__ starts) {
  const uint64_t r = static_cast<uint64_t>(first) + global_thread();
  if (r >= last) return;
  const uint32_t e = offsets[r];
  atomicOr(starts + (e >> 5), 1u << (e & 31));
}

// chunk_row[c] for the chunks of [0, offsets[last]): the largest row r in
// [first, last) with offsets[r] <= c * 512, minus one if offsets[r] == c * 512
// (unsigned: first - 1 wraps when first = 0, and the popcounts wrap it back).
extern "C" __global__ __launch_bounds__(256) void pr_merge_chunks(const uint32_t* __restrict__ offsets,
                                                                  uint32_t first, uint32_t last,
                                                                  uint32_t chunks,
                                                                  uint32_t* __restrict__ chunk_row) {
  const uint64_t c = global_thread();
  if (c >= chunks) return;
  const uint32_t e0 = static_cast<uint32_t>(c) * PRM_CHUNK;
  uint32_t lo = first, hi = last - 1;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo + 1) >> 1);
    if (offsets[mid] <= e0) lo = mid;
    else hi = mid - 1;
  }
  chunk_row[c] = lo - (offsets[lo] == e0 ? 1u : 0u);
}

typedef uint32_t prm_u32x4 __attribute__((ext_vector_type(4)));

// One chunk on one wave (every lane calls it: it shuffles).
__device__ __forceinline__ void prm_chunk(uint32_t c, uint32_t entries, const uint32_t* __restrict__ targets,
                                          const uint8_t* __restrict__ starts, const uint32_t* __restrict__ chunk_row,
                                          const float* __restrict__ contrib_in, const float* __restrict__ s_hot,
                                          uint32_t hot_base, unsigned long long* __restrict__ s_push,
                                          uint32_t push_base, const float* __restrict__ rank_in,
                                          float* __restrict__ contrib_out, float* __restrict__ rank_out,
                                          double base, double damping, double& diff,
                                          double* __restrict__ head_partial, double* __restrict__ tail_partial,
                                          uint32_t* __restrict__ tail_row) {
  const uint32_t lane = lane_id();
  const uint32_t e = c * PRM_CHUNK + lane * 8;
  const uint32_t nv = e < entries ? min(entries - e, 8u) : 0u;
  uint32_t t[8];
  if (nv == 8) {
    const prm_u32x4* p = reinterpret_cast<const prm_u32x4*>(targets + e);
    const prm_u32x4 a = __builtin_nontemporal_load(p);
    const prm_u32x4 b = __builtin_nontemporal_load(p + 1);
    t[0] = a.x; t[1] = a.y; t[2] = a.z; t[3] = a.w;
    t[4] = b.x; t[5] = b.y; t[6] = b.z; t[7] = b.w;
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = static_cast<uint32_t>(k) < nv ? targets[e + k] : hot_base;
  }
  uint32_t bits = nv ? static_cast<uint32_t>(starts[e >> 3]) : 0u;
  bits &= (1u << nv) - 1u;
  float x[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = t[k] >= hot_base ? s_hot[t[k] - hot_base] : contrib_in[t[k]];
  // rows of the lane's entries: rbase + starts in [e, e + k]; the exclusive
  // prefix of the lanes' start counts (0..8) from four bit-slice ballots
  const uint32_t pc = __popc(bits);
  uint32_t below = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) below += lanes_below(__ballot((pc >> b) & 1u)) << b;
  const uint32_t rbase = chunk_row[c] + below;
  // each row's own contribution for its pushes (loaded at entry 0 and at every
  // start) and the input rank of every row that starts here, all in flight together
  float cst[8], rin[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint32_t rk = rbase + __popc(bits & ((2u << k) - 1u));
    const bool start = (bits >> k) & 1u;
    cst[k] = (k == 0 || start) ? contrib_in[rk] : 0.0f;
    rin[k] = start ? rank_in[rk] : 0.0f;
  }
  float cur = cst[0];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if ((bits >> k) & 1u) cur = cst[k];
    if (static_cast<uint32_t>(k) < nv && t[k] >= push_base) atomicAdd(s_push + (t[k] - push_base), pr_to_fixed(cur));
  }
  // the parts of the lane's rows, last entry first: seg[k] / len[k] = sum and
  // count of entries k .. (next start or the lane's end) - 1
  double seg[8];
  uint32_t len[8];
  double acc = 0.0;
  uint32_t cnt = 0;
#pragma unroll
  for (int k = 7; k >= 0; --k) {
    if (static_cast<uint32_t>(k) < nv) {
      acc += static_cast<double>(x[k]);
      ++cnt;
    }
    seg[k] = acc;
    len[k] = cnt;
    if ((bits >> k) & 1u) {
      acc = 0.0;
      cnt = 0;
    }
  }
  // segmented suffix scan over the lanes: C = sum of the heads (entries before
  // a lane's first start) from this lane up to and including the first lane at
  // or after it that holds a start; which lanes hold one comes from a ballot
  const uint64_t smask = __ballot(bits != 0u);
  const uint64_t from_me = smask >> lane;  // bit i: lane + i holds a start
  double C = acc;
#pragma unroll
  for (int o = 1; o < GC_WAVE; o <<= 1) {
    const double oc = __shfl_down(C, o);
    const uint64_t window = from_me & ((1ull << o) - 1ull);  // a start in [lane, lane + o)
    if (window == 0ull && lane + o < GC_WAVE) C += oc;
  }
  double nc = __shfl_down(C, 1);
  const uint64_t after = lane + 1 < GC_WAVE ? smask >> (lane + 1) : 0ull;
  const uint32_t next = lane + 1 + (after ? static_cast<uint32_t>(__builtin_ctzll(after)) : 0u);
  const uint32_t first_pos = bits ? static_cast<uint32_t>(__builtin_ctz(bits)) : 0u;
  const uint32_t next_pos = __shfl(first_pos, static_cast<int>(min(next, GC_WAVE - 1u)));
  if (lane == GC_WAVE - 1) nc = 0.0;
  if (lane == 0) {
    head_partial[c] = C;
    if (smask == 0ull) tail_row[c] = PRM_NONE;
  }
  const uint32_t last = bits ? 31u - __clz(bits) : 8u;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if ((bits >> k) & 1u) {
      const uint32_t row = rbase + __popc(bits & ((2u << k) - 1u));
      double total = seg[k];
      uint32_t deg = len[k];
      bool whole = true;
      if (static_cast<uint32_t>(k) == last) {
        total += nc;
        // the row runs to the first start of lane `next`
        deg = (8u * next + next_pos) - (8u * lane + last);
        whole = after != 0ull;
      }
      if (whole) {
        const float r = static_cast<float>(base + damping * total);
        rank_out[row] = r;
        contrib_out[row] = r / static_cast<float>(deg);
        diff += fabs(static_cast<double>(r) - static_cast<double>(rin[k]));
      } else {
        tail_row[c] = row;
        tail_partial[c] = total;
      }
    }
  }
}

// Hub row h (one wave): push c[h] into the hubs of its list inside P (from
// psplit, its first such entry), as pr_spmv_push does.
__device__ __forceinline__ void prm_hub_item(uint32_t h, const uint32_t* __restrict__ offsets,
                                             const uint32_t* __restrict__ targets,
                                             const uint32_t* __restrict__ psplit, const float* __restrict__ s_hot,
                                             uint32_t hot_base, unsigned long long* __restrict__ s_push,
                                             uint32_t push_base) {
  const unsigned long long cfix = pr_to_fixed(s_hot[h - hot_base]);
  const uint32_t e = offsets[h + 1];
  uint32_t i = psplit[h - push_base] + lane_id();
  for (; i + 7 * GC_WAVE < e; i += 8 * GC_WAVE) {
    uint32_t t[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) t[j] = load_stream(targets + i + j * GC_WAVE);
#pragma unroll
    for (int j = 0; j < 8; ++j) atomicAdd(s_push + (t[j] - push_base), cfix);
  }
  for (; i < e; i += GC_WAVE) atomicAdd(s_push + (load_stream(targets + i) - push_base), cfix);
}

// Rows of degree 0, run `item` of 512 (eight per lane): rank = base,
// contribution 0, dangling.
__device__ __forceinline__ void prm_zero_item(uint32_t item, uint32_t zero_rows, double base,
                                              const float* __restrict__ rank_in, float* __restrict__ rank_out,
                                              float* __restrict__ contrib_out, double& dangling, double& diff) {
  const uint32_t r0 = item * PRM_CHUNK + lane_id() * 8;
  const float r = static_cast<float>(base);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint32_t row = r0 + k;
    if (row < zero_rows) {
      rank_out[row] = r;
      contrib_out[row] = 0.0f;
      dangling += static_cast<double>(r);
      diff += fabs(static_cast<double>(r) - static_cast<double>(rank_in[row]));
    }
  }
}

// End of a push-pull workgroup: its dangling and L1 partials and its LDS
// accumulators go out (one slot per workgroup).
__device__ __forceinline__ void prm_end(uint32_t wave, const unsigned long long* __restrict__ s_push,
                                        uint32_t push_count, unsigned long long* __restrict__ push_partials,
                                        double* __restrict__ partials, double dangling, double diff,
                                        double (*s_part)[PRP_THREADS / GC_WAVE]) {
  dangling = wave_sum(dangling);
  diff = wave_sum(diff);
  if (lane_id() == 0) {
    s_part[0][wave] = dangling;
    s_part[1][wave] = diff;
  }
  __syncthreads();  // also: every push of this workgroup has landed
  unsigned long long* out = push_partials + static_cast<uint64_t>(blockIdx.x) * push_count;
  for (uint32_t i = threadIdx.x; i < push_count; i += PRP_THREADS) out[i] = s_push[i];
  if (threadIdx.x == 0) {
    double d0 = 0.0, d1 = 0.0;
    for (int w = 0; w < PRP_THREADS / GC_WAVE; ++w) {
      d0 += s_part[0][w];
      d1 += s_part[1][w];
    }
    partials[2 * static_cast<uint64_t>(blockIdx.x)] = d0;
    partials[2 * static_cast<uint64_t>(blockIdx.x) + 1] = d1;
  }
}

// Waves take items round-robin over the whole grid: the push_count hub rows
// (prm_hub_item), then the chunks, then the rows of degree 0 in runs of 512
// (prm_zero_item). Per-workgroup dangling and L1 partials and LDS
// accumulators go out as in pr_spmv_push.
extern "C" __global__ __launch_bounds__(PRP_THREADS) void pr_spmv_merge(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const float* __restrict__ contrib_in, const float* __restrict__ rank_in,
    float* __restrict__ contrib_out, float* __restrict__ rank_out, uint32_t n, double damping,
    const PrState* __restrict__ state, uint32_t hot_count, uint32_t push_count,
    const uint32_t* __restrict__ psplit, uint32_t entries, uint32_t chunks, uint32_t zero_rows,
    const uint8_t* __restrict__ starts, const uint32_t* __restrict__ chunk_row,
    double* __restrict__ head_partial, double* __restrict__ tail_partial, uint32_t* __restrict__ tail_row,
    unsigned long long* __restrict__ push_partials, double* __restrict__ partials) {
  extern __shared__ unsigned long long s_dyn[];
  unsigned long long* s_push = s_dyn;
  float* s_hot = reinterpret_cast<float*>(s_dyn + push_count);
  __shared__ double s_part[2][PRP_THREADS / GC_WAVE];
  if (state->done) return;
  const uint32_t hot_base = n - hot_count;
  const uint32_t push_base = n - push_count;
  lds_fill_f32<16>(s_hot, contrib_in + hot_base, hot_count, PRP_THREADS);
  for (uint32_t i = threadIdx.x; i < push_count; i += PRP_THREADS) s_push[i] = 0ull;
  const double inv_n = 1.0 / static_cast<double>(n);
  const double base = (1.0 - damping) * inv_n + damping * state->dangling * inv_n;
  double dangling = 0.0, diff = 0.0;
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / GC_WAVE);
  const uint32_t waves = gridDim.x * (PRP_THREADS / GC_WAVE);
  const uint32_t zero_items = (zero_rows + PRM_CHUNK - 1) / PRM_CHUNK;
  const uint32_t items = push_count + chunks + zero_items;
  for (uint32_t item = blockIdx.x * (PRP_THREADS / GC_WAVE) + wave; item < items; item += waves) {
    if (item < push_count) {
      prm_hub_item(n - 1 - item, offsets, targets, psplit, s_hot, hot_base, s_push, push_base);
    } else if (item < push_count + chunks) {
      prm_chunk(item - push_count, entries, targets, starts, chunk_row, contrib_in, s_hot, hot_base, s_push,
                push_base, rank_in, contrib_out, rank_out, base, damping, diff, head_partial, tail_partial,
                tail_row);
    } else {
      prm_zero_item(item - push_count - chunks, zero_rows, base, rank_in, rank_out, contrib_out, dangling, diff);
    }
  }
  prm_end(wave, s_push, push_count, push_partials, partials, dangling, diff, s_part);
}

// The hub rows as pr_push_finish (blocks [0, hub_blocks), 16 hubs each), then
// one lane per chunk (blocks from hub_blocks on, 256 chunks each): a chunk's
// tail row gets the heads of the chunks after it, up to and including the
// first chunk that holds a start, and is finished there. Per-block dangling
// and L1 partials go to partials from partial_base on.
extern "C" __global__ __launch_bounds__(256) void pr_merge_finish(
    const uint32_t* __restrict__ offsets, const float* __restrict__ rank_in,
    float* __restrict__ contrib_out, float* __restrict__ rank_out, uint32_t n, double damping,
    const PrState* __restrict__ state, uint32_t push_count,
    const unsigned long long* __restrict__ push_partials, uint32_t parts, uint32_t hub_blocks,
    uint32_t chunks, const double* __restrict__ head_partial, const double* __restrict__ tail_partial,
    const uint32_t* __restrict__ tail_row, double* __restrict__ partials, uint32_t partial_base) {
  __shared__ unsigned long long s_sum[256];
  __shared__ double s_part[2][256 / GC_WAVE];
  if (state->done) return;
  const double inv_n = 1.0 / static_cast<double>(n);
  const double base = (1.0 - damping) * inv_n + damping * state->dangling * inv_n;
  double dangling = 0.0, diff = 0.0;
  if (blockIdx.x < hub_blocks) {
    const uint32_t hub = threadIdx.x % PRF_HUBS, j = threadIdx.x / PRF_HUBS;
    const uint32_t i = blockIdx.x * PRF_HUBS + hub;
    unsigned long long sum = 0ull;
    if (i < push_count) {
      for (uint32_t w = j; w < parts; w += 256 / PRF_HUBS) sum += push_partials[static_cast<uint64_t>(w) * push_count + i];
    }
    s_sum[threadIdx.x] = sum;
    __syncthreads();
    if (j == 0 && i < push_count) {
      unsigned long long y = 0ull;
      for (uint32_t q = 0; q < 256 / PRF_HUBS; ++q) y += s_sum[q * PRF_HUBS + hub];
      const uint32_t row = n - push_count + i;
      pr_finish_row(row, offsets[row + 1] - offsets[row], static_cast<double>(y) / PRP_FIX_SCALE, base, damping,
                    rank_in, rank_out, contrib_out, dangling, diff);
    }
  } else {
    const uint32_t c = (blockIdx.x - hub_blocks) * 256 + threadIdx.x;
    if (c < chunks) {
      const uint32_t row = tail_row[c];
      if (row != PRM_NONE) {
        double total = tail_partial[c];
        for (uint32_t k = c + 1; k < chunks; ++k) {
          total += head_partial[k];
          if (tail_row[k] != PRM_NONE) break;
        }
        pr_finish_row(row, offsets[row + 1] - offsets[row], total, base, damping, rank_in, rank_out,
                      contrib_out, dangling, diff);
      }
    }
  }
  dangling = wave_sum(dangling);
  diff = wave_sum(diff);
  if (lane_id() == 0) {
    s_part[0][threadIdx.x / GC_WAVE] = dangling;
    s_part[1][threadIdx.x / GC_WAVE] = diff;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double d0 = 0.0, d1 = 0.0;
    for (int w = 0; w < 256 / GC_WAVE; ++w) {
      d0 += s_part[0][w];
      d1 += s_part[1][w];
    }
    const uint64_t slot = partial_base + blockIdx.x;
    partials[2 * slot] = d0;
    partials[2 * slot + 1] = d1;
  }
}

// ---- binned cold entries (symmetric graphs, the default there) -------------
//
// In the push-pull iteration the entries of the rows below P whose target lies
// below H (the cold entries: 47 M of the 128 M entries at R-MAT scale 22) are
// random 4-byte gathers from global memory, and those run at about 190 G/s
// whatever cache level holds the line: the L1 looks up one tag per distinct
// line, up to 64 per gather instruction. Here c[t] is propagated instead, along
// a layout fixed once per graph:
//   pr_bin_scatter: the cold sources (rows z .. H-1, in groups of 64
//     consecutive ids; entry j of source z + 64 g + lane sits in slot row
//     (g, j) of a static array of 16-bit LDS slots, eight slot rows to one
//     16-byte load per lane) are cut into sub-ranges of consecutive slot rows
//     holding at most `stage` real entries. A workgroup stores the c[t] of
//     each entry of its sub-range into that entry's LDS slot (the entries
//     ordered by destination bin), then copies each bin's run to its place in
//     the bin's part of the value buffer.
//   pr_bin_accum: one workgroup per bin (a range of at most 16,384 rows and
//     about 64 K values) adds each value into the 64-bit fixed-point LDS
//     accumulator of its row (a static 16-bit row index per value) and writes
//     every row's sum of cold contributions.
//   pr_spmv_bin: the nonzero-split pull of pr_spmv_merge over the other
//     entries only, stored as 16-bit offsets from H (a row with none holds one
//     offset equal to hot_count, which reads a zero), each row adding its cold
//     sum. Hub rows, pushes and the rows of degree 0 are as in pr_spmv_merge.
// Per cold entry an iteration moves 2 (slot) + 4 (value stored) + 4 (value
// read) + 2 (row index) bytes in coalesced streams instead of one gather; the
// fixed-point sums are exact, so the result stays deterministic.

#define PRB_THREADS 1024
#define PRB_WAVES (PRB_THREADS / GC_WAVE)

typedef uint16_t prb_u16x8 __attribute__((ext_vector_type(8)));
typedef float prb_f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t prb_u32x2 __attribute__((ext_vector_type(2)));

// Raw buffer stores whose lanes that must not store take an offset past the
// buffer's end, which the hardware drops: the store is issued by every lane
// with no branch around it, so the count of memory operations in flight is
// the same on every path and the waits for earlier loads stay exact (the
// vector memory counter counts stores too). Buffers stay below 2 GB.
#define PRB_OOB 0x80000000u
#define PRB_RSRC_FLAGS 0x00020000  // raw buffer, 32-bit data format (gfx9)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t prb_rsrc(void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(base, 0, static_cast<int>(min(bytes, PRB_OOB)), PRB_RSRC_FLAGS);
}

__device__ __forceinline__ void prb_st32(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, bool on, uint32_t v) {
  __builtin_amdgcn_raw_buffer_store_b32(v, r, static_cast<int>(on ? byte_off : PRB_OOB), 0, 0);
}

__device__ __forceinline__ void prb_st64(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, bool on, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(prb_u32x2, v), r,
                                        static_cast<int>(on ? byte_off : PRB_OOB), 0, 0);
}

// The output buffers of pr_spmv_bin's chunks.
struct PrbOut {
  __amdgpu_buffer_rsrc_t rank, contrib, head, tail, tail_row;
};

// Workgroup s: slot blocks [sub[s], sub[s+1]) (eight slot rows of one group
// each: block b holds, for lane l, the slots of its rows at slots[64 b + l];
// block_group[b] its group); then runs [run_first[s], run_first[s+1]): run r
// copies staging [info & 0xFFFF, + (info >> 16)) to vals[run_dst[r] ...].
// Dynamic LDS: stage + 1 floats (slot `stage` takes the padding). Each wave
// takes a contiguous share of the blocks, PRB_BATCH blocks per step with the
// groups of the next step loaded in the same round trip, and a contiguous
// share of the runs, whose descriptors it loads 64 at a time.
#define PRB_BATCH 4
extern "C" __global__ __launch_bounds__(PRB_THREADS) void pr_bin_scatter(
    const float* __restrict__ contrib_in, uint32_t src_base, uint32_t src_end,
    const prb_u16x8* __restrict__ slots, const uint32_t* __restrict__ block_group,
    const uint32_t* __restrict__ sub, const uint32_t* __restrict__ run_first,
    const uint32_t* __restrict__ run_info, const uint32_t* __restrict__ run_dst, float* __restrict__ vals,
    const PrState* __restrict__ state) {
  extern __shared__ float s_stage[];
  if (state->done) return;
  const uint32_t s = blockIdx.x;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / GC_WAVE);
  const uint32_t lane = lane_id();
  const uint32_t b0 = sub[s], b1 = sub[s + 1];
  const uint32_t per = (b1 - b0 + PRB_WAVES - 1) / PRB_WAVES;
  const uint32_t w0 = min(b1, b0 + wave * per), w1 = min(b1, w0 + per);
  uint32_t g[PRB_BATCH];
#pragma unroll
  for (int q = 0; q < PRB_BATCH; ++q) g[q] = w0 + q < w1 ? block_group[w0 + q] : 0u;
  for (uint32_t blk = w0; blk < w1; blk += PRB_BATCH) {
    prb_u16x8 sl[PRB_BATCH];
    float c[PRB_BATCH];
#pragma unroll
    for (int q = 0; q < PRB_BATCH; ++q) {
      if (blk + q < w1) {
        const uint32_t t = src_base + g[q] * GC_WAVE + lane;
        c[q] = t < src_end ? contrib_in[t] : 0.0f;
        sl[q] = __builtin_nontemporal_load(slots + static_cast<uint64_t>(blk + q) * GC_WAVE + lane);
      }
    }
#pragma unroll
    for (int q = 0; q < PRB_BATCH; ++q) {
      const uint32_t nb = blk + PRB_BATCH + q;
      g[q] = nb < w1 ? block_group[nb] : 0u;
    }
#pragma unroll
    for (int q = 0; q < PRB_BATCH; ++q) {
      if (blk + q < w1) {
#pragma unroll
        for (int i = 0; i < 8; ++i) s_stage[sl[q][i]] = c[q];
      }
    }
  }
  __syncthreads();
  const uint32_t r0 = run_first[s], r1 = run_first[s + 1];
  const uint32_t rper = (r1 - r0 + PRB_WAVES - 1) / PRB_WAVES;
  const uint32_t q0 = min(r1, r0 + wave * rper), q1 = min(r1, q0 + rper);
  for (uint32_t at = q0; at < q1; at += GC_WAVE) {
    const uint32_t mine = at + lane;
    const uint32_t info = mine < q1 ? run_info[mine] : 0u;
    const uint32_t dst = mine < q1 ? run_dst[mine] : 0u;
    const uint32_t cnt = min(static_cast<uint32_t>(GC_WAVE), q1 - at);
#pragma unroll 4
    for (uint32_t j = 0; j < cnt; ++j) {
      const uint32_t inf = __builtin_amdgcn_readlane(info, j);
      const uint32_t d = __builtin_amdgcn_readlane(dst, j);
      const uint32_t from = inf & 0xFFFFu, len = inf >> 16;
      for (uint32_t i = lane; i < len; i += GC_WAVE) vals[d + i] = s_stage[from + i];
    }
  }
}

// Workgroup b: rows [bin_row[b], bin_row[b+1]) and values [bin_val[b],
// bin_val[b+1]) (multiples of eight; padding values are 0 with row index
// equal to the bin's row count). cold_sum[row - row_base] = sum of the row's
// values. Dynamic LDS: (rows + 1) u64.
extern "C" __global__ __launch_bounds__(PRB_THREADS) void pr_bin_accum(
    const float* __restrict__ vals, const uint16_t* __restrict__ uloc, const uint32_t* __restrict__ bin_row,
    const uint32_t* __restrict__ bin_val, uint32_t row_base, double* __restrict__ cold_sum,
    const PrState* __restrict__ state) {
  extern __shared__ unsigned long long s_acc[];
  if (state->done) return;
  const uint32_t b = blockIdx.x;
  const uint32_t row0 = bin_row[b], rows = bin_row[b + 1] - row0;
  const uint32_t v0 = bin_val[b], v1 = bin_val[b + 1];
  for (uint32_t i = threadIdx.x; i <= rows; i += PRB_THREADS) s_acc[i] = 0ull;
  __syncthreads();
  constexpr uint32_t kStep = PRB_THREADS * 8;
  uint32_t p = v0 + threadIdx.x * 8;
  // two groups of eight values per lane in flight
  for (; p + kStep < v1; p += 2 * kStep) {
    const prb_f32x4 a0 = __builtin_nontemporal_load(reinterpret_cast<const prb_f32x4*>(vals + p));
    const prb_f32x4 a1 = __builtin_nontemporal_load(reinterpret_cast<const prb_f32x4*>(vals + p + 4));
    const prb_u16x8 u0 = __builtin_nontemporal_load(reinterpret_cast<const prb_u16x8*>(uloc + p));
    const prb_f32x4 c0 = __builtin_nontemporal_load(reinterpret_cast<const prb_f32x4*>(vals + p + kStep));
    const prb_f32x4 c1 = __builtin_nontemporal_load(reinterpret_cast<const prb_f32x4*>(vals + p + kStep + 4));
    const prb_u16x8 u1 = __builtin_nontemporal_load(reinterpret_cast<const prb_u16x8*>(uloc + p + kStep));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      atomicAdd(s_acc + u0[k], pr_to_fixed(a0[k]));
      atomicAdd(s_acc + u0[k + 4], pr_to_fixed(a1[k]));
      atomicAdd(s_acc + u1[k], pr_to_fixed(c0[k]));
      atomicAdd(s_acc + u1[k + 4], pr_to_fixed(c1[k]));
    }
  }
  if (p < v1) {
    const prb_f32x4 a0 = __builtin_nontemporal_load(reinterpret_cast<const prb_f32x4*>(vals + p));
    const prb_f32x4 a1 = __builtin_nontemporal_load(reinterpret_cast<const prb_f32x4*>(vals + p + 4));
    const prb_u16x8 u0 = __builtin_nontemporal_load(reinterpret_cast<const prb_u16x8*>(uloc + p));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      atomicAdd(s_acc + u0[k], pr_to_fixed(a0[k]));
      atomicAdd(s_acc + u0[k + 4], pr_to_fixed(a1[k]));
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < rows; i += PRB_THREADS)
    cold_sum[row0 - row_base + i] = static_cast<double>(s_acc[i]) / PRP_FIX_SCALE;
}

// A chunk's loads: the lane's eight 16-bit entries and its row-start byte,
// and the data of row R0 + lane (R0 = chunk_row[c], passed in): its own
// contribution, input rank, cold sum and degree, which serve the chunk's
// first 64 rows. The stream and its start bits are padded to whole chunks
// (entries hot_count, which read a zero, and no start) and the rows are
// clamped to the pull rows, so every load is in bounds and none needs a
// branch: a branch would make the waits for these loads in flight
// conservative.
struct PrbChunkIn {
  prb_u16x8 t;
  uint32_t bits;
  uint32_t row0;
  float cst, rin;
  double cold;
  uint32_t off0, off1;
};

__device__ __forceinline__ PrbChunkIn prb_load(uint32_t c, uint32_t row0, const uint16_t* __restrict__ hot16,
                                               const uint8_t* __restrict__ starts,
                                               const uint32_t* __restrict__ offsets,
                                               const double* __restrict__ cold_sum,
                                               const float* __restrict__ contrib_in,
                                               const float* __restrict__ rank_in, uint32_t row_base,
                                               uint32_t row_last) {
  PrbChunkIn in;
  const uint32_t lane = lane_id();
  const uint32_t e = c * PRM_CHUNK + lane * 8;
  in.t = __builtin_nontemporal_load(reinterpret_cast<const prb_u16x8*>(hot16 + e));
  in.bits = starts[e >> 3];
  in.row0 = row0;
  const uint32_t row = min(row0 + lane, row_last);  // R0 = -1 (first row starts chunk 0) clamps too
  in.cst = contrib_in[row];
  in.rin = rank_in[row];
  in.cold = cold_sum[max(row, row_base) - row_base];
  in.off0 = offsets[row];
  in.off1 = offsets[row + 1];
  return in;
}

// One chunk of the 16-bit stream on one wave (every lane calls it: it
// shuffles), from its loads. The segment sums are those of prm_chunk, with
// entries t = id - H (t < hot_count: s_hot[t]; t == hot_count: the zero of a
// row without such entries). The chunk's rows are R0 + rel, rel = 0 .. S (S
// row starts in the chunk; rel 0 is the row running in from the chunk
// before, rel S the tail, which may run on). Instead of loading the data of
// each start's row where the start lies (one dependent load per start), the
// wave loads rows R0 + rel0 + lane for pages of 64 rows at once: their own
// contributions (passed to the entries for the pushes with a lane shuffle),
// input ranks, degrees and cold sums. The lanes holding starts stage their
// rows' totals in the wave's 64 LDS doubles (s_tot), and the row lanes finish
// rows 1 .. S-1 and hand the tail's partial (with its cold sum) to
// pr_merge_finish.
__device__ __forceinline__ void prb_chunk(uint32_t c, const PrbChunkIn& in, const uint32_t* __restrict__ offsets,
                                          const double* __restrict__ cold_sum, uint32_t row_base,
                                          const float* __restrict__ contrib_in, const float* __restrict__ s_hot,
                                          uint32_t hot_count, unsigned long long* __restrict__ s_push,
                                          uint32_t push_off, const float* __restrict__ rank_in,
                                          const PrbOut& out, double base, double damping, double& diff,
                                          double* __restrict__ s_tot) {
  const uint32_t lane = lane_id();
  const uint32_t bits = in.bits;
  uint32_t t[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) t[k] = in.t[k];
  float x[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = s_hot[t[k]];
  const uint32_t pc = __popc(bits);
  uint32_t below = 0, starts_total = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const uint64_t bb = __ballot((pc >> b) & 1u);
    below += lanes_below(bb) << b;
    starts_total += static_cast<uint32_t>(__popcll(bb)) << b;
  }
  const uint32_t S = __builtin_amdgcn_readfirstlane(starts_total);
  // the parts of the lane's rows, last entry first: seg[k] = sum of entries
  // k .. (next start or the lane's end) - 1
  double seg[8];
  double acc = 0.0;
#pragma unroll
  for (int k = 7; k >= 0; --k) {
    acc += static_cast<double>(x[k]);  // entries past the stream read the zero
    seg[k] = acc;
    if ((bits >> k) & 1u) acc = 0.0;
  }
  const uint64_t smask = __ballot(bits != 0u);
  const uint64_t from_me = smask >> lane;
  double C = acc;
#pragma unroll
  for (int o = 1; o < GC_WAVE; o <<= 1) {
    const double oc = __shfl_down(C, o);
    const uint64_t window = from_me & ((1ull << o) - 1ull);
    if (window == 0ull && lane + o < GC_WAVE) C += oc;
  }
  double nc = __shfl_down(C, 1);
  if (lane == GC_WAVE - 1) nc = 0.0;
  prb_st64(out.head, c * 8u, lane == 0, C);
  prb_st32(out.tail_row, c * 4u, lane == 0 && smask == 0ull, PRM_NONE);
  const uint32_t last = bits ? 31u - __clz(bits) : 8u;
  uint32_t relk[8];  // row of entry k, relative to R0
#pragma unroll
  for (int k = 0; k < 8; ++k) relk[k] = below + __popc(bits & ((2u << k) - 1u));
  const uint32_t R0 = in.row0;
  // rows R0 + rel0 + lane: the first 64 from the chunk's loads, further pages (chunks
  // of more than 63 rows) loaded here
  auto page = [&](uint32_t rel0, uint32_t row, float cst_l, float rin_l, double cold_l, uint32_t deg_l) {
    const uint32_t rel = rel0 + lane;
    // each row lane converts its row's contribution once; the entries take it by shuffle
    const unsigned long long cfix_l = pr_to_fixed(cst_l);
    const uint32_t cfix_lo = static_cast<uint32_t>(cfix_l), cfix_hi = static_cast<uint32_t>(cfix_l >> 32);
    const uint32_t push_count = hot_count - push_off;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int src = static_cast<int>((relk[k] - rel0) & (GC_WAVE - 1));
      const uint32_t lo = __shfl(cfix_lo, src), hi = __shfl(cfix_hi, src);
      if (relk[k] - rel0 < GC_WAVE && t[k] - push_off < push_count)
        atomicAdd(s_push + (t[k] - push_off), (static_cast<unsigned long long>(hi) << 32) | lo);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (((bits >> k) & 1u) && relk[k] - rel0 < GC_WAVE)
        s_tot[relk[k] - rel0] = seg[k] + (static_cast<uint32_t>(k) == last ? nc : 0.0);
    }
    wave_lds_sync();
    const bool fin = rel >= 1 && rel < S;  // rows that end in the chunk
    const bool tail = rel >= 1 && rel == S;
    const double total = s_tot[lane] + cold_l;
    const float r = static_cast<float>(base + damping * total);
    prb_st32(out.rank, row * 4u, fin, __builtin_bit_cast(uint32_t, r));
    prb_st32(out.contrib, row * 4u, fin, __builtin_bit_cast(uint32_t, r / static_cast<float>(deg_l)));
    diff += fin ? fabs(static_cast<double>(r) - static_cast<double>(rin_l)) : 0.0;
    prb_st32(out.tail_row, c * 4u, tail, row);
    prb_st64(out.tail, c * 8u, tail, total);
    wave_lds_sync();
  };
  page(0, R0 + lane, in.cst, in.rin, in.cold, in.off1 - in.off0);
  for (uint32_t rel0 = GC_WAVE; rel0 <= S; rel0 += GC_WAVE) {
    const uint32_t row = R0 + min(rel0 + lane, S);  // lanes past the last row repeat it
    page(rel0, row, contrib_in[row], rank_in[row], cold_sum[row - row_base], offsets[row + 1] - offsets[row]);
  }
}

// Hub piece (one wave): up to 512 entries [b, min(b + 512, end of row h)) of
// hub row h that point into P, each taking c[h] (pushes into the LDS
// accumulators, as prm_hub_item); eight loads per lane in flight, clamped to
// the piece.
__device__ __forceinline__ void prb_hub_piece(uint32_t h, uint32_t b, const uint32_t* __restrict__ offsets,
                                              const uint32_t* __restrict__ targets, const float* __restrict__ s_hot,
                                              uint32_t hot_base, unsigned long long* __restrict__ s_push,
                                              uint32_t push_base) {
  const unsigned long long cfix = pr_to_fixed(s_hot[h - hot_base]);
  const uint32_t e = min(b + PRM_CHUNK, offsets[h + 1]);
  uint32_t t[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) t[j] = load_stream(targets + min(b + j * GC_WAVE + lane_id(), e - 1));
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (b + j * GC_WAVE + lane_id() < e) atomicAdd(s_push + (t[j] - push_base), cfix);
}

// pr_spmv_merge over the 16-bit stream (prb_chunk; the stream and its start
// bits padded to whole chunks). Dynamic LDS: push_count
// u64 accumulators, then hot_count + 1 contributions (the last one 0); static
// LDS: 8.25 KB (the waves' staged row totals and the partials).
extern "C" __global__ __launch_bounds__(PRP_THREADS) void pr_spmv_bin(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const uint16_t* __restrict__ hot16, const double* __restrict__ cold_sum,
    const float* __restrict__ contrib_in, const float* __restrict__ rank_in,
    float* __restrict__ contrib_out, float* __restrict__ rank_out, uint32_t n, double damping,
    const PrState* __restrict__ state, uint32_t hot_count, uint32_t push_count,
    const uint32_t* __restrict__ pieces, uint32_t piece_count, uint32_t chunks, uint32_t zero_rows,
    const uint8_t* __restrict__ starts, const uint32_t* __restrict__ chunk_row,
    double* __restrict__ head_partial, double* __restrict__ tail_partial, uint32_t* __restrict__ tail_row,
    unsigned long long* __restrict__ push_partials, double* __restrict__ partials) {
  extern __shared__ unsigned long long s_dyn[];
  unsigned long long* s_push = s_dyn;
  float* s_hot = reinterpret_cast<float*>(s_dyn + push_count);
  __shared__ double s_part[2][PRP_THREADS / GC_WAVE];
  __shared__ double s_tot[PRP_THREADS / GC_WAVE][GC_WAVE];  // each wave's staged row totals
  if (state->done) return;
  const uint32_t hot_base = n - hot_count;
  const uint32_t push_base = n - push_count;
  lds_fill_f32<16>(s_hot, contrib_in + hot_base, hot_count, PRP_THREADS);
  if (threadIdx.x == 0) s_hot[hot_count] = 0.0f;
  for (uint32_t i = threadIdx.x; i < push_count; i += PRP_THREADS) s_push[i] = 0ull;
  const double inv_n = 1.0 / static_cast<double>(n);
  const double base = (1.0 - damping) * inv_n + damping * state->dangling * inv_n;
  double dangling = 0.0, diff = 0.0;
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / GC_WAVE);
  const uint32_t waves = gridDim.x * (PRP_THREADS / GC_WAVE);
  const uint32_t zero_items = (zero_rows + PRM_CHUNK - 1) / PRM_CHUNK;
  const uint32_t items = piece_count + chunks + zero_items;
  // the wave's items in order: hub pieces, then chunks, then rows of degree 0
  uint32_t item = blockIdx.x * (PRP_THREADS / GC_WAVE) + wave;
  for (; item < piece_count; item += waves)
    prb_hub_piece(pieces[2 * item], pieces[2 * item + 1], offsets, targets, s_hot, hot_base, s_push, push_base);
  if (item < piece_count + chunks) {
    uint32_t c = item - piece_count;
    PrbOut out;
    out.rank = prb_rsrc(rank_out, n * 4u);
    out.contrib = prb_rsrc(contrib_out, n * 4u);
    out.head = prb_rsrc(head_partial, chunks * 8u);
    out.tail = prb_rsrc(tail_partial, chunks * 8u);
    out.tail_row = prb_rsrc(tail_row, chunks * 4u);
    auto process = [&](uint32_t ch, const PrbChunkIn& in) {
      prb_chunk(ch, in, offsets, cold_sum, zero_rows, contrib_in, s_hot, hot_count, s_push, push_base - hot_base,
                rank_in, out, base, damping, diff, s_tot[wave]);
    };
    // Loads run two chunks ahead of the processing (the last chunk's loads are
    // issued again past the end), and the first chunk is processed before the
    // loop, so that both ways into the loop have the same operations after
    // the loads it waits for: the waits then cover those loads and not the
    // stores of the chunk before (the vector memory counter is in order). The
    // first rows of the wave's chunks c, c + waves, ... come 64 at a time in
    // one load (lane k: the k-th of them).
    const uint32_t last = chunks - 1;
    const uint32_t row_last = push_base - 1;
    const uint32_t c0 = c;
    uint32_t k = 0;  // index of chunk c among the wave's chunks
    uint32_t rows_v = 0;
    auto rows_batch = [&](uint32_t from) {
      const uint64_t ck = static_cast<uint64_t>(from) + static_cast<uint64_t>(lane_id()) * waves;
      rows_v = chunk_row[min(ck, static_cast<uint64_t>(last))];
    };
    // row0 of the wave's chunk number j (j - k < 64 after the batch of chunk k was loaded)
    auto row_of = [&](uint32_t j) {
      return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rows_v), static_cast<int>(j & 63u)));
    };
    auto load = [&](uint32_t j) {
      const uint32_t ch = min(c0 + j * waves, last);
      return prb_load(ch, row_of(j), hot16, starts, offsets, cold_sum, contrib_in, rank_in, zero_rows, row_last);
    };
    rows_batch(c0);
    PrbChunkIn a = load(0);
    PrbChunkIn b = load(1);
    process(c, a);
    a = load(2);
    c += waves;
    ++k;
    while (c < chunks) {
      process(c, b);
      if (((k + 2) & 63u) == 0u) rows_batch(c0 + (k + 2) * waves);
      b = load(k + 2);
      c += waves;
      ++k;
      if (c >= chunks) break;
      process(c, a);
      if (((k + 2) & 63u) == 0u) rows_batch(c0 + (k + 2) * waves);
      a = load(k + 2);
      c += waves;
      ++k;
    }
    item = piece_count + c;
  }
  for (; item < items; item += waves)
    prm_zero_item(item - piece_count - chunks, zero_rows, base, rank_in, rank_out, contrib_out, dangling, diff);
  prm_end(wave, s_push, push_count, push_partials, partials, dangling, diff, s_part);
}

// ---- sliced hot stream (GC_PR_KERNEL=slice) --------------------------------
//
// The pull over the hot entries with one lane per row: slice s holds rows
// z + 64 s + l (lane l), its entries interleaved so that group g (eight
// entries per lane, 16 bytes) sits at [512 g + 8 l ..] and, after the G
// groups, the remainder of R entries per lane (R = 0, 2, 4, 6) at
// [512 G + R l ..]; entries hot_count (a row's end, rows past the last pull
// row) read the zero held after the hot set. A lane sums its row in FP64 and
// pushes its own c into the LDS accumulators for entries in P, with no
// exchange between lanes, where the chunked stream of pr_spmv_bin needs a
// segmented scan across the wave and a shuffle per entry. Widths of the 64
// rows differ, and the padding is what that costs: 1.24x the hot entries at
// R-MAT scale 22.
//
// Items (four words, pr_slice_layout.hpp): a whole slice (its rows finished
// here), a piece of a wide slice (its 64 sums go to slice_partial, and
// pr_slice_finish adds a row's pieces in order) or a hub piece (pushes only,
// as prb_hub_piece). Wave w takes items [wave_first[w], wave_first[w + 1]),
// dealt on the host longest first to the wave with the least work, its hub
// pieces first; then the rows of degree 0 in runs of 512, round-robin.
//
// The slice items run one ahead: the next item's row data, its first four
// groups (clamped to its last group) and its remainder are loaded before the
// current item is summed, so a wave keeps ten loads in flight while it works;
// an item of more than four groups loads the rest itself. The descriptors
// come by scalar loads two items ahead. The row data and the stores carry no
// branch (clamped loads, buffer stores dropped out of range), so that the
// waits for the loads ahead stay exact.
typedef uint32_t prs_u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

struct PrsIn {
  prb_u16x8 v[4];
  prs_u32x4a4 rem;
  float cst, rin;
  double cold;
  uint32_t off0, off1;
};

__device__ __forceinline__ PrsIn prs_load(uint32_t x, uint32_t y, uint32_t zw, const uint16_t* __restrict__ data,
                                          const float* __restrict__ contrib_in, const float* __restrict__ rank_in,
                                          const double* __restrict__ cold_sum, const uint32_t* __restrict__ offsets,
                                          uint32_t zero_rows, uint32_t push_base) {
  PrsIn in;
  const uint32_t lane = lane_id();
  const uint32_t rc = min(zero_rows + 64u * (x & 0x3FFFFFFFu) + lane, push_base - 1);
  const uint32_t ng = zw & 0xFFFFu, R = zw >> 16;
  const uint32_t glast = ng > 0 ? ng - 1 : 0u;
  const prb_u16x8* grp = reinterpret_cast<const prb_u16x8*>(data + y) + lane;
#pragma unroll
  for (int q = 0; q < 4; ++q) in.v[q] = __builtin_nontemporal_load(grp + 64u * min(static_cast<uint32_t>(q), glast));
  in.rem = *reinterpret_cast<const prs_u32x4a4*>(data + y + 512u * ng + R * lane);
  in.cst = contrib_in[rc];
  in.rin = rank_in[rc];
  in.cold = cold_sum[rc - zero_rows];
  in.off0 = offsets[rc];
  in.off1 = offsets[rc + 1];
  return in;
}

__device__ __forceinline__ void prs_entry(uint32_t t, const float* __restrict__ s_hot,
                                          unsigned long long* __restrict__ s_push, uint32_t push_off,
                                          uint32_t push_count, unsigned long long cfix, double& acc) {
  acc += static_cast<double>(s_hot[t]);
  if (t - push_off < push_count) atomicAdd(s_push + (t - push_off), cfix);
}

// Sums one slice item from its loads: the first four groups, the remainder,
// then (loaded here) the further groups; a whole slice's rows are finished, a
// piece's sums written to its slot.
__device__ __forceinline__ void prs_item(uint32_t x, uint32_t y, uint32_t zw, uint32_t slot, const PrsIn& in,
                                         const uint16_t* __restrict__ data, const float* __restrict__ s_hot,
                                         unsigned long long* __restrict__ s_push, uint32_t push_off,
                                         uint32_t push_count, uint32_t zero_rows, uint32_t push_base,
                                         const PrbOut& out, double* __restrict__ slice_partial, double base,
                                         double damping, double& diff) {
  const uint32_t lane = lane_id();
  const uint32_t ng = zw & 0xFFFFu, R = zw >> 16;
  const unsigned long long cfix = pr_to_fixed(in.cst);
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (static_cast<uint32_t>(q) < ng) {
#pragma unroll
      for (int i = 0; i < 8; ++i) prs_entry(in.v[q][i], s_hot, s_push, push_off, push_count, cfix, acc);
    }
  }
  const prb_u16x8 rem = __builtin_bit_cast(prb_u16x8, in.rem);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if (static_cast<uint32_t>(i) < R) prs_entry(rem[i], s_hot, s_push, push_off, push_count, cfix, acc);
  }
  // the row data is taken here, before any further loads, while the waits for it are exact on
  // every path
  __asm__ volatile("" ::"v"(in.cold), "v"(in.rin), "v"(in.off0), "v"(in.off1));
  if (ng > 4) {
    const prb_u16x8* grp = reinterpret_cast<const prb_u16x8*>(data + y) + lane;
    for (uint32_t g = 4; g < ng; g += 4) {
      prb_u16x8 v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = __builtin_nontemporal_load(grp + 64u * min(g + q, ng - 1));
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (g + q < ng) {
#pragma unroll
          for (int i = 0; i < 8; ++i) prs_entry(v[q][i], s_hot, s_push, push_off, push_count, cfix, acc);
        }
      }
    }
  }
  const uint32_t row = zero_rows + 64u * (x & 0x3FFFFFFFu) + lane;
  if ((x >> 30) == 0u) {
    const bool valid = row < push_base;
    const float r = static_cast<float>(base + damping * (acc + in.cold));
    prb_st32(out.rank, row * 4u, valid, __builtin_bit_cast(uint32_t, r));
    prb_st32(out.contrib, row * 4u, valid,
             __builtin_bit_cast(uint32_t, r / static_cast<float>(in.off1 - in.off0)));
    diff += valid ? fabs(static_cast<double>(r) - static_cast<double>(in.rin)) : 0.0;
  } else {
    slice_partial[static_cast<uint64_t>(slot) * GC_WAVE + lane] = acc;
  }
}

// Hub piece (one wave) of the sliced layout: its up to 512 entries in P as
// 16-bit offsets from push_base (0xFFFF past the piece's end), eight per lane
// in one 16-byte load, each taking c[h] (the pushes of prb_hub_piece).
__device__ __forceinline__ void prs_hub_piece(uint32_t h, uint32_t first, const uint16_t* __restrict__ hub16,
                                              const float* __restrict__ s_hot, uint32_t hot_base,
                                              unsigned long long* __restrict__ s_push) {
  const unsigned long long cfix = pr_to_fixed(s_hot[h - hot_base]);
  const prb_u16x8 t = __builtin_nontemporal_load(reinterpret_cast<const prb_u16x8*>(hub16 + first) + lane_id());
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (t[k] != 0xFFFFu) atomicAdd(s_push + t[k], cfix);
}

// Dynamic LDS: push_count u64 accumulators, then hot_count + 1 contributions
// (the last one 0), as pr_spmv_bin.
extern "C" __global__ __launch_bounds__(PRP_THREADS) void pr_spmv_slice(
    const uint32_t* __restrict__ offsets, const uint16_t* __restrict__ hub16,
    const uint16_t* __restrict__ data, const uint4* __restrict__ items, const uint32_t* __restrict__ wave_first,
    const double* __restrict__ cold_sum, const float* __restrict__ contrib_in, const float* __restrict__ rank_in,
    float* __restrict__ contrib_out, float* __restrict__ rank_out, uint32_t n, double damping,
    const PrState* __restrict__ state, uint32_t hot_count, uint32_t push_count, uint32_t zero_rows,
    double* __restrict__ slice_partial, unsigned long long* __restrict__ push_partials,
    double* __restrict__ partials) {
  extern __shared__ unsigned long long s_dyn[];
  unsigned long long* s_push = s_dyn;
  float* s_hot = reinterpret_cast<float*>(s_dyn + push_count);
  __shared__ double s_part[2][PRP_THREADS / GC_WAVE];
  if (state->done) return;
  const uint32_t hot_base = n - hot_count;
  const uint32_t push_base = n - push_count;
  const uint32_t push_off = hot_count - push_count;
  lds_fill_f32<16>(s_hot, contrib_in + hot_base, hot_count, PRP_THREADS);
  if (threadIdx.x == 0) s_hot[hot_count] = 0.0f;
  for (uint32_t i = threadIdx.x; i < push_count; i += PRP_THREADS) s_push[i] = 0ull;
  const double inv_n = 1.0 / static_cast<double>(n);
  const double base = (1.0 - damping) * inv_n + damping * state->dangling * inv_n;
  double dangling = 0.0, diff = 0.0;
  __syncthreads();
  const uint32_t lane = lane_id();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / GC_WAVE);
  const uint32_t gw = blockIdx.x * (PRP_THREADS / GC_WAVE) + wave;
  const uint32_t k0 = wave_first[gw], k1 = wave_first[gw + 1];
  if (k0 < k1) {
    PrbOut out;
    out.rank = prb_rsrc(rank_out, n * 4u);
    out.contrib = prb_rsrc(contrib_out, n * 4u);
    uint32_t k = k0;
    for (; k < k1; ++k) {
      const uint4 it = items[k];
      if ((it.x >> 30) != 2u) break;
      prs_hub_piece(it.y, it.w, hub16, s_hot, hot_base, s_push);
    }
    if (k < k1) {
      // slice items, one ahead; the last item's loads are issued again past the end
      uint4 da = items[k];
      uint4 db = items[min(k + 1, k1 - 1)];
      PrsIn a = prs_load(da.x, da.y, da.z, data, contrib_in, rank_in, cold_sum, offsets, zero_rows, push_base);
      while (true) {
        const uint4 dc = items[min(k + 2, k1 - 1)];
        PrsIn b = prs_load(db.x, db.y, db.z, data, contrib_in, rank_in, cold_sum, offsets, zero_rows, push_base);
        prs_item(da.x, da.y, da.z, da.w, a, data, s_hot, s_push, push_off, push_count, zero_rows, push_base, out,
                 slice_partial, base, damping, diff);
        if (++k >= k1) break;
        da = db;
        db = dc;
        a = b;
      }
    }
  }
  const uint32_t zero_items = (zero_rows + PRM_CHUNK - 1) / PRM_CHUNK;
  for (uint32_t item = gw; item < zero_items; item += gridDim.x * (PRP_THREADS / GC_WAVE))
    prm_zero_item(item, zero_rows, base, rank_in, rank_out, contrib_out, dangling, diff);
  prm_end(wave, s_push, push_count, push_partials, partials, dangling, diff, s_part);
}

// The hub rows as pr_push_finish (blocks [0, hub_blocks), 16 hubs each), then
// the rows of the wide slices (blocks from hub_blocks on, four slices of 64
// rows each): long_slices holds (slice, first partial slot, pieces); a row's
// sum is its pieces' sums in order plus its cold sum. Per-block dangling and
// L1 partials go to partials from partial_base on.
extern "C" __global__ __launch_bounds__(256) void pr_slice_finish(
    const uint32_t* __restrict__ offsets, const float* __restrict__ rank_in,
    float* __restrict__ contrib_out, float* __restrict__ rank_out, uint32_t n, double damping,
    const PrState* __restrict__ state, uint32_t push_count,
    const unsigned long long* __restrict__ push_partials, uint32_t parts, uint32_t hub_blocks,
    const uint32_t* __restrict__ long_slices, uint32_t long_count, const double* __restrict__ slice_partial,
    const double* __restrict__ cold_sum, uint32_t zero_rows, double* __restrict__ partials,
    uint32_t partial_base) {
  __shared__ unsigned long long s_sum[256];
  __shared__ double s_part[2][256 / GC_WAVE];
  if (state->done) return;
  const double inv_n = 1.0 / static_cast<double>(n);
  const double base = (1.0 - damping) * inv_n + damping * state->dangling * inv_n;
  double dangling = 0.0, diff = 0.0;
  if (blockIdx.x < hub_blocks) {
    const uint32_t hub = threadIdx.x % PRF_HUBS, j = threadIdx.x / PRF_HUBS;
    const uint32_t i = blockIdx.x * PRF_HUBS + hub;
    unsigned long long sum = 0ull;
    if (i < push_count) {
      for (uint32_t w = j; w < parts; w += 256 / PRF_HUBS) sum += push_partials[static_cast<uint64_t>(w) * push_count + i];
    }
    s_sum[threadIdx.x] = sum;
    __syncthreads();
    if (j == 0 && i < push_count) {
      unsigned long long y = 0ull;
      for (uint32_t q = 0; q < 256 / PRF_HUBS; ++q) y += s_sum[q * PRF_HUBS + hub];
      const uint32_t row = n - push_count + i;
      pr_finish_row(row, offsets[row + 1] - offsets[row], static_cast<double>(y) / PRP_FIX_SCALE, base, damping,
                    rank_in, rank_out, contrib_out, dangling, diff);
    }
  } else {
    const uint32_t q = (blockIdx.x - hub_blocks) * (256 / GC_WAVE) + threadIdx.x / GC_WAVE;
    if (q < long_count) {
      const uint32_t s = long_slices[3 * q], p0 = long_slices[3 * q + 1], np = long_slices[3 * q + 2];
      const uint32_t row = zero_rows + 64u * s + lane_id();
      if (row < n - push_count) {
        double acc = 0.0;
        for (uint32_t p = 0; p < np; ++p) acc += slice_partial[static_cast<uint64_t>(p0 + p) * GC_WAVE + lane_id()];
        pr_finish_row(row, offsets[row + 1] - offsets[row], acc + cold_sum[row - zero_rows], base, damping, rank_in,
                      rank_out, contrib_out, dangling, diff);
      }
    }
  }
  dangling = wave_sum(dangling);
  diff = wave_sum(diff);
  if (lane_id() == 0) {
    s_part[0][threadIdx.x / GC_WAVE] = dangling;
    s_part[1][threadIdx.x / GC_WAVE] = diff;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double d0 = 0.0, d1 = 0.0;
    for (int w = 0; w < 256 / GC_WAVE; ++w) {
      d0 += s_part[0][w];
      d1 += s_part[1][w];
    }
    const uint64_t slot = partial_base + blockIdx.x;
    partials[2 * slot] = d0;
    partials[2 * slot + 1] = d1;
  }
}
