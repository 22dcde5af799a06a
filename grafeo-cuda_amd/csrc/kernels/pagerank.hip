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
#define PR_MAX_SEGMENTS 10

struct PrSegment {
  uint32_t row_begin;
  uint32_t row_end;
  uint32_t group;        // lanes per row: 1, 2, ..., 64, or 256 (one workgroup)
  uint32_t block_begin;  // first workgroup of the segment
};

struct PrSegments {
  PrSegment seg[PR_MAX_SEGMENTS];
  uint32_t count;
  uint32_t total_blocks;
};

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
    for (uint32_t i = b + g; i < e; i += G) acc += static_cast<double>(contrib_in[load_stream(targets + i)]);
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (valid && g == 0) {
    pr_finish_row(row, e - b, acc, base, damping, rank_in, rank_out, contrib_out, dangling, diff);
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
  double acc = 0.0;
  for (uint32_t i = b + threadIdx.x; i < e; i += PR_THREADS) {
    acc += static_cast<double>(contrib_in[load_stream(targets + i)]);
  }
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

extern "C" __global__ __launch_bounds__(PR_THREADS) void pr_spmv(
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ targets,
    const float* __restrict__ contrib_in, const float* __restrict__ rank_in,
    float* __restrict__ contrib_out, float* __restrict__ rank_out, uint32_t n, double damping,
    PrSegments segs, const PrState* __restrict__ state, double* __restrict__ partials) {
  __shared__ double s_red[PR_THREADS / GC_WAVE];
  __shared__ double s_part[2][PR_THREADS / GC_WAVE];
  if (state->done) return;
  const uint32_t blk = blockIdx.x;
  uint32_t s = 0;
  while (s + 1 < segs.count && blk >= segs.seg[s + 1].block_begin) ++s;
  const PrSegment seg = segs.seg[s];
  const uint32_t local = blk - seg.block_begin;
  const double inv_n = 1.0 / static_cast<double>(n);
  const double base = (1.0 - damping) * inv_n + damping * state->dangling * inv_n;
  double dangling = 0.0, diff = 0.0;
  switch (seg.group) {
    case 1: pr_rows_group<1>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 2: pr_rows_group<2>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 4: pr_rows_group<4>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 8: pr_rows_group<8>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 16: pr_rows_group<16>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 32: pr_rows_group<32>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    case 64: pr_rows_group<64>(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff); break;
    default: pr_row_block(seg, local, offsets, targets, contrib_in, rank_in, contrib_out, rank_out, base, damping, dangling, diff, s_red); break;
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
    partials[2 * static_cast<uint64_t>(blk)] = d0;
    partials[2 * static_cast<uint64_t>(blk) + 1] = d1;
  }
}

// Adds the per-workgroup partials in a fixed order and advances the state.
extern "C" __global__ __launch_bounds__(1024) void pr_finalize(PrState* __restrict__ state,
                                                               const double* __restrict__ partials,
                                                               uint32_t blocks, double tolerance) {
  __shared__ double s_a[1024 / GC_WAVE];
  __shared__ double s_b[1024 / GC_WAVE];
  if (state->done) return;
  double d0 = 0.0, d1 = 0.0;
  for (uint32_t i = threadIdx.x; i < blocks; i += 1024) {
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
    for (int w = 0; w < 1024 / GC_WAVE; ++w) {
      a += s_a[w];
      b += s_b[w];
    }
    state->dangling = a;
    state->diff = b;
    state->iterations += 1;
    if (b < tolerance) state->done = 1;
  }
}
