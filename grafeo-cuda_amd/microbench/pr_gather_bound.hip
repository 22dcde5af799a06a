// Microbenchmark behind PageRank's SpMV bound on MI355X (not part of the library).
//
// 1. Random 4-byte gathers table[idx[i]] over tables of 16 K .. 16 M floats (inside the L1,
//    inside one XCD's 4 MB L2, beyond it): the gather rate each level of the hierarchy sustains.
// 2. The access pattern of pr_spmv_push on R-MAT scale 22 (the library's degree-ordered layout),
//    flattened into one stream with no row structure: the target stream the kernel reads, each
//    entry gathered from LDS (the top `hot` ids) or from global memory, with and without the LDS
//    pushes for the top `push` ids; then with every gather served by LDS, then the stream alone.
// 3. The library's own push-pull SpMV time per iteration on the same graph (pr_spmv_merge, the
//    default, and pr_spmv_push, the row-tiled kernel, with GC_PR_KERNEL=push).
//
// Build and run (repo root, after build.py):
//   hipcc --offload-arch=gfx950 -O3 -std=c++17 -Xclang -target-feature -Xclang -packed-fp32-ops \
//     -Iinclude grafeo-cuda_amd/microbench/pr_gather_bound.hip -Lgrafeo-cuda_amd/grafeo_cuda \
//     -lgrafeo_hip -Wl,-rpath,grafeo-cuda_amd/grafeo_cuda -o gpurun_out/pr_gather_bound
//   gpurun_out/pr_gather_bound
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "grafeo_hip.h"

#define CK(x)                                                                          \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);       \
      std::exit(1);                                                                    \
    }                                                                                  \
  } while (0)
#define GK(x)                                                                          \
  do {                                                                                 \
    if ((x) != GC_OK) {                                                                \
      std::printf("library error %s at line %d\n", gc_last_error(), __LINE__);         \
      std::exit(1);                                                                    \
    }                                                                                  \
  } while (0)

namespace {

// splitmix64 finaliser: uniform indices from a counter
__device__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void uniform_indices(uint32_t* idx, uint64_t n, uint32_t range, uint64_t seed) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull)
    idx[i] = static_cast<uint32_t>(mix64(i ^ seed) % range);
}

__global__ void fill_table(float* t, uint64_t n) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) t[i] = 1.0f / (1 + (i & 1023));
}

// eight independent gathers in flight per lane; out holds one value per thread of the grid
__global__ __launch_bounds__(256) void gather8(const uint32_t* __restrict__ idx, uint64_t n,
                                               const float* __restrict__ table, float* __restrict__ out) {
  const uint64_t tid = blockIdx.x * 256ull + threadIdx.x, stride = gridDim.x * 256ull;
  float acc = 0.f;
  uint64_t i = tid;
  for (; i + 7 * stride < n; i += 8 * stride) {
    uint32_t t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = __builtin_nontemporal_load(idx + i + k * stride);
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += table[t[k]];
  }
  for (; i < n; i += stride) acc += table[idx[i]];
  out[tid] = acc;
}

// MODE 0: LDS for t >= hot_base, global otherwise, LDS pushes for t >= push_base;
// MODE 1: the same without pushes; MODE 2: every gather from LDS; MODE 3: the stream alone.
// Dynamic LDS: push_count u64 then hot_count f32; out holds one value per workgroup.
template <int MODE>
__global__ __launch_bounds__(1024) void push_pattern(const uint32_t* __restrict__ idx, uint64_t n,
                                                     const float* __restrict__ table, uint32_t hot_base,
                                                     uint32_t hot_count, uint32_t push_base, uint32_t push_count,
                                                     float* __restrict__ out) {
  extern __shared__ unsigned long long s_dyn[];
  unsigned long long* s_push = s_dyn;
  float* s_hot = reinterpret_cast<float*>(s_dyn + push_count);
  for (uint32_t i = threadIdx.x; i < hot_count; i += 1024) s_hot[i] = table[hot_base + i];
  for (uint32_t i = threadIdx.x; i < push_count; i += 1024) s_push[i] = 0;
  __syncthreads();
  const uint64_t tid = blockIdx.x * 1024ull + threadIdx.x, stride = gridDim.x * 1024ull;
  double acc = 0;
  uint64_t i = tid;
  for (; i + 7 * stride < n; i += 8 * stride) {
    uint32_t t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = __builtin_nontemporal_load(idx + i + k * stride);
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (MODE == 3) x[k] = static_cast<float>(t[k] & 1);
      else if (MODE == 2) x[k] = s_hot[t[k] % hot_count];
      else x[k] = t[k] >= hot_base ? s_hot[t[k] - hot_base] : table[t[k]];
    }
    if (MODE == 0) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (t[k] >= push_base) atomicAdd(s_push + (t[k] - push_base), 1ull);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += x[k];
  }
  for (; i < n; i += stride) acc += table[idx[i]];
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = static_cast<float>(acc) + (MODE == 0 ? static_cast<float>(s_push[0]) : 0.f);
}

template <class F>
float best_ms(F f, int reps = 5) {
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  f();
  CK(hipDeviceSynchronize());
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    CK(hipEventRecord(a));
    f();
    CK(hipEventRecord(b));
    CK(hipEventSynchronize(b));
    float ms;
    CK(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, ms);
  }
  CK(hipEventDestroy(a));
  CK(hipEventDestroy(b));
  return best;
}

}  // namespace

int main() {
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, 0));
  const int cu = prop.multiProcessorCount;
  std::printf("%s: %d CUs, L2 %d bytes per XCD\n", prop.gcnArchName, cu, prop.l2CacheSize);

  // 1. uniform random gathers, 128 M per launch, 8 workgroups of 256 per CU
  const uint64_t N = 128ull << 20;
  const uint64_t max_table = 16ull << 20;
  const int grid = cu * 8;
  uint32_t* idx;
  float* table;
  float* out;
  CK(hipMalloc(&idx, N * 4));
  CK(hipMalloc(&table, max_table * 4));
  CK(hipMalloc(&out, static_cast<size_t>(grid) * 256 * 4));
  fill_table<<<4096, 256>>>(table, max_table);
  CK(hipGetLastError());
  for (uint32_t S : {16384u, 262144u, 1u << 20, 4u << 20, 16u << 20}) {
    uniform_indices<<<4096, 256>>>(idx, N, S, 12345);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    const float ms = best_ms([&] { gather8<<<grid, 256>>>(idx, N, table, out); });
    std::printf("uniform gathers over %8u floats (%6.2f MB): %.3f ms = %.1f G gathers/s\n", S, S * 4.0 / 1e6, ms,
                N / ms / 1e6);
  }
  CK(hipFree(idx));
  CK(hipFree(table));
  CK(hipFree(out));

  // 2. the push kernel's pattern on R-MAT 22
  gc_device* dev;
  GK(gc_device_new(0, &dev));
  uint32_t* pairs;
  size_t np;
  GK(gc_gen_rmat(22, 16, 0.57, 0.19, 0.19, 1, 1, &pairs, &np));
  gc_graph* g;
  GK(gc_graph_from_edge_pairs(dev, pairs, np, 1u << 22, 1, &g));
  gc_free(pairs);
  double secs;
  GK(gc_graph_prepare(g, &secs));
  const uint32_t *roff, *rtgt, *perm, *fwd;
  GK(gc_graph_relabelled(g, &roff, &rtgt, &perm, &fwd));
  const uint32_t n = gc_graph_num_nodes(g);
  const uint64_t m = gc_graph_num_edges(g);
  std::vector<uint32_t> off(n + 1ull), tgt(m);
  CK(hipMemcpy(off.data(), roff, 4ull * (n + 1), hipMemcpyDeviceToHost));
  CK(hipMemcpy(tgt.data(), rtgt, 4ull * m, hipMemcpyDeviceToHost));
  // an LDS split of the push-pull kernels at this size (8192 pushed hubs, 22528 contributions
  // cached; the default is 6144 and 26624, the same 152 KB)
  const uint32_t push_count = 8192, hot_count = 22528;
  const uint32_t push_base = n - push_count, hot_base = n - hot_count;
  // the entries pr_spmv_push streams: every row's list, but only the part of a pushed hub's list
  // that points at other pushed hubs (the rest arrives by push)
  std::vector<uint32_t> stream;
  stream.reserve(m);
  for (uint32_t r = 0; r < n; ++r) {
    uint32_t b = off[r];
    const uint32_t e = off[r + 1];
    if (r >= push_base) b = static_cast<uint32_t>(std::lower_bound(tgt.begin() + b, tgt.begin() + e, push_base) - tgt.begin());
    for (uint32_t i = b; i < e; ++i) stream.push_back(tgt[i]);
  }
  uint64_t cold = 0;
  for (uint32_t t : stream) cold += t < hot_base;
  std::printf("R-MAT 22: n = %u, m = %llu; streamed %zu entries (%.1f%% of m), %llu global gathers\n", n,
              static_cast<unsigned long long>(m), stream.size(), 100.0 * stream.size() / m,
              static_cast<unsigned long long>(cold));
  uint32_t* d_stream;
  float* d_table;
  float* d_out;
  CK(hipMalloc(&d_stream, 4 * stream.size()));
  CK(hipMalloc(&d_table, 4ull * n));
  CK(hipMalloc(&d_out, 4096 * 4));
  CK(hipMemcpy(d_stream, stream.data(), 4 * stream.size(), hipMemcpyHostToDevice));
  CK(hipMemset(d_table, 0, 4ull * n));
  const uint32_t shmem = 8 * push_count + 4 * hot_count;  // 155,648 bytes of the 160 KB
  const uint64_t S = stream.size();
  const float m0 = best_ms([&] {
    push_pattern<0><<<cu, 1024, shmem>>>(d_stream, S, d_table, hot_base, hot_count, push_base, push_count, d_out);
  });
  const float m1 = best_ms([&] {
    push_pattern<1><<<cu, 1024, shmem>>>(d_stream, S, d_table, hot_base, hot_count, push_base, push_count, d_out);
  });
  const float m2 = best_ms([&] {
    push_pattern<2><<<cu, 1024, shmem>>>(d_stream, S, d_table, hot_base, hot_count, push_base, push_count, d_out);
  });
  const float m3 = best_ms([&] {
    push_pattern<3><<<cu, 1024, shmem>>>(d_stream, S, d_table, hot_base, hot_count, push_base, push_count, d_out);
  });
  std::printf("flattened push pattern: full %.1f us, without pushes %.1f us, every gather from LDS %.1f us, "
              "stream alone %.1f us\n", m0 * 1e3, m1 * 1e3, m2 * 1e3, m3 * 1e3);
  std::printf("global gathers in the full pattern: %.1f G/s over the time above the stream alone\n",
              cold / ((m0 - m3) * 1e-3) / 1e9);

  // 3. the library kernel on the same graph, per iteration that ran
  float* d_rank;
  CK(hipMalloc(&d_rank, 4ull * n));
  uint32_t it;
  int32_t conv;
  GK(gc_pagerank_device(g, 0.85f, 20, 1e-6f, d_rank, &it, &conv, &secs));
  GK(gc_profile_enable(dev, 1));
  GK(gc_pagerank_device(g, 0.85f, 20, 1e-6f, d_rank, &it, &conv, &secs));
  for (const char* name : {"pr_spmv_merge", "pr_spmv_push"}) {
    double total_ms;
    uint64_t launches;
    GK(gc_profile_query(dev, name, &total_ms, &launches));
    if (launches != 0)
      std::printf("library %s: %.1f us per iteration that ran (%u iterations, %llu launches)\n", name,
                  1000.0 * total_ms / it, it, static_cast<unsigned long long>(launches));
  }
  CK(hipFree(d_rank));
  CK(hipFree(d_stream));
  CK(hipFree(d_table));
  CK(hipFree(d_out));
  gc_graph_free(g);
  gc_device_release(dev);
  return 0;
}
