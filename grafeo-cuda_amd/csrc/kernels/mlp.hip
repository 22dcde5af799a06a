// The scheduler's network: a 12 -> 32 -> 32 -> 3 FP32 MLP (ReLU hidden layers,
// linear outputs = predicted log-times of three choices), with forward,
// per-sample backward, a fixed-order gradient reduction and an Adam step.
//
// Parameter layout (1,571 floats): W1[32][12], b1[32], W2[32][32], b2[32],
// W3[3][32], b3[3]. Inputs are normalised as (x - mean) * scale before the
// first layer. The loss of a batch is the mean over observed outputs
// (mask = 1) of the squared error: L = sum mask (y - t)^2 / sum mask.
#include "device.hpp"

using namespace gcd;

#define MLP_IN 12
#define MLP_H 32
#define MLP_OUT 3
#define MLP_W1 0
#define MLP_B1 (MLP_W1 + MLP_H * MLP_IN)
#define MLP_W2 (MLP_B1 + MLP_H)
#define MLP_B2 (MLP_W2 + MLP_H * MLP_H)
#define MLP_W3 (MLP_B2 + MLP_H)
#define MLP_B3 (MLP_W3 + MLP_OUT * MLP_H)
#define MLP_PARAMS (MLP_B3 + MLP_OUT)  // 1571

struct MlpNorm {
  float mean[MLP_IN];
  float scale[MLP_IN];
};

// Copies the parameters into LDS (every thread of the block takes part).
__device__ __forceinline__ void mlp_stage(const float* __restrict__ params, float* s_p) {
  for (uint32_t i = threadIdx.x; i < MLP_PARAMS; i += blockDim.x) s_p[i] = params[i];
  __syncthreads();
}

__device__ __forceinline__ void mlp_forward_one(const float* p, const float* x, const MlpNorm& norm,
                                                float* xn, float* h1, float* h2, float* y) {
#pragma unroll
  for (int i = 0; i < MLP_IN; ++i) xn[i] = (x[i] - norm.mean[i]) * norm.scale[i];
#pragma unroll 4
  for (int j = 0; j < MLP_H; ++j) {
    float a = p[MLP_B1 + j];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) a = fmaf(p[MLP_W1 + j * MLP_IN + i], xn[i], a);
    h1[j] = a > 0.0f ? a : 0.0f;
  }
#pragma unroll 4
  for (int j = 0; j < MLP_H; ++j) {
    float a = p[MLP_B2 + j];
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) a = fmaf(p[MLP_W2 + j * MLP_H + i], h1[i], a);
    h2[j] = a > 0.0f ? a : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < MLP_OUT; ++k) {
    float a = p[MLP_B3 + k];
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) a = fmaf(p[MLP_W3 + k * MLP_H + i], h2[i], a);
    y[k] = a;
  }
}

// y[s][0..3) for samples s < count (one thread per sample).
extern "C" __global__ __launch_bounds__(256) void mlp_forward(const float* __restrict__ params,
                                                              MlpNorm norm,
                                                              const float* __restrict__ x,
                                                              uint32_t count,
                                                              float* __restrict__ y) {
  __shared__ float s_p[MLP_PARAMS];
  mlp_stage(params, s_p);
  for (uint64_t s = global_thread(); s < count; s += grid_threads()) {
    float xs[MLP_IN], xn[MLP_IN], h1[MLP_H], h2[MLP_H], ys[MLP_OUT];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) xs[i] = x[s * MLP_IN + i];
    mlp_forward_one(s_p, xs, norm, xn, h1, h2, ys);
#pragma unroll
    for (int k = 0; k < MLP_OUT; ++k) y[s * MLP_OUT + k] = ys[k];
  }
}

// Per-sample gradient of L (all 1,571 parameters) into grads[s][.], and the
// sample's share of the loss into losses[s]; inv_obs = 1 / sum(mask).
extern "C" __global__ __launch_bounds__(256) void mlp_backward(
    const float* __restrict__ params, MlpNorm norm, const float* __restrict__ x,
    const float* __restrict__ t, const float* __restrict__ mask, uint32_t count, float inv_obs,
    float* __restrict__ grads, float* __restrict__ losses) {
  __shared__ float s_p[MLP_PARAMS];
  mlp_stage(params, s_p);
  for (uint64_t s = global_thread(); s < count; s += grid_threads()) {
    float xs[MLP_IN], xn[MLP_IN], h1[MLP_H], h2[MLP_H], y[MLP_OUT];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) xs[i] = x[s * MLP_IN + i];
    mlp_forward_one(s_p, xs, norm, xn, h1, h2, y);
    float* g = grads + s * MLP_PARAMS;
    float dy[MLP_OUT];
    float loss = 0.0f;
#pragma unroll
    for (int k = 0; k < MLP_OUT; ++k) {
      const float e = (y[k] - t[s * MLP_OUT + k]) * mask[s * MLP_OUT + k];
      loss = fmaf(e, e, loss);
      dy[k] = 2.0f * e * inv_obs;
      g[MLP_B3 + k] = dy[k];
#pragma unroll 8
      for (int i = 0; i < MLP_H; ++i) g[MLP_W3 + k * MLP_H + i] = dy[k] * h2[i];
    }
    losses[s] = loss * inv_obs;
    float dh2[MLP_H];
#pragma unroll 4
    for (int i = 0; i < MLP_H; ++i) {
      float a = 0.0f;
#pragma unroll
      for (int k = 0; k < MLP_OUT; ++k) a = fmaf(s_p[MLP_W3 + k * MLP_H + i], dy[k], a);
      dh2[i] = h2[i] > 0.0f ? a : 0.0f;
      g[MLP_B2 + i] = dh2[i];
    }
    for (int j = 0; j < MLP_H; ++j) {
#pragma unroll 8
      for (int i = 0; i < MLP_H; ++i) g[MLP_W2 + j * MLP_H + i] = dh2[j] * h1[i];
    }
    for (int i = 0; i < MLP_H; ++i) {
      float a = 0.0f;
#pragma unroll 8
      for (int j = 0; j < MLP_H; ++j) a = fmaf(s_p[MLP_W2 + j * MLP_H + i], dh2[j], a);
      const float d1 = h1[i] > 0.0f ? a : 0.0f;
      g[MLP_B1 + i] = d1;
#pragma unroll
      for (int c = 0; c < MLP_IN; ++c) g[MLP_W1 + i * MLP_IN + c] = d1 * xn[c];
    }
  }
}

// grad[p] = sum over samples of grads[s][p] in sample order, one thread per
// parameter (deterministic); the thread after the last parameter sums the
// losses in sample order.
extern "C" __global__ __launch_bounds__(256) void mlp_reduce(const float* __restrict__ grads,
                                                             const float* __restrict__ losses,
                                                             uint32_t count,
                                                             float* __restrict__ grad,
                                                             float* __restrict__ loss) {
  const uint64_t p = global_thread();
  if (p < MLP_PARAMS) {
    double a = 0.0;
    for (uint32_t s = 0; s < count; ++s) a += grads[static_cast<uint64_t>(s) * MLP_PARAMS + p];
    grad[p] = static_cast<float>(a);
  } else if (p == MLP_PARAMS) {
    double a = 0.0;
    for (uint32_t s = 0; s < count; ++s) a += losses[s];
    *loss = static_cast<float>(a);
  }
}

// One Adam step (Kingma and Ba, 2015) with bias correction; step >= 1.
extern "C" __global__ __launch_bounds__(256) void mlp_adam(float* __restrict__ params,
                                                           const float* __restrict__ grad,
                                                           float* __restrict__ m1,
                                                           float* __restrict__ m2, uint32_t step,
                                                           float lr, float beta1, float beta2,
                                                           float eps) {
  const uint64_t p = global_thread();
  if (p >= MLP_PARAMS) return;
  const float g = grad[p];
  const float a = beta1 * m1[p] + (1.0f - beta1) * g;
  const float b = beta2 * m2[p] + (1.0f - beta2) * g * g;
  m1[p] = a;
  m2[p] = b;
  const float c1 = 1.0f - powf(beta1, static_cast<float>(step));
  const float c2 = 1.0f - powf(beta2, static_cast<float>(step));
  params[p] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// ---- BFS direction per level ------------------------------------------------
//
// The state of a BFS at the start of a level and the 12 features derived from
// it (all log2 of counts + 1 unless noted):
//   0 n, 1 m, 2 frontier vertices n_f, 3 frontier edges m_f, 4 edges of the
//   vertices not reached yet m_u, 5 vertices reached, 6 log2((m_f+1)/(m_u+1)),
//   7 log2((n_f+1)/(previous n_f+1)), 8 level index, 9 frontier held as a
//   bitmap (0/1), 10 average degree m/n, 11 maximum degree.
// Choices: 0 top-down, 1 bottom-up, 2 device-chained levels (captured HIP
// graphs of many levels, read back once per chain).

struct BfsLevelState {
  unsigned long long m, m_f, m_u, reached;
  uint32_t n, n_f, prev_n_f, level, bitmap, max_degree, pad0, pad1;
};

struct BfsCountersView {
  uint32_t next_size;
  uint32_t pad;
  unsigned long long next_edges;
};

struct SchedDecision {
  float features[MLP_IN];
  float predicted[MLP_OUT];
  uint32_t choice;
};

// BfsPublished of bfs.hip: the level counters, this decision, and the
// sequence word the host polls.
struct BfsPublishedView {
  BfsCountersView counters;
  SchedDecision decision;
  uint32_t seq;
  uint32_t pad[3];
};

__device__ __forceinline__ float lg(double x) { return static_cast<float>(log2(x + 1.0)); }

// Features of the level about to run, from `st`, or - with use_counters - from
// the level that just ran (counters) and the state before it. With `params`
// null only the features are written; otherwise the allowed choice
// (allowed_mask bit c; choice 2 also needs n_f <= chain_cap and m_f <=
// chain_edge_cap) with the
// smallest predicted log-time. The counters are the sum of `slots` counter
// slots (slots <= 64). With `host` (and use_counters), the kernel also
// publishes the counters and the decision there as bfs_publish does, clears
// the slots for the next level and sets host->seq = seq last.
extern "C" __global__ __launch_bounds__(64) void sched_bfs_decide(
    const float* __restrict__ params, MlpNorm norm, BfsLevelState st,
    BfsCountersView* __restrict__ counters, uint32_t slots, uint32_t use_counters, uint32_t allowed_mask,
    uint32_t chain_cap, unsigned long long chain_edge_cap, SchedDecision* __restrict__ out,
    BfsPublishedView* __restrict__ host, uint32_t seq) {
  if (blockIdx.x != 0) return;
  const uint32_t lane = lane_id();
  // every lane derives the same state and features
  BfsCountersView c{0u, 0u, 0ull};
  if (use_counters) {
    uint32_t size = 0;
    unsigned long long edges = 0;
    if (lane < slots) {
      size = counters[lane].next_size;
      edges = counters[lane].next_edges;
    }
    c.next_size = wave_sum(size);
    c.next_edges = wave_sum(edges);
    if (host != nullptr && lane < slots) {  // each lane clears the slot it read
      counters[lane].next_size = 0;
      counters[lane].pad = 0;  // the top-down queue counter
      counters[lane].next_edges = 0;
    }
    st.prev_n_f = st.n_f;
    st.n_f = c.next_size;
    st.m_f = c.next_edges;
    st.m_u = st.m_u > c.next_edges ? st.m_u - c.next_edges : 0ull;
    st.reached += c.next_size;
    st.level += 1;
  }
  // The graph's size enters only as the fractions of it the state holds (a
  // network trained on smaller graphs then sees a larger graph's states within
  // the ranges it was trained on); the frontier's own size stays absolute, as
  // it decides between per-level overheads and per-level work.
  const double n1 = static_cast<double>(st.n) + 1.0, m1 = static_cast<double>(st.m) + 1.0;
  float f[MLP_IN];
  f[0] = static_cast<float>(log2((static_cast<double>(st.n_f) + 1.0) / n1));
  f[1] = static_cast<float>(log2((static_cast<double>(st.m_f) + 1.0) / m1));
  f[2] = lg(st.n_f);
  f[3] = lg(static_cast<double>(st.m_f));
  f[4] = static_cast<float>(log2((static_cast<double>(st.m_u) + 1.0) / m1));
  f[5] = static_cast<float>(log2((static_cast<double>(st.reached) + 1.0) / n1));
  f[6] = static_cast<float>(log2((static_cast<double>(st.m_f) + 1.0) / (static_cast<double>(st.m_u) + 1.0)));
  f[7] = static_cast<float>(log2((static_cast<double>(st.n_f) + 1.0) / (static_cast<double>(st.prev_n_f) + 1.0)));
  f[8] = lg(st.level);
  f[9] = st.bitmap ? 1.0f : 0.0f;
  f[10] = lg(static_cast<double>(st.m) / (st.n > 0 ? st.n : 1));
  f[11] = static_cast<float>(log2((static_cast<double>(st.max_degree) + 1.0) * n1 / m1));  // hub size / mean degree
  if (lane < MLP_IN) out->features[lane] = f[lane];
  // chained levels take bounded frontiers only
  if (st.n_f > chain_cap || st.m_f > chain_edge_cap) allowed_mask &= ~4u;
  if (params == nullptr) {
    if (lane == 0) out->choice = 0;
    return;
  }
  // The forward pass with one hidden neuron per lane; each neuron adds its
  // inputs in the order of mlp_forward_one, so the predictions are the same
  // numbers (a single lane did the whole pass before: about 1,500 dependent
  // multiply-adds per level).
  float xn[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; ++i) xn[i] = (f[i] - norm.mean[i]) * norm.scale[i];
  const uint32_t j = lane < MLP_H ? lane : 0u;
  float a = params[MLP_B1 + j];
#pragma unroll
  for (int i = 0; i < MLP_IN; ++i) a = fmaf(params[MLP_W1 + j * MLP_IN + i], xn[i], a);
  const float h1 = a > 0.0f ? a : 0.0f;
  a = params[MLP_B2 + j];
#pragma unroll 8
  for (int i = 0; i < MLP_H; ++i) a = fmaf(params[MLP_W2 + j * MLP_H + i], __shfl(h1, i), a);
  const float h2 = a > 0.0f ? a : 0.0f;
  float y[MLP_OUT];
#pragma unroll
  for (int k = 0; k < MLP_OUT; ++k) {
    float o = params[MLP_B3 + k];
#pragma unroll 8
    for (int i = 0; i < MLP_H; ++i) o = fmaf(params[MLP_W3 + k * MLP_H + i], __shfl(h2, i), o);
    y[k] = o;
  }
  if (lane != 0) return;
  uint32_t best = 0xFFFFFFFFu;
  for (uint32_t k = 0; k < MLP_OUT; ++k) {
    out->predicted[k] = y[k];
    if ((allowed_mask >> k) & 1u) {
      if (best == 0xFFFFFFFFu || y[k] < y[best]) best = k;
    }
  }
  out->choice = best == 0xFFFFFFFFu ? 0u : best;
  if (host != nullptr && use_counters) {
    host->counters = c;
    for (int i = 0; i < MLP_IN; ++i) host->decision.features[i] = f[i];
    for (int k = 0; k < MLP_OUT; ++k) host->decision.predicted[k] = y[k];
    host->decision.choice = best == 0xFFFFFFFFu ? 0u : best;
    __threadfence_system();
    __hip_atomic_store(&host->seq, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
