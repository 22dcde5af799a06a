// Batched triangle counting (an addition with no reference function; per
// graph it returns what triangle_counting::triangle_count returns,
// crates/grafeo-cuda-algos/src/triangle_counting.rs:40-152).
//
// A batch owns the concatenated degree-oriented DAGs of many symmetric
// graphs (kernels of batch.hip). A count is one replay of a captured HIP
// graph (reset kernel + one persistent counting kernel). The task plan (which
// rows use the LDS bitmap, how many rows or DAG edges one task holds, how many
// workgroups run) is a BatchConfig; gc_batch_tune times candidate configs on
// the batch itself and keeps the fastest. Every config gives the same counts.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "../runtime/common.hpp"

namespace gc {

namespace {

constexpr uint32_t kHeavy = 0x80000000u;      // BT_HEAVY of batch.hip
constexpr uint32_t kLocalLimit = 65536;       // BT_LOCAL_BITS: graphs with bitmap rows
constexpr uint32_t kBatchThreads = 256;       // BT_WAVES * 64

struct BatchTask {
  uint32_t graph;
  uint32_t begin;
  uint32_t end;
  uint32_t base;
};

struct BatchConfig {
  uint32_t heavy_degree;   // rows with out-degree >= this use the bitmap (0: never)
  uint32_t light_edges;    // DAG edges per light task (rows are added until reached)
  uint32_t heavy_rows;     // rows per bitmap task
  uint32_t blocks_per_cu;  // workgroups per CU of the persistent launch
};

constexpr BatchConfig kDefaultConfig{32, 2048, 4, 8};

}  // namespace

struct Batch {
  Device* dev = nullptr;  // holds one reference
  uint32_t num_graphs = 0;
  uint32_t total_nodes = 0;
  std::vector<uint32_t> node_base;  // num_graphs + 1
  std::vector<uint32_t> dag_off;    // total_nodes + 1 (host copy, for planning)
  std::vector<uint64_t> row_work;   // per row: d+(u) + sum over v in L(u) of d+(v)
  DeviceMemory d_off, d_tgt, d_rows, d_tasks, d_counts, d_work;
  BatchConfig config = kDefaultConfig;
  bool tuned = false;
  uint32_t num_tasks = 0;
  uint32_t grid = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void drop_graph() {
    if (exec != nullptr) (void)hipGraphExecDestroy(exec);
    if (graph != nullptr) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~Batch() { drop_graph(); }
};

namespace {

Batch* batch_from_handle(gc_batch* h) {
  if (h == nullptr) fail_kernel("batch is null");
  return reinterpret_cast<Batch*>(h);
}

struct BatchCall {
  Batch* b;
  std::unique_lock<std::recursive_mutex> lock;
  explicit BatchCall(gc_batch* h) : b(batch_from_handle(h)), lock(b->dev->mu) { b->dev->make_current(); }
};

// Degree-oriented DAG of one symmetric graph with ids base + rank, where rank
// orders vertices by (degree, id); rows sorted; self-loops dropped.
void build_dag(const uint32_t* off, const uint32_t* tgt, uint32_t n, uint32_t base,
               uint32_t* row_len, std::vector<uint32_t>& out_tgt, std::vector<uint32_t>& row_start) {
  std::vector<uint32_t> rank(n), order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return off[a + 1] - off[a] < off[b + 1] - off[b]; });
  for (uint32_t i = 0; i < n; ++i) rank[order[i]] = i;
  row_start.assign(n + 1, 0);
  for (uint32_t r = 0; r < n; ++r) {
    const uint32_t u = order[r];
    uint32_t c = 0;
    for (uint32_t e = off[u]; e < off[u + 1]; ++e) c += rank[tgt[e]] > r;
    row_len[r] = c;
    row_start[r + 1] = row_start[r] + c;
  }
  out_tgt.resize(row_start[n]);
  for (uint32_t r = 0; r < n; ++r) {
    const uint32_t u = order[r];
    uint32_t* dst = out_tgt.data() + row_start[r];
    uint32_t c = 0;
    for (uint32_t e = off[u]; e < off[u + 1]; ++e) {
      const uint32_t x = rank[tgt[e]];
      if (x > r) dst[c++] = base + x;
    }
    std::sort(dst, dst + c);
  }
}

bool is_symmetric(const uint32_t* off, const uint32_t* tgt, uint32_t n) {
  for (uint32_t u = 0; u < n; ++u) {
    for (uint32_t e = off[u]; e < off[u + 1]; ++e) {
      const uint32_t v = tgt[e];
      if (!std::binary_search(tgt + off[v], tgt + off[v + 1], u)) return false;
    }
  }
  return true;
}

// Rows and tasks for `cfg`; returns the number of tasks (uploaded).
uint32_t plan(Batch& b, const BatchConfig& cfg) {
  std::vector<uint32_t> rows;
  struct Planned {
    BatchTask task;
    uint64_t work;
  };
  std::vector<Planned> tasks;
  rows.reserve(b.total_nodes);
  for (uint32_t g = 0; g < b.num_graphs; ++g) {
    const uint32_t lo = b.node_base[g], hi = b.node_base[g + 1];
    const bool bitmap_ok = hi - lo <= kLocalLimit && cfg.heavy_degree > 0;
    // bitmap rows
    if (bitmap_ok) {
      const uint32_t first = static_cast<uint32_t>(rows.size());
      for (uint32_t u = lo; u < hi; ++u)
        if (b.dag_off[u + 1] - b.dag_off[u] >= std::max(cfg.heavy_degree, 2u)) rows.push_back(u);
      const uint32_t last = static_cast<uint32_t>(rows.size());
      for (uint32_t r = first; r < last; r += cfg.heavy_rows) {
        const uint32_t e = std::min(last, r + cfg.heavy_rows);
        uint64_t w = 0;
        for (uint32_t i = r; i < e; ++i) w += b.row_work[rows[i]];
        tasks.push_back({{g | kHeavy, r, e, lo}, w});
      }
    }
    // merge rows (out-degree 2 or more: a row with one out-neighbour closes no triangle)
    const uint32_t first = static_cast<uint32_t>(rows.size());
    for (uint32_t u = lo; u < hi; ++u) {
      const uint32_t d = b.dag_off[u + 1] - b.dag_off[u];
      if (d >= 2 && !(bitmap_ok && d >= cfg.heavy_degree)) rows.push_back(u);
    }
    const uint32_t last = static_cast<uint32_t>(rows.size());
    uint32_t r = first;
    while (r < last) {
      uint32_t e = r;
      uint64_t edges = 0, w = 0;
      while (e < last && (edges < cfg.light_edges || e == r)) {
        edges += b.dag_off[rows[e] + 1] - b.dag_off[rows[e]];
        w += b.row_work[rows[e]];
        ++e;
      }
      tasks.push_back({{g, r, e, lo}, w});
      r = e;
    }
  }
  std::stable_sort(tasks.begin(), tasks.end(),
                   [](const Planned& x, const Planned& y) { return x.work > y.work; });
  std::vector<BatchTask> flat(tasks.size());
  for (size_t i = 0; i < tasks.size(); ++i) flat[i] = tasks[i].task;
  Device& dev = *b.dev;
  b.d_rows.ensure(sizeof(uint32_t) * std::max<size_t>(rows.size(), 1));
  b.d_tasks.ensure(sizeof(BatchTask) * std::max<size_t>(flat.size(), 1));
  if (!rows.empty()) copy_to_device(dev, b.d_rows.get(), rows.data(), sizeof(uint32_t) * rows.size());
  if (!flat.empty()) copy_to_device(dev, b.d_tasks.get(), flat.data(), sizeof(BatchTask) * flat.size());
  return static_cast<uint32_t>(flat.size());
}

void launch_raw(Device& dev, const Kernel& k, uint32_t grid, uint32_t block, void** params) {
  hip_check(hipModuleLaunchKernel(k.fn, grid, 1, 1, block, 1, 1, 0, dev.stream, params, nullptr),
            k.name);
}

// Plans `cfg` and captures reset + count into an executable graph.
void instantiate(Batch& b, const BatchConfig& cfg) {
  Device& dev = *b.dev;
  b.drop_graph();
  b.config = cfg;
  b.num_tasks = plan(b, cfg);
  b.grid = std::max(1u, static_cast<uint32_t>(std::max(1, dev.cu_count)) * cfg.blocks_per_cu);
  Kernel k_reset = get_kernel(dev, "batch", "bt_reset");
  Kernel k_count = get_kernel(dev, "batch", "bt_count");
  void* counts = b.d_counts.get();
  void* work = b.d_work.get();
  uint32_t num_graphs = b.num_graphs;
  const void* off = b.d_off.get();
  const void* tgt = b.d_tgt.get();
  const void* rows = b.d_rows.get();
  const void* tasks = b.d_tasks.get();
  uint32_t num_tasks = b.num_tasks;
  void* reset_params[] = {&counts, &num_graphs, &work};
  void* count_params[] = {&off, &tgt, &rows, &tasks, &num_tasks, &work, &counts};
  const uint32_t reset_grid = std::max(1u, std::min<uint32_t>((num_graphs + 255) / 256, 1024));
  dev.sync();
  hip_check(hipStreamBeginCapture(dev.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  hipError_t err = hipModuleLaunchKernel(k_reset.fn, reset_grid, 1, 1, 256, 1, 1, 0, dev.stream,
                                         reset_params, nullptr);
  if (err == hipSuccess) {
    err = hipModuleLaunchKernel(k_count.fn, b.grid, 1, 1, kBatchThreads, 1, 1, 0, dev.stream,
                                count_params, nullptr);
  }
  hipGraph_t graph = nullptr;
  const hipError_t end = hipStreamEndCapture(dev.stream, &graph);
  hip_check(err, "batched triangle count capture");
  hip_check(end, "hipStreamEndCapture");
  b.graph = graph;
  hip_check(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

// One replay; with `ms`, its device time from events around the launch.
void replay(Batch& b, float* ms) {
  Device& dev = *b.dev;
  hipEvent_t start = nullptr, stop = nullptr;
  if (ms != nullptr) {
    hip_check(hipEventCreate(&start), "hipEventCreate");
    hip_check(hipEventCreate(&stop), "hipEventCreate");
    hip_check(hipEventRecord(start, dev.stream), "hipEventRecord");
  }
  hipEvent_t prof = nullptr;
  profile_begin(dev, "bt_count", &prof);
  hip_check(hipGraphLaunch(b.exec, dev.stream), "hipGraphLaunch");
  profile_end(dev, "bt_count", prof);
  if (ms != nullptr) {
    hip_check(hipEventRecord(stop, dev.stream), "hipEventRecord");
    hip_check(hipEventSynchronize(stop), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(ms, start, stop), "hipEventElapsedTime");
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
  }
}

float time_config(Batch& b, const BatchConfig& cfg) {
  instantiate(b, cfg);
  replay(b, nullptr);  // warm
  float best = 1e30f;
  for (int i = 0; i < 3; ++i) {
    float ms = 0.0f;
    replay(b, &ms);
    best = std::min(best, ms);
  }
  return best;
}

}  // namespace

}  // namespace gc

using gc::guard;

extern "C" {

gc_status gc_batch_from_csrs(gc_device* dev_handle, uint32_t num_graphs,
                             const uint32_t* const* offsets, const uint32_t* num_nodes,
                             const uint32_t* const* targets, const uint64_t* num_targets,
                             gc_batch** out) {
  return guard([&] {
    if (out == nullptr) gc::fail_kernel("out is null");
    *out = nullptr;
    if (dev_handle == nullptr) gc::fail_kernel("device is null");
    if (num_graphs > 0 && (offsets == nullptr || num_nodes == nullptr || targets == nullptr ||
                           num_targets == nullptr)) {
      gc::fail_kernel("null array with non-zero length");
    }
    // validate every graph first (nothing is allocated on the device before)
    uint64_t total_nodes = 0;
    for (uint32_t g = 0; g < num_graphs; ++g) {
      if (offsets[g] == nullptr || (targets[g] == nullptr && num_targets[g] > 0)) {
        gc::fail_kernel("graph " + std::to_string(g) + ": null array with non-zero length");
      }
      try {
        gc::validate_csr(offsets[g], size_t(num_nodes[g]) + 1, targets[g], num_targets[g]);
      } catch (gc::Error& e) {
        e.message += " (graph " + std::to_string(g) + ")";
        throw;
      }
      if (!gc::is_symmetric(offsets[g], targets[g], num_nodes[g])) {
        gc::fail_kernel("batched triangle counting needs symmetric CSRs (graph " +
                        std::to_string(g) + ")");
      }
      total_nodes += num_nodes[g];
    }
    if (total_nodes >= 0xFFFFFFFFull) gc::fail_kernel("batch has too many vertices for u32 ids");
    auto* dev = reinterpret_cast<gc::Device*>(dev_handle);
    std::unique_ptr<gc::Batch> b(new gc::Batch());
    b->num_graphs = num_graphs;
    b->total_nodes = static_cast<uint32_t>(total_nodes);
    b->node_base.assign(num_graphs + 1, 0);
    for (uint32_t g = 0; g < num_graphs; ++g) b->node_base[g + 1] = b->node_base[g] + num_nodes[g];
    // per-graph DAGs, built in parallel, then concatenated
    std::vector<std::vector<uint32_t>> tg(num_graphs), starts(num_graphs);
    std::vector<uint32_t> row_len(b->total_nodes);
    gc::parallel_for(
        num_graphs,
        [&](uint64_t lo, uint64_t hi) {
          for (uint64_t g = lo; g < hi; ++g) {
            gc::build_dag(offsets[g], targets[g], num_nodes[g], b->node_base[g],
                          row_len.data() + b->node_base[g], tg[g], starts[g]);
          }
        },
        1);
    uint64_t total_entries = 0;
    for (uint32_t g = 0; g < num_graphs; ++g) total_entries += tg[g].size();
    if (total_entries >= 0xFFFFFFFFull) gc::fail_kernel("batch has too many edges for u32 offsets");
    b->dag_off.assign(b->total_nodes + 1, 0);
    for (uint32_t u = 0; u < b->total_nodes; ++u) b->dag_off[u + 1] = b->dag_off[u] + row_len[u];
    std::vector<uint32_t> dag_tgt(total_entries);
    for (uint32_t g = 0; g < num_graphs; ++g) {
      std::copy(tg[g].begin(), tg[g].end(), dag_tgt.begin() + b->dag_off[b->node_base[g]]);
      std::vector<uint32_t>().swap(tg[g]);
    }
    b->row_work.assign(b->total_nodes, 0);
    gc::parallel_for(b->total_nodes, [&](uint64_t lo, uint64_t hi) {
      for (uint64_t u = lo; u < hi; ++u) {
        uint64_t w = b->dag_off[u + 1] - b->dag_off[u];
        for (uint32_t e = b->dag_off[u]; e < b->dag_off[u + 1]; ++e) {
          const uint32_t v = dag_tgt[e];
          w += b->dag_off[v + 1] - b->dag_off[v];
        }
        b->row_work[u] = w;
      }
    });
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    b->d_off.allocate(sizeof(uint32_t) * b->dag_off.size());
    b->d_tgt.allocate(sizeof(uint32_t) * std::max<uint64_t>(total_entries, 1));
    b->d_counts.allocate(sizeof(unsigned long long) * std::max<uint32_t>(num_graphs, 1));
    b->d_work.allocate(256);
    gc::copy_to_device(*dev, b->d_off.get(), b->dag_off.data(), sizeof(uint32_t) * b->dag_off.size());
    if (total_entries > 0)
      gc::copy_to_device(*dev, b->d_tgt.get(), dag_tgt.data(), sizeof(uint32_t) * total_entries);
    b->dev = reinterpret_cast<gc::Device*>(gc_device_retain(dev_handle));
    *out = reinterpret_cast<gc_batch*>(b.release());
  });
}

void gc_batch_free(gc_batch* h) {
  if (h == nullptr) return;
  auto* b = reinterpret_cast<gc::Batch*>(h);
  gc::Device* dev = b->dev;
  {
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    (void)hipSetDevice(dev->ordinal);
    (void)hipStreamSynchronize(dev->stream);
    delete b;
  }
  gc_device_release(reinterpret_cast<gc_device*>(dev));
}

uint32_t gc_batch_num_graphs(const gc_batch* h) {
  return h ? reinterpret_cast<const gc::Batch*>(h)->num_graphs : 0;
}

gc_status gc_batch_set_config(gc_batch* h, const uint32_t* config) {
  return guard([&] {
    gc::BatchCall call(h);
    if (config == nullptr) gc::fail_kernel("config is null");
    const gc::BatchConfig cfg{config[0], config[1], config[2], config[3]};
    if (cfg.light_edges == 0 || cfg.heavy_rows == 0 || cfg.blocks_per_cu == 0 || cfg.blocks_per_cu > 64) {
      gc::fail_kernel("invalid batch config (light_edges, heavy_rows >= 1; blocks_per_cu in 1..64)");
    }
    gc::instantiate(*call.b, cfg);
    call.b->tuned = true;
  });
}

gc_status gc_batch_get_config(gc_batch* h, uint32_t* config) {
  return guard([&] {
    gc::BatchCall call(h);
    if (config == nullptr) gc::fail_kernel("config is null");
    const gc::BatchConfig& c = call.b->config;
    config[0] = c.heavy_degree;
    config[1] = c.light_edges;
    config[2] = c.heavy_rows;
    config[3] = c.blocks_per_cu;
  });
}

gc_status gc_batch_tune(gc_batch* h, uint32_t* config, double* best_ms) {
  return guard([&] {
    gc::BatchCall call(h);
    gc::Batch& b = *call.b;
    // coordinate descent over the four knobs, from the default
    gc::BatchConfig best = gc::kDefaultConfig;
    float best_t = gc::time_config(b, best);
    const uint32_t heavy_opts[] = {0, 8, 16, 32, 64, 128};
    const uint32_t light_opts[] = {256, 1024, 2048, 8192};
    const uint32_t rows_opts[] = {1, 4, 16};
    const uint32_t blocks_opts[] = {4, 8, 16};
    for (int pass = 0; pass < 2; ++pass) {
      for (uint32_t v : heavy_opts) {
        gc::BatchConfig c = best;
        c.heavy_degree = v;
        const float t = gc::time_config(b, c);
        if (t < best_t) best_t = t, best = c;
      }
      for (uint32_t v : light_opts) {
        gc::BatchConfig c = best;
        c.light_edges = v;
        const float t = gc::time_config(b, c);
        if (t < best_t) best_t = t, best = c;
      }
      for (uint32_t v : rows_opts) {
        gc::BatchConfig c = best;
        c.heavy_rows = v;
        const float t = gc::time_config(b, c);
        if (t < best_t) best_t = t, best = c;
      }
      for (uint32_t v : blocks_opts) {
        gc::BatchConfig c = best;
        c.blocks_per_cu = v;
        const float t = gc::time_config(b, c);
        if (t < best_t) best_t = t, best = c;
      }
    }
    gc::instantiate(b, best);
    b.tuned = true;
    if (config != nullptr) {
      config[0] = best.heavy_degree;
      config[1] = best.light_edges;
      config[2] = best.heavy_rows;
      config[3] = best.blocks_per_cu;
    }
    if (best_ms != nullptr) *best_ms = best_t;
  });
}

gc_status gc_batch_triangle_count(gc_batch* h, uint64_t* counts, double* secs) {
  return guard([&] {
    gc::Stopwatch sw;
    gc::BatchCall call(h);
    gc::Batch& b = *call.b;
    if (counts == nullptr && b.num_graphs > 0) gc::fail_kernel("counts is null");
    if (b.num_graphs > 0) {
      if (b.exec == nullptr) gc::instantiate(b, b.config);
      gc::replay(b, nullptr);
      gc::copy_to_host(*b.dev, counts, b.d_counts.get(), sizeof(uint64_t) * b.num_graphs);
    }
    if (secs) *secs = sw.seconds();
  });
}

}  // extern "C"
