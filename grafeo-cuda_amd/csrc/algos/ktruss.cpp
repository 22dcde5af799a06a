// k-truss driver.
//
// Replaces k_truss::k_truss (crates/grafeo-cuda-algos/src/k_truss.rs:42-161):
// k < 3 is a Kernel error with the reference's message; threshold k-2; an
// edgeless graph returns no edges and 0 iterations; rounds repeat until one
// removes nothing and `iterations` counts that last round too; the result
// lists each surviving edge once as (u, v), u < v, in CSR order.
// Symmetric graphs run the DAG rounds on the degree-ordered relabelling
// (support from one enumeration of each triangle); any other CSR runs the
// reference's per-entry rounds directly. GC_KTRUSS_DIRECT=1 forces the
// per-entry rounds (used by the tests to check one against the other).
#include <algorithm>
#include <cstdlib>

#include "../runtime/common.hpp"

namespace gc {

namespace {

uint32_t read_u32(Device& dev, const uint32_t* p) {
  uint32_t v = 0;
  copy_to_host(dev, dev.pinned, p, sizeof(uint32_t));
  std::memcpy(&v, dev.pinned, sizeof(v));
  return v;
}

bool force_direct() {
  const char* env = std::getenv("GC_KTRUSS_DIRECT");
  return env != nullptr && env[0] == '1';
}

// Compacts the flagged original entries into (u, v) pairs (host, malloc'd).
void emit_pairs(Graph& g, const uint8_t* flags, uint32_t** edges, size_t* num_edges) {
  Device& dev = *g.dev;
  const uint64_t m = g.m;
  const uint32_t per_block = 256 * 16;
  const uint32_t blocks = static_cast<uint32_t>((m + per_block - 1) / per_block);
  uint32_t* block_counts = g.buffer<uint32_t>("kt.block_counts", sizeof(uint32_t) * (blocks + 1));
  uint32_t* total_dev = block_counts + blocks;
  launch(dev, get_kernel(dev, "util", "compact_count"), blocks, 256u, 0, flags, m, block_counts);
  launch(dev, get_kernel(dev, "util", "exclusive_scan_small"), 1u, 1024u, 0, block_counts, blocks,
         total_dev);
  const uint32_t total = read_u32(dev, total_dev);
  auto* host = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * 2 * std::max<size_t>(1, total)));
  if (host == nullptr) throw std::bad_alloc();
  if (total > 0) {
    uint32_t* pairs = g.buffer<uint32_t>("kt.pairs", sizeof(uint32_t) * 2 * static_cast<size_t>(total));
    launch(dev, get_kernel(dev, "util", "compact_emit_pairs"), blocks, 256u, 0, flags, m,
           static_cast<const uint32_t*>(block_counts), g.offsets.as<const uint32_t>(),
           g.targets.as<const uint32_t>(), g.n, pairs);
    try {
      copy_to_host(dev, host, pairs, sizeof(uint32_t) * 2 * static_cast<size_t>(total));
    } catch (...) {
      std::free(host);
      throw;
    }
  }
  *edges = host;
  *num_edges = total;
}

uint32_t rounds_direct(Graph& g, uint32_t threshold, uint8_t* flags) {
  Device& dev = *g.dev;
  const uint32_t n = g.n, m = g.m;
  uint32_t* support = g.buffer<uint32_t>("kt.support", sizeof(uint32_t) * m);
  uint8_t* removed = g.buffer<uint8_t>("kt.removed", (static_cast<size_t>(m) + 3) / 4 * 4);
  uint32_t* count = g.buffer<uint32_t>("kt.count", 16);
  fill_words(dev, removed, (static_cast<uint64_t>(m) + 3) / 4, 0);
  Kernel k_support = get_kernel(dev, "ktruss", "kt_support_direct");
  Kernel k_filter = get_kernel(dev, "ktruss", "kt_filter");
  Grid1D eg = thread_per_item(m);
  uint32_t iterations = 0;
  for (;;) {
    fill_words(dev, support, m, 0);
    launch(dev, k_support, eg.blocks, eg.threads, 0, g.offsets.as<const uint32_t>(),
           g.targets.as<const uint32_t>(), static_cast<const uint8_t*>(removed), n, m, support);
    fill_words(dev, count, 1, 0);
    launch(dev, k_filter, eg.blocks, eg.threads, 0, static_cast<const uint32_t*>(support), removed,
           m, threshold, count);
    ++iterations;
    if (read_u32(dev, count) == 0) break;
  }
  launch(dev, get_kernel(dev, "ktruss", "kt_emit_flags_direct"), eg.blocks, eg.threads, 0,
         g.offsets.as<const uint32_t>(), g.targets.as<const uint32_t>(), n, m,
         static_cast<const uint8_t*>(removed), flags);
  return iterations;
}

uint32_t rounds_dag(Graph& g, uint32_t threshold, uint8_t* flags) {
  Device& dev = *g.dev;
  const uint32_t n = g.n, m = g.m;
  const uint32_t* off = g.rl.offsets.as<const uint32_t>();
  const uint32_t* tgt = g.rl.targets.as<const uint32_t>();
  const uint32_t* fwd = g.rl.fwd.as<const uint32_t>();
  Grid1D vg = thread_per_item(n);
  if (!g.ktruss_rev_ready) {
    g.ktruss_rev.allocate(sizeof(uint32_t) * std::max<uint32_t>(m, 1));
    launch(dev, get_kernel(dev, "ktruss", "kt_reverse_index"), vg.blocks, vg.threads, 0, off, tgt,
           fwd, n, g.ktruss_rev.as<uint32_t>());
    g.ktruss_rev_ready = true;
  }
  uint32_t* support = g.buffer<uint32_t>("kt.support", sizeof(uint32_t) * m);
  uint8_t* removed = g.buffer<uint8_t>("kt.removed", (static_cast<size_t>(m) + 3) / 4 * 4);
  uint32_t* count = g.buffer<uint32_t>("kt.count", 16);  // [0] removed, [1..2] work counters
  fill_words(dev, removed, (static_cast<uint64_t>(m) + 3) / 4, 0);
  Kernel k_support = get_kernel(dev, "ktruss", "kt_wedge_support");
  Kernel k_finish = get_kernel(dev, "ktruss", "kt_dag_finish");
  const uint32_t bits = wedge_table_bits(g.rl);
  const uint32_t grid = std::max(1, dev.cu_count) * 8u;
  const uint32_t hub_chunk = Relabelled::kHubChunk;
  uint32_t iterations = 0;
  for (;;) {
    fill_words(dev, support, m, 0);
    fill_words(dev, count, 4, 0);
    launch(dev, k_support, grid, kWaveSize, (1u << bits) * 8u, off, tgt, fwd,
           static_cast<const uint8_t*>(removed), n, g.rl.hub_units.as<const uint32_t>(),
           g.rl.num_hub_units, hub_chunk, bits, count + 1, support);
    const uint32_t round1 = iterations == 0 ? 1 : 0;
    launch(dev, k_finish, vg.blocks, vg.threads, 0, off, tgt, fwd,
           g.ktruss_rev.as<const uint32_t>(), removed, n, static_cast<const uint32_t*>(support),
           threshold, round1, count);
    ++iterations;
    if (read_u32(dev, count) == 0) break;
  }
  Grid1D eg = thread_per_item(m);
  launch(dev, get_kernel(dev, "ktruss", "kt_emit_flags_relabelled"), eg.blocks, eg.threads, 0,
         g.offsets.as<const uint32_t>(), g.targets.as<const uint32_t>(), n, m,
         g.rl.inv.as<const uint32_t>(), off, tgt, static_cast<const uint8_t*>(removed), flags);
  return iterations;
}

}  // namespace

}  // namespace gc

using gc::guard;

extern "C" {

gc_status gc_k_truss(gc_graph* h, uint32_t k, uint32_t** edges, size_t* num_edges,
                     uint32_t* iterations, double* secs) {
  return guard([&] {
    if (k < 3) gc::fail_kernel("k must be >= 3 for k-truss decomposition");
    gc::Stopwatch sw;
    gc::GraphCall call(h);
    if (edges == nullptr || num_edges == nullptr) gc::fail_kernel("edges is null");
    gc::Graph& g = *call.g;
    *edges = nullptr;
    *num_edges = 0;
    uint32_t it = 0;
    if (g.m == 0) {
      *edges = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * 2));
      if (*edges == nullptr) throw std::bad_alloc();
    } else {
      gc::prepare(g);
      const uint32_t threshold = k - 2;
      uint8_t* flags = g.buffer<uint8_t>("kt.flags", g.m);
      it = (g.symmetric && !gc::force_direct()) ? gc::rounds_dag(g, threshold, flags)
                                                : gc::rounds_direct(g, threshold, flags);
      gc::emit_pairs(g, flags, edges, num_edges);
    }
    if (iterations) *iterations = it;
    if (secs) *secs = sw.seconds();
  });
}

}  // extern "C"
