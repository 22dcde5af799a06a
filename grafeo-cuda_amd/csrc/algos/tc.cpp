// Triangle counting drivers.
//
// Replace triangle_counting::{triangle_count, triangle_count_oriented}
// (crates/grafeo-cuda-algos/src/triangle_counting.rs:40-210): same results,
// same early return for an edgeless graph, no forward-edge ceiling (the
// reference refuses more than 134,217,727 forward edges, launch.rs:27-41).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "../runtime/common.hpp"

namespace gc {

namespace {

// Runs tc_count over the lists L(x) = targets[lo[x] .. hi[x]) and returns the count.
uint64_t run_tc_count(Graph& g, const uint32_t* lo, const uint32_t* hi, const uint32_t* targets,
                      bool descending, bool after_only) {
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  // [0..1] u64 result, [2..3] work counters
  uint32_t* scratch = g.buffer<uint32_t>("tc.scratch", 16);
  fill_words(dev, scratch, 4, 0);
  const uint32_t heavy = descending ? std::min<uint32_t>(n, 65536) : 0;
  const uint32_t light_chunk = descending ? 64 : 16;
  const uint32_t waves = std::max(1, dev.cu_count) * 32u;
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(waves, n));
  Kernel k = get_kernel(dev, "tc", "tc_count");
  auto* result = reinterpret_cast<unsigned long long*>(scratch);
  uint32_t descending_u = descending ? 1 : 0, after_u = after_only ? 1 : 0;
  launch(dev, k, grid, kWaveSize, 0, lo, hi, targets, n, descending_u, after_u, heavy, light_chunk,
         scratch + 2, result);
  uint64_t total = 0;
  copy_to_host(dev, dev.pinned, result, sizeof(uint64_t));
  std::memcpy(&total, dev.pinned, sizeof(uint64_t));
  return total;
}

}  // namespace

// Triangles of a symmetric graph by enumeration at the middle vertex of the
// relabelled DAG (tc_wedge).
uint64_t run_tc_wedge(Graph& g) {
  Device& dev = *g.dev;
  uint32_t* scratch = g.buffer<uint32_t>("tc.scratch", 16);
  fill_words(dev, scratch, 4, 0);
  const uint32_t bits = wedge_table_bits(g.rl);
  const uint32_t grid = std::max(1, dev.cu_count) * 8u;
  auto* result = reinterpret_cast<unsigned long long*>(scratch);
  const uint32_t hub_chunk = Relabelled::kHubChunk;
  launch(dev, get_kernel(dev, "tc", "tc_wedge"), grid, kWaveSize, (1u << bits) * 4u,
         g.rl.offsets.as<const uint32_t>(), g.rl.targets.as<const uint32_t>(),
         g.rl.fwd.as<const uint32_t>(), g.n, g.rl.hub_units.as<const uint32_t>(), g.rl.num_hub_units,
         hub_chunk, bits, scratch + 2, result);
  uint64_t total = 0;
  copy_to_host(dev, dev.pinned, result, sizeof(uint64_t));
  std::memcpy(&total, dev.pinned, sizeof(uint64_t));
  return total;
}

// Triangles of a symmetric graph by hash probes of each DAG row (tc_hash).
uint64_t run_tc_hash(Graph& g) {
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  uint32_t* scratch = g.buffer<uint32_t>("tc.scratch", 16);
  fill_words(dev, scratch, 4, 0);
  const uint32_t heavy = std::min<uint32_t>(n, 65536);
  const uint32_t light_chunk = 64;
  const uint32_t grid = std::max(1, dev.cu_count) * 8u;
  auto* result = reinterpret_cast<unsigned long long*>(scratch);
  launch(dev, get_kernel(dev, "tc", "tc_hash"), grid, 4 * kWaveSize, 0,
         g.rl.fwd.as<const uint32_t>(), g.rl.offsets.as<const uint32_t>() + 1,
         g.rl.targets.as<const uint32_t>(), n, heavy, light_chunk, scratch + 2, result);
  uint64_t total = 0;
  copy_to_host(dev, dev.pinned, result, sizeof(uint64_t));
  std::memcpy(&total, dev.pinned, sizeof(uint64_t));
  return total;
}

// Triangles of a symmetric graph with the top-id bitmap of each DAG row (tc_bitmap).
uint64_t run_tc_bitmap(Graph& g) {
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  constexpr uint32_t kHotBits = 32768;  // TCB_HOT_BITS of tc.hip
  // hot range [hot_base, n): at most kHotBits ids, hot_base a multiple of 32
  const uint32_t hot_base = n > kHotBits ? ((n - kHotBits) + 31u) & ~31u : 0u;
  uint32_t* scratch = g.buffer<uint32_t>("tc.scratch", 16);
  fill_words(dev, scratch, 4, 0);
  const uint32_t heavy = std::min<uint32_t>(n, 65536);
  const uint32_t light_chunk = 64;
  const uint32_t grid = std::max(1, dev.cu_count) * 8u;
  auto* result = reinterpret_cast<unsigned long long*>(scratch);
  launch(dev, get_kernel(dev, "tc", "tc_bitmap"), grid, 4 * kWaveSize, 0,
         g.rl.fwd.as<const uint32_t>(), g.rl.offsets.as<const uint32_t>() + 1,
         g.rl.targets.as<const uint32_t>(), n, hot_base, heavy, light_chunk, scratch + 2, result);
  uint64_t total = 0;
  copy_to_host(dev, dev.pinned, result, sizeof(uint64_t));
  std::memcpy(&total, dev.pinned, sizeof(uint64_t));
  return total;
}

// Which symmetric-graph method runs (GC_TC_METHOD): "bitmap" (default,
// tc_bitmap), "hash" (tc_hash), "lists" (binary-search intersection, tc_count)
// or "wedge" (enumeration at the middle vertex, tc_wedge). All four count the
// same triangles.
static const char* tc_method() {
  const char* env = std::getenv("GC_TC_METHOD");
  return env != nullptr ? env : "bitmap";
}

uint64_t triangle_count_symmetric(Graph& g) {
  prepare(g);
  if (g.symmetric) {
    const char* method = tc_method();
    if (std::strcmp(method, "hash") == 0) return run_tc_hash(g);
    if (std::strcmp(method, "wedge") == 0) return run_tc_wedge(g);
    if (std::strcmp(method, "lists") == 0) {
      // Degree-oriented DAG of the relabelled graph, hubs first.
      return run_tc_count(g, g.rl.fwd.as<const uint32_t>(), g.rl.offsets.as<const uint32_t>() + 1,
                          g.rl.targets.as<const uint32_t>(), true, true);
    }
    return run_tc_bitmap(g);
  }
  // Any other CSR: the suffix of each list above its row, in the caller's ids.
  ensure_fwd_original(g);
  return run_tc_count(g, g.fwd_original.as<const uint32_t>(), g.offsets.as<const uint32_t>() + 1,
                      g.targets.as<const uint32_t>(), false, true);
}

uint64_t triangle_count_oriented(Graph& g) {
  prepare(g);
  // sum over entries u->v of |N(u) intersect N(v)| is invariant under the
  // relabelling; the degree order lets the heaviest rows start first.
  return run_tc_count(g, g.rl.offsets.as<const uint32_t>(), g.rl.offsets.as<const uint32_t>() + 1,
                      g.rl.targets.as<const uint32_t>(), true, false);
}

}  // namespace gc

using gc::guard;

extern "C" {

gc_status gc_triangle_count(gc_graph* h, uint64_t* total_triangles, double* secs) {
  return guard([&] {
    gc::Stopwatch sw;
    gc::GraphCall call(h);
    if (total_triangles == nullptr) gc::fail_kernel("total_triangles is null");
    *total_triangles = call.g->m == 0 ? 0 : gc::triangle_count_symmetric(*call.g);
    if (secs) *secs = sw.seconds();
  });
}

gc_status gc_triangle_count_oriented(gc_graph* h, uint64_t* total_triangles, double* secs) {
  return guard([&] {
    gc::Stopwatch sw;
    gc::GraphCall call(h);
    if (total_triangles == nullptr) gc::fail_kernel("total_triangles is null");
    *total_triangles = call.g->m == 0 ? 0 : gc::triangle_count_oriented(*call.g);
    if (secs) *secs = sw.seconds();
  });
}

}  // extern "C"
