// BFS driver.
//
// Replaces bfs::bfs (crates/grafeo-cuda-algos/src/bfs.rs:42-131): source
// checked first (>= num_nodes is a Kernel error with the reference's
// message); distances -1 for unreachable nodes, 0 at the source; `reachable`
// counts distances >= 0; `max_depth` is the last non-empty level. The
// traversal runs on the degree-ordered relabelling and switches between
// top-down (queue) and bottom-up (bitmap) steps per level with the rule of
// Beamer, Asanovic and Patterson (SC 2012): go bottom-up when the frontier's
// edges exceed 1/alpha of the unvisited vertices' edges, back top-down when
// the frontier holds fewer than n/beta vertices (alpha = 14, beta = 24).
// Distances do not depend on the schedule, so both directions give the
// reference's result bit for bit. A bottom-up step looks for parents in a
// vertex's own list, which holds its in-neighbours only when the CSR is
// symmetric; other CSRs stay top-down.
#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "../runtime/common.hpp"

namespace gc {

namespace {

struct BfsCounters {
  uint32_t next_size;
  uint32_t pad;
  unsigned long long next_edges;
};

constexpr double kAlpha = 14.0;
constexpr double kBeta = 24.0;

}  // namespace

struct BfsOutcome {
  uint32_t reachable = 0;
  int32_t max_depth = 0;
  uint32_t levels_bottom_up = 0;
};

// Runs BFS from `source` (caller's id) and writes caller-ordered distances to
// the device array `d_out`.
BfsOutcome bfs(Graph& g, uint32_t source, int32_t* d_out) {
  Device& dev = *g.dev;
  prepare(g);
  const uint32_t n = g.n;
  const uint64_t words64 = (static_cast<uint64_t>(n) + 63) / 64;
  int32_t* dist = g.buffer<int32_t>("bfs.dist", sizeof(int32_t) * n);
  uint32_t* queue[2] = {g.buffer<uint32_t>("bfs.q0", sizeof(uint32_t) * n),
                        g.buffer<uint32_t>("bfs.q1", sizeof(uint32_t) * n)};
  uint64_t* bitmap[2] = {g.buffer<uint64_t>("bfs.b0", sizeof(uint64_t) * words64),
                         g.buffer<uint64_t>("bfs.b1", sizeof(uint64_t) * words64)};
  auto* counters = g.buffer<BfsCounters>("bfs.counters", 2 * sizeof(BfsCounters));
  const uint32_t* off = g.rl.offsets.as<const uint32_t>();
  const uint32_t* tgt = g.rl.targets.as<const uint32_t>();
  const std::vector<uint32_t>& hoff = g.rl.host_offsets;

  const uint32_t src = g.rl.host_inv[source];
  Kernel k_init = get_kernel(dev, "bfs", "bfs_init");
  Kernel k_td = get_kernel(dev, "bfs", "bfs_topdown_edges");
  Kernel k_deg = get_kernel(dev, "bfs", "bfs_frontier_degrees");
  uint32_t* start = g.buffer<uint32_t>("bfs.start", sizeof(uint32_t) * n);
  uint32_t* scan_buf = g.buffer<uint32_t>("bfs.scan", sizeof(uint32_t) * scan_scratch_words(n));
  Kernel k_bu = get_kernel(dev, "bfs", "bfs_bottomup");
  Kernel k_q2b = get_kernel(dev, "bfs", "bfs_queue_to_bitmap");
  Kernel k_b2q = get_kernel(dev, "bfs", "bfs_bitmap_to_queue");
  Grid1D vgrid = thread_per_item(n);
  launch(dev, k_init, vgrid.blocks, vgrid.threads, 0, dist, n, src, queue[0]);

  const uint32_t cu = static_cast<uint32_t>(std::max(1, dev.cu_count));
  uint64_t n_f = 1;
  uint64_t m_f = hoff[src + 1] - hoff[src];
  uint64_t m_u = static_cast<uint64_t>(g.m) - m_f;
  bool top_down = true;
  int cur = 0;  // index of the current frontier buffer (queue or bitmap)
  int32_t level = 0;
  uint64_t reachable = 1;
  BfsOutcome outcome;
  const char* trace_env = std::getenv("GC_BFS_TRACE");
  const bool trace = trace_env != nullptr && trace_env[0] == '1';
  Stopwatch level_clock;
  for (;;) {
    if (top_down && g.symmetric && static_cast<double>(m_f) > static_cast<double>(m_u) / kAlpha) {
      fill_words(dev, bitmap[cur], 2 * words64, 0);
      Grid1D qg = thread_per_item(n_f);
      launch(dev, k_q2b, qg.blocks, qg.threads, 0, static_cast<const uint32_t*>(queue[cur]),
             static_cast<uint32_t>(n_f), reinterpret_cast<uint32_t*>(bitmap[cur]));
      top_down = false;
    } else if (!top_down && static_cast<double>(n_f) < static_cast<double>(n) / kBeta) {
      fill_words(dev, &counters[1], sizeof(BfsCounters) / 4, 0);
      Grid1D bg = wave_per_item((words64 + 63) / 64);
      launch(dev, k_b2q, std::min(bg.blocks, cu * 16), bg.threads, 0,
             static_cast<const uint64_t*>(bitmap[cur]), words64, queue[cur], &counters[1].next_size);
      top_down = true;
    }
    fill_words(dev, counters, sizeof(BfsCounters) / 4, 0);
    const int32_t next_level = level + 1;
    if (top_down) {
      // Edge-balanced expansion: prefix sum of the frontier's degrees, then
      // 256 consecutive frontier edges per wave.
      Grid1D qg = thread_per_item(n_f);
      launch(dev, k_deg, qg.blocks, qg.threads, 0, off, static_cast<const uint32_t*>(queue[cur]),
             static_cast<uint32_t>(n_f), start);
      exclusive_scan_u32(dev, start, n_f, scan_buf);
      Grid1D tg = wave_per_item((m_f + 255) / 256);
      launch(dev, k_td, std::min(tg.blocks, cu * 32), tg.threads, 0, off, tgt, dist,
             static_cast<const uint32_t*>(queue[cur]), static_cast<uint32_t>(n_f),
             static_cast<const uint32_t*>(start), static_cast<uint32_t>(m_f), queue[cur ^ 1],
             counters, next_level);
    } else {
      Grid1D bg = wave_per_item(words64);
      launch(dev, k_bu, std::min(bg.blocks, cu * 32), bg.threads, 0, off, tgt, dist,
             reinterpret_cast<const uint32_t*>(bitmap[cur]), bitmap[cur ^ 1], n, counters,
             next_level);
      ++outcome.levels_bottom_up;
    }
    BfsCounters c;
    copy_to_host(dev, dev.pinned, counters, sizeof(BfsCounters));
    std::memcpy(&c, dev.pinned, sizeof(c));
    if (trace) {
      std::fprintf(stderr, "bfs level %d %s frontier %llu edges %llu -> next %u edges %llu (%.3f ms)\n",
                   level, top_down ? "top-down " : "bottom-up", static_cast<unsigned long long>(n_f),
                   static_cast<unsigned long long>(m_f), c.next_size, c.next_edges,
                   level_clock.seconds() * 1e3);
      level_clock = Stopwatch();
    }
    if (c.next_size == 0) break;
    level = next_level;
    reachable += c.next_size;
    m_u -= std::min<uint64_t>(m_u, c.next_edges);
    m_f = c.next_edges;
    n_f = c.next_size;
    cur ^= 1;
  }
  launch(dev, get_kernel(dev, "util", "gather_u32"), vgrid.blocks, vgrid.threads, 0,
         reinterpret_cast<const uint32_t*>(dist), g.rl.inv.as<const uint32_t>(),
         reinterpret_cast<uint32_t*>(d_out), static_cast<uint64_t>(n));
  dev.sync();
  outcome.reachable = static_cast<uint32_t>(reachable);
  outcome.max_depth = level;
  return outcome;
}

static void check_source(const Graph& g, uint32_t source) {
  if (source >= g.n) {
    fail_kernel("source node " + std::to_string(source) + " out of range (num_nodes = " +
                std::to_string(g.n) + ")");
  }
}

}  // namespace gc

using gc::guard;

extern "C" {

gc_status gc_bfs_device(gc_graph* h, uint32_t source, int32_t* d_distances, uint32_t* reachable,
                        int32_t* max_depth, double* secs) {
  return guard([&] {
    gc::Stopwatch sw;
    gc::GraphCall call(h);
    gc::check_source(*call.g, source);
    if (d_distances == nullptr) gc::fail_kernel("distances is null");
    gc::BfsOutcome r = gc::bfs(*call.g, source, d_distances);
    if (reachable) *reachable = r.reachable;
    if (max_depth) *max_depth = r.max_depth;
    if (secs) *secs = sw.seconds();
  });
}

gc_status gc_bfs(gc_graph* h, uint32_t source, int32_t* distances, uint32_t* reachable,
                 int32_t* max_depth, double* secs) {
  return guard([&] {
    gc::Stopwatch sw;
    gc::GraphCall call(h);
    gc::check_source(*call.g, source);
    if (distances == nullptr) gc::fail_kernel("distances is null");
    gc::Graph& g = *call.g;
    int32_t* d_out = g.buffer<int32_t>("bfs.out", sizeof(int32_t) * g.n);
    gc::BfsOutcome r = gc::bfs(g, source, d_out);
    gc::copy_to_host(call.dev(), distances, d_out, sizeof(int32_t) * g.n);
    if (reachable) *reachable = r.reachable;
    if (max_depth) *max_depth = r.max_depth;
    if (secs) *secs = sw.seconds();
  });
}

}  // extern "C"
