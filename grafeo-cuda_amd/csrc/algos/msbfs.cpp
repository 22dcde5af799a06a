// Multi-source BFS driver (an addition with no reference function).
//
// Each source gets exactly what bfs::bfs (crates/grafeo-cuda-algos/src/bfs.rs:42-131)
// returns for it: i32 distances (-1 unreachable, 0 at the source), the number
// of reachable nodes and the last non-empty level; sources are checked first,
// with bfs's message. Up to 128 sources run together in one traversal of the
// relabelled graph (kernels of msbfs.hip); more sources run 128 at a time.
// The direction of each level follows the rule of Beamer, Asanovic and
// Patterson (SC 2012) applied to the union of the frontiers (alpha = 14,
// beta = 24); non-symmetric CSRs stay top-down.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../runtime/common.hpp"

namespace gc {

namespace {

struct MsCounters {
  uint32_t frontier_vertices;
  uint32_t pad;
  unsigned long long frontier_edges;
  unsigned long long unexplored_edges;
};

// One 128-bit source mask per vertex (ulonglong2 in the kernels).
struct alignas(16) SourceMask {
  uint64_t lo, hi;
};

constexpr uint32_t kMaxSources = 128;
constexpr uint32_t kWaveDegree = 32;  // rows of at least this degree get a wave in bottom-up steps
constexpr double kAlpha = 14.0;
constexpr double kBeta = 24.0;
constexpr int32_t kMaxByteLevel = 254;  // MS_MAX_BYTE_LEVEL of msbfs.hip: 0xFF marks "not reached"

// One traversal of up to 128 sources (caller ids). d_dist holds count rows of
// n distances; d_reach / d_depth count entries, zeroed here.
void msbfs_chunk(Graph& g, const uint32_t* sources, uint32_t count, int32_t* d_dist,
                 uint32_t* d_reach, int32_t* d_depth) {
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  const uint32_t* off = g.rl.offsets.as<const uint32_t>();
  const uint32_t* tgt = g.rl.targets.as<const uint32_t>();
  const std::vector<uint32_t>& hoff = g.rl.host_offsets;
  const uint64_t mask_words = 4ull * n;  // u32 words of one 128-bit mask array
  auto* seen = g.buffer<SourceMask>("ms.seen", sizeof(SourceMask) * n);
  auto* visit = g.buffer<SourceMask>("ms.visit", sizeof(SourceMask) * n);
  auto* next = g.buffer<SourceMask>("ms.next", sizeof(SourceMask) * n);
  uint32_t* queue = g.buffer<uint32_t>("ms.queue", sizeof(uint32_t) * n);
  uint32_t* src_dev = g.buffer<uint32_t>("ms.sources", sizeof(uint32_t) * 2 * kMaxSources);
  auto* counters = g.buffer<MsCounters>("ms.counters", sizeof(MsCounters));
  fill_words(dev, seen, mask_words, 0);
  fill_words(dev, visit, mask_words, 0);
  fill_words(dev, next, mask_words, 0);
  // the levels as bytes (ms_write_levels), written out as distances at the end
  auto* lv = g.buffer<uint8_t>("ms.levels", 128ull * std::max<uint32_t>(n, 1));
  fill_words(dev, lv, 32ull * n, 0xFFFFFFFFu);
  bool byte_levels = true;

  std::vector<uint32_t> staged(2 * kMaxSources, 0);
  std::vector<uint32_t> frontier;
  uint64_t m_f = 0;
  for (uint32_t s = 0; s < count; ++s) {
    staged[s] = sources[s];
    staged[kMaxSources + s] = g.rl.host_inv[sources[s]];
    frontier.push_back(staged[kMaxSources + s]);
  }
  std::sort(frontier.begin(), frontier.end());
  frontier.erase(std::unique(frontier.begin(), frontier.end()), frontier.end());
  for (uint32_t v : frontier) m_f += hoff[v + 1] - hoff[v];
  copy_to_device(dev, src_dev, staged.data(), sizeof(uint32_t) * staged.size());
  copy_to_device(dev, queue, frontier.data(), sizeof(uint32_t) * frontier.size());
  launch(dev, get_kernel(dev, "msbfs", "ms_seed"), 1u, 64u, 0,
         static_cast<const uint32_t*>(src_dev + kMaxSources), count, seen, visit);
  launch(dev, get_kernel(dev, "msbfs", "ms_seed_levels"), (count + 63) / 64, 64u, 0,
         static_cast<const uint32_t*>(src_dev + kMaxSources), count, lv);

  const unsigned long long active_lo = count >= 64 ? ~0ull : ((1ull << count) - 1);
  const unsigned long long active_hi =
      count <= 64 ? 0ull : (count >= 128 ? ~0ull : ((1ull << (count - 64)) - 1));
  // first row whose degree reaches kWaveDegree (rows are sorted by degree)
  uint32_t split = 0;
  {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (hoff[mid + 1] - hoff[mid] >= kWaveDegree) hi = mid;
      else lo = mid + 1;
    }
    split = lo;
  }
  Kernel k_td = get_kernel(dev, "msbfs", "ms_topdown");
  Kernel k_but = get_kernel(dev, "msbfs", "ms_bottomup_thread");
  Kernel k_buw = get_kernel(dev, "msbfs", "ms_bottomup_wave");
  Kernel k_fin = get_kernel(dev, "msbfs", "ms_finalize");
  Kernel k_dist = get_kernel(dev, "msbfs", "ms_write_distances");
  Kernel k_lv = get_kernel(dev, "msbfs", "ms_write_levels");
  Kernel k_lv2d = get_kernel(dev, "msbfs", "ms_levels_to_dist");
  const uint32_t cu = static_cast<uint32_t>(std::max(1, dev.cu_count));
  const uint32_t resident = cu * 8;  // 256-lane workgroups
  const uint32_t out_blocks = std::min<uint32_t>((n + 255) / 256, resident * 4);
  // four distances per 16-byte store where n and the caller's buffer allow it
  const uint32_t vec4 = (n % 4 == 0 && (reinterpret_cast<uintptr_t>(d_dist) & 15u) == 0) ? 1u : 0u;
  auto levels_to_dist = [&] {
    launch(dev, k_lv2d, out_blocks, 256u, 0, g.rl.inv.as<const uint32_t>(), n, count,
           static_cast<const uint8_t*>(lv), d_dist, vec4);
  };
  uint64_t n_f = frontier.size();
  uint64_t m_u = static_cast<uint64_t>(g.m) - std::min<uint64_t>(g.m, m_f);
  bool top_down = true;
  int32_t level = 0;
  for (;;) {
    if (top_down && g.symmetric && static_cast<double>(m_f) > static_cast<double>(m_u) / kAlpha) {
      top_down = false;
    } else if (!top_down && static_cast<double>(n_f) < static_cast<double>(n) / kBeta) {
      top_down = true;
    }
    if (top_down) {
      const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((n_f + 3) / 4, resident));
      launch(dev, k_td, blocks, 256u, 0, off, tgt, static_cast<const uint32_t*>(queue),
             static_cast<uint32_t>(n_f), static_cast<const SourceMask*>(visit),
             static_cast<const SourceMask*>(seen), next);
    } else {
      if (split > 0) {
        const uint32_t blocks = std::min<uint32_t>((split + 255) / 256, resident * 4);
        launch(dev, k_but, blocks, 256u, 0, off, tgt, 0u, split, active_lo, active_hi,
               static_cast<const SourceMask*>(visit), static_cast<const SourceMask*>(seen), next);
      }
      if (split < n) {
        const uint32_t blocks = std::min<uint32_t>((n - split + 3) / 4, resident);
        launch(dev, k_buw, blocks, 256u, 0, off, tgt, split, n, active_lo, active_hi,
               static_cast<const SourceMask*>(visit), static_cast<const SourceMask*>(seen), next);
      }
    }
    fill_words(dev, counters, sizeof(MsCounters) / 4, 0);
    // one workgroup per MSF_RANGE (4096) vertices
    launch(dev, k_fin, static_cast<uint32_t>(std::max<uint64_t>(1, (static_cast<uint64_t>(n) + 4095) / 4096)),
           256u, 0, off, n, active_lo, active_hi, seen, visit, next, queue, counters);
    MsCounters c;
    copy_to_host(dev, dev.pinned, counters, sizeof(MsCounters));
    std::memcpy(&c, dev.pinned, sizeof(c));
    if (c.frontier_vertices == 0) break;
    ++level;
    if (byte_levels && level > kMaxByteLevel) {
      // deeper than a byte holds: the levels so far become distances, and later
      // levels write distances directly
      levels_to_dist();
      byte_levels = false;
    }
    if (byte_levels) {
      launch(dev, k_lv, out_blocks, 256u, 0, n, static_cast<const SourceMask*>(visit),
             static_cast<uint32_t>(level), lv);
    } else {
      launch(dev, k_dist, out_blocks, 256u, 0, g.rl.inv.as<const uint32_t>(), n,
             static_cast<const SourceMask*>(visit), level, d_dist);
    }
    n_f = c.frontier_vertices;
    m_f = c.frontier_edges;
    m_u = c.unexplored_edges;
  }
  if (byte_levels) levels_to_dist();
  fill_words(dev, d_reach, count, 0);
  fill_words(dev, d_depth, count, 0);
  const uint32_t per_source = std::max<uint32_t>(1, std::min<uint32_t>((n + 255) / 256, 64));
  launch(dev, get_kernel(dev, "msbfs", "ms_summarize"), per_source * count, 256u, 0,
         static_cast<const int32_t*>(d_dist), n, per_source, d_reach, d_depth);
}

void check_sources(const Graph& g, const uint32_t* sources, uint32_t count) {
  if (count > 0 && sources == nullptr) fail_kernel("sources is null");
  for (uint32_t s = 0; s < count; ++s) {
    if (sources[s] >= g.n) {
      fail_kernel("source node " + std::to_string(sources[s]) + " out of range (num_nodes = " +
                  std::to_string(g.n) + ")");
    }
  }
}

// All sources, 128 at a time; per-source results to host arrays (nullable).
void msbfs(Graph& g, const uint32_t* sources, uint32_t count, int32_t* d_dist, uint32_t* reachable,
           int32_t* max_depth) {
  prepare(g);
  Device& dev = *g.dev;
  uint32_t* d_stats = g.buffer<uint32_t>("ms.stats", sizeof(uint32_t) * 2 * kMaxSources);
  std::vector<uint32_t> stats(2 * kMaxSources);
  for (uint32_t first = 0; first < count; first += kMaxSources) {
    const uint32_t chunk = std::min(kMaxSources, count - first);
    msbfs_chunk(g, sources + first, chunk, d_dist + static_cast<uint64_t>(first) * g.n, d_stats,
                reinterpret_cast<int32_t*>(d_stats + kMaxSources));
    copy_to_host(dev, stats.data(), d_stats, sizeof(uint32_t) * 2 * kMaxSources);
    for (uint32_t s = 0; s < chunk; ++s) {
      if (reachable) reachable[first + s] = stats[s];
      if (max_depth) max_depth[first + s] = static_cast<int32_t>(stats[kMaxSources + s]);
    }
  }
  dev.sync();
}

}  // namespace

}  // namespace gc

using gc::guard;

extern "C" {

gc_status gc_msbfs_device(gc_graph* h, const uint32_t* sources, uint32_t num_sources,
                          int32_t* d_distances, uint32_t* reachable, int32_t* max_depth,
                          double* secs) {
  return guard([&] {
    gc::Stopwatch sw;
    gc::GraphCall call(h);
    gc::check_sources(*call.g, sources, num_sources);
    if (num_sources > 0 && d_distances == nullptr) gc::fail_kernel("distances is null");
    if (num_sources > 0) gc::msbfs(*call.g, sources, num_sources, d_distances, reachable, max_depth);
    if (secs) *secs = sw.seconds();
  });
}

gc_status gc_msbfs(gc_graph* h, const uint32_t* sources, uint32_t num_sources, int32_t* distances,
                   uint32_t* reachable, int32_t* max_depth, double* secs) {
  return guard([&] {
    gc::Stopwatch sw;
    gc::GraphCall call(h);
    gc::check_sources(*call.g, sources, num_sources);
    if (num_sources > 0 && distances == nullptr) gc::fail_kernel("distances is null");
    if (num_sources > 0) {
      gc::Graph& g = *call.g;
      const uint64_t words = static_cast<uint64_t>(num_sources) * g.n;
      int32_t* d_out = g.buffer<int32_t>("ms.out", sizeof(int32_t) * std::max<uint64_t>(words, 1));
      gc::msbfs(g, sources, num_sources, d_out, reachable, max_depth);
      gc::copy_to_host(call.dev(), distances, d_out, sizeof(int32_t) * words);
    }
    if (secs) *secs = sw.seconds();
  });
}

}  // extern "C"
