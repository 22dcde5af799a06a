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
#include <cstring>

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

// Compacts the flagged original entries into (u, v) pairs (host, malloc'd),
// and with `values`, values[entry] of the same entries into *out_values.
void emit_pairs(Graph& g, const uint8_t* flags, uint32_t** edges, size_t* num_edges,
                const uint32_t* values = nullptr, uint32_t** out_values = nullptr) {
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
  uint32_t* host_values = nullptr;
  if (values != nullptr) {
    host_values = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * std::max<size_t>(1, total)));
    if (host_values == nullptr) {
      std::free(host);
      throw std::bad_alloc();
    }
  }
  if (total > 0) {
    try {
      uint32_t* pairs = g.buffer<uint32_t>("kt.pairs", sizeof(uint32_t) * 2 * static_cast<size_t>(total));
      uint32_t* dev_values =
          values != nullptr ? g.buffer<uint32_t>("kt.values", sizeof(uint32_t) * static_cast<size_t>(total))
                            : nullptr;
      launch(dev, get_kernel(dev, "util", "compact_emit_pairs"), blocks, 256u, 0, flags, m,
             static_cast<const uint32_t*>(block_counts), g.offsets.as<const uint32_t>(),
             g.targets.as<const uint32_t>(), g.n, pairs, values, dev_values);
      copy_to_host(dev, host, pairs, sizeof(uint32_t) * 2 * static_cast<size_t>(total));
      if (values != nullptr)
        copy_to_host(dev, host_values, dev_values, sizeof(uint32_t) * static_cast<size_t>(total));
    } catch (...) {
      std::free(host);
      std::free(host_values);
      throw;
    }
  }
  *edges = host;
  *num_edges = total;
  if (out_values != nullptr) *out_values = host_values;
}

// Reverse entry of every forward relabelled entry (built once per graph).
void ensure_reverse_index(Graph& g) {
  if (g.ktruss_rev_ready) return;
  Device& dev = *g.dev;
  Grid1D vg = thread_per_item(g.n);
  g.ktruss_rev.allocate(sizeof(uint32_t) * std::max<uint32_t>(g.m, 1));
  launch(dev, get_kernel(dev, "ktruss", "kt_reverse_index"), vg.blocks, vg.threads, 0,
         g.rl.offsets.as<const uint32_t>(), g.rl.targets.as<const uint32_t>(),
         g.rl.fwd.as<const uint32_t>(), g.n, g.ktruss_rev.as<uint32_t>());
  g.ktruss_rev_ready = true;
}

// The middle side of the supports (kt_mid_support) unless GC_KTRUSS_MID=0: an
// edge u->v with v among the hot ids goes there when its walk of L(u) above v,
// plus GC_KTRUSS_MID_COST entries (default 16), is shorter than L(v);
// GC_KTRUSS_MID_UNIT (default 8192) entries of a hot row's in-list make one
// work unit. Built once per graph (again if the cost or the unit change). At
// R-MAT scale 22 (k = 4): supports 69.5 ms without it, 15.7 + 21.0 ms with it
// (kt_bitmap_support + kt_mid_support), the call 0.186 -> 0.153 s; costs 0, 16
// and 64 and units of 2048 and 8192 entries within 1%.
bool kt_mid_on(uint32_t hot_base, uint32_t n) {
  const char* env = std::getenv("GC_KTRUSS_MID");
  return (env == nullptr || std::strcmp(env, "0") != 0) && hot_base < n;
}

void ensure_kt_mid(Graph& g, uint32_t hot_base) {
  const char* ce = std::getenv("GC_KTRUSS_MID_COST");
  const char* ue = std::getenv("GC_KTRUSS_MID_UNIT");
  const uint32_t cost = ce == nullptr ? 16u : static_cast<uint32_t>(std::strtoul(ce, nullptr, 10));
  const uint32_t unit = std::max(64u, ue == nullptr ? 8192u : static_cast<uint32_t>(std::strtoul(ue, nullptr, 10)));
  if (g.kt_mid_ready && g.kt_mid_key[0] == cost && g.kt_mid_key[1] == unit) return;
  g.kt_mid_ready = false;
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  const std::vector<uint32_t>& off = g.rl.host_offsets;
  const uint64_t m = off[n];
  const uint32_t rec_base = off[hot_base];
  const uint64_t recs = m - rec_base;
  g.kt_mid_rec.ensure(sizeof(uint32_t) * 2 * std::max<uint64_t>(recs, 1));
  const uint64_t bit_words = (m + 31) / 32 + 1;
  g.kt_midbits.ensure(sizeof(uint32_t) * bit_words);
  fill_words(dev, g.kt_midbits.as<uint32_t>(), bit_words, 0);
  const uint64_t walk_words = (static_cast<uint64_t>(n - hot_base) + 3) / 4;
  DeviceMemory walks;
  walks.allocate(4 * std::max<uint64_t>(walk_words, 1));
  fill_words(dev, walks.as<uint32_t>(), walk_words, 0);
  std::vector<uint8_t> has_walk(n - hot_base, 0);
  if (recs > 0) {
    Grid1D rg = thread_per_item(recs);
    launch(dev, get_kernel(dev, "ktruss", "kt_mid_records"), rg.blocks, rg.threads, 0,
           g.rl.offsets.as<const uint32_t>(), g.rl.targets.as<const uint32_t>(), g.rl.fwd.as<const uint32_t>(), n,
           hot_base, cost, g.kt_mid_rec.as<uint32_t>() /* uint2 */, g.kt_midbits.as<uint32_t>(), walks.as<uint8_t>());
    copy_to_host(dev, has_walk.data(), walks.as<uint8_t>(), has_walk.size());
  }
  std::vector<uint32_t> units;
  for (uint32_t v = n; v-- > hot_base;) {
    if (!has_walk[v - hot_base]) continue;
    for (uint32_t b = off[v]; b < off[v + 1]; b += unit) {
      units.insert(units.end(), {v, b, std::min<uint32_t>(off[v + 1], b + unit), 0u});
    }
  }
  g.kt_mid_num_units = static_cast<uint32_t>(units.size() / 4);
  g.kt_mid_units.ensure(sizeof(uint32_t) * std::max<size_t>(units.size(), 4));
  if (!units.empty()) copy_to_device(dev, g.kt_mid_units.as<uint32_t>(), units.data(), sizeof(uint32_t) * units.size());
  dev.sync();  // walks is released on return
  g.kt_mid_rec_base = rec_base;
  g.kt_mid_key[0] = cost;
  g.kt_mid_key[1] = unit;
  g.kt_mid_ready = true;
}

// support[e] += the triangles of every forward entry e of the relabelled CSR
// (support zeroed, every entry active): the walk on the hub bitmap
// (kt_bitmap_support) unless GC_KTRUSS_SUPPORT=wedge (the enumeration at the
// middle vertex, kt_wedge_support). `counters` holds 16 zeroed words. At
// R-MAT scale 22: bitmap 69.5 ms, wedge 289-295 ms (same supports).
void support_pass(Graph& g, uint32_t* support, uint32_t* counters) {
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  const uint32_t* off = g.rl.offsets.as<const uint32_t>();
  const uint32_t* tgt = g.rl.targets.as<const uint32_t>();
  const uint32_t* fwd = g.rl.fwd.as<const uint32_t>();
  const char* env = std::getenv("GC_KTRUSS_SUPPORT");
  if (env != nullptr && std::strcmp(env, "wedge") == 0) {
    const uint32_t bits = wedge_table_bits(g.rl);
    const uint32_t grid = std::max(1, dev.cu_count) * 8u;
    launch(dev, get_kernel(dev, "ktruss", "kt_wedge_support"), grid, kWaveSize, (1u << bits) * 8u, off, tgt,
           fwd, static_cast<const uint8_t*>(nullptr), n, g.rl.hub_units.as<const uint32_t>(), g.rl.num_hub_units,
           Relabelled::kHubChunk, bits, counters + 1, support);
    return;
  }
  uint32_t hot_base = 0, heavy = 0, parts = 1, split_min = 0;
  tc_bitmap_schedule(g, &hot_base, &heavy, &parts, &split_min);
  constexpr uint32_t kStride = 64, kMaxParts = 8;  // KTB_COUNTER_STRIDE of ktruss.hip; parts <= 8
  uint32_t* work = g.buffer<uint32_t>("kt.work", sizeof(uint32_t) * kStride * kMaxParts);
  fill_words(dev, work, kStride * kMaxParts, 0);
  bool mid = kt_mid_on(hot_base, n);
  if (mid) ensure_kt_mid(g, hot_base);
  mid = mid && g.kt_mid_num_units > 0;  // no walk anywhere: the bits would only skip empty ones
  // three 43 KB workgroups fit a CU's LDS; one more each waits for a slot
  const uint32_t grid = std::max(1, dev.cu_count) * 4u;
  launch(dev, get_kernel(dev, "ktruss", "kt_bitmap_support"), grid, 256u, 0, fwd, off + 1, tgt, n, hot_base,
         heavy, 64u, parts, split_min, work, support,
         mid ? static_cast<const uint32_t*>(g.kt_midbits.as<uint32_t>()) : nullptr);
  if (mid) {
    // one-wave workgroups (10.2 KB of LDS each), on work counter word 2 of part 0's line
    const uint32_t blocks = std::min<uint32_t>(std::max(1, dev.cu_count) * 12u, g.kt_mid_num_units);
    launch(dev, get_kernel(dev, "ktruss", "kt_mid_support"), blocks, kWaveSize, 0, off, tgt, fwd, hot_base,
           g.kt_mid_rec.as<const uint32_t>() /* uint2 */, g.kt_mid_rec_base,
           g.kt_mid_units.as<const uint32_t>() /* uint4 */, g.kt_mid_num_units, work + 2, support);
  }
}

// Lanes per frontier edge of kt_truss_peel over `blocks` 256-lane workgroups:
// 16 when the batch gives every 16-lane group an edge, else a whole wave.
uint32_t peel_width(uint32_t size, uint32_t blocks) {
  return static_cast<uint64_t>(size) >= static_cast<uint64_t>(blocks) * 256u / 16u ? 16u : 64u;
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

// The reference's rounds on a symmetric graph, with the supports computed
// once: every triangle is enumerated at its middle vertex, and each later
// round only destroys the triangles of the edges it removes (kt_truss_peel:
// the surviving edges of a destroyed triangle lose one support each, once per
// triangle). Round r removes every live edge whose support, among live edges,
// is below the threshold; the first round also counts the self-loops of an
// edge's endpoints and removes the self-loops. An edge falls below the
// threshold in round r >= 3 only by losing support in round r - 1, so the
// peel's list of such edges is the next frontier; round 2 adds the edges that
// only the self-loops held above it. `iterations` counts the last round,
// which removes nothing. (Recomputing every support each round, as the
// reference does, cost 8 triangle enumerations for the 4-truss of R-MAT
// scale 22.)
uint32_t rounds_dag(Graph& g, uint32_t threshold, uint8_t* flags) {
  Device& dev = *g.dev;
  const uint32_t n = g.n, m = g.m;
  const uint32_t* off = g.rl.offsets.as<const uint32_t>();
  const uint32_t* tgt = g.rl.targets.as<const uint32_t>();
  const uint32_t* fwd = g.rl.fwd.as<const uint32_t>();
  ensure_reverse_index(g);
  const uint64_t m_o = g.rl.oriented_entries;
  uint32_t* support = g.buffer<uint32_t>("kt.support", sizeof(uint32_t) * m);
  uint8_t* state = g.buffer<uint8_t>("kt.removed", (static_cast<size_t>(m) + 3) / 4 * 4);
  uint32_t* canon = g.buffer<uint32_t>("kt.canon", sizeof(uint32_t) * m);
  uint32_t* queue[2] = {g.buffer<uint32_t>("kt.queue0", sizeof(uint32_t) * std::max<uint64_t>(m_o, 1)),
                        g.buffer<uint32_t>("kt.queue1", sizeof(uint32_t) * std::max<uint64_t>(m_o, 1))};
  uint32_t* counters = g.buffer<uint32_t>("kt.count", 16);
  fill_words(dev, state, (static_cast<uint64_t>(m) + 3) / 4, 0);
  fill_words(dev, support, m, 0);
  fill_words(dev, counters, 4, 0);
  const uint32_t grid = std::max(1, dev.cu_count) * 8u;
  support_pass(g, support, counters);  // every entry is active: no removal flags to read
  Grid1D vg = thread_per_item(n);
  launch(dev, get_kernel(dev, "ktruss", "kt_canonical_index"), vg.blocks, vg.threads, 0, off, fwd, n, canon,
         state);
  launch(dev, get_kernel(dev, "ktruss", "kt_canonical_reverse"), vg.blocks, vg.threads, 0, off, fwd,
         g.ktruss_rev.as<const uint32_t>(), n, canon);

  Kernel k_frontier = get_kernel(dev, "ktruss", "kt_truss_frontier");
  Kernel k_peel = get_kernel(dev, "ktruss", "kt_truss_peel");
  Kernel k_commit = get_kernel(dev, "ktruss", "kt_truss_commit");
  const uint32_t t = threshold;
  const uint32_t frontier_blocks = static_cast<uint32_t>(std::max<uint64_t>(1, (static_cast<uint64_t>(m) + 4095) / 4096));
  // appends the live edges below t (with the self-loop terms when `loops`)
  // to queue[q] after its first `already` entries; returns the new length
  auto frontier = [&](int q, uint32_t already, uint32_t loops) {
    const uint32_t init[2] = {already, 0xFFFFFFFFu};
    copy_to_device(dev, counters, init, sizeof(init));
    launch(dev, k_frontier, frontier_blocks, 256u, 0, off, tgt, fwd, n, loops, state,
           static_cast<const uint32_t*>(support), static_cast<uint64_t>(m), t, queue[q], counters);
    uint32_t size = 0;
    copy_to_host(dev, dev.pinned, counters, sizeof(uint32_t));
    std::memcpy(&size, dev.pinned, sizeof(size));
    return size;
  };
  int cur = 0;
  uint32_t size = frontier(cur, 0, 1);
  uint32_t iterations = 1;
  if (size > 0 || g.self_loops > 0) {  // round 1 removed something
    for (uint32_t round = 1;; ++round) {
      uint32_t next = 0;
      if (size > 0) {
        fill_words(dev, counters + 2, 1, 0);
        launch(dev, k_peel, grid, 256u, 0, off, tgt, static_cast<const uint32_t*>(canon), n,
               static_cast<const uint32_t*>(queue[cur]), size, state, support, t, queue[cur ^ 1], counters + 2,
               peel_width(size, grid));
        copy_to_host(dev, dev.pinned, counters + 2, sizeof(uint32_t));
        std::memcpy(&next, dev.pinned, sizeof(next));
        Grid1D cg = thread_per_item(std::max(size, next));
        launch(dev, k_commit, cg.blocks, cg.threads, 0, static_cast<const uint32_t*>(queue[cur]), size,
               static_cast<const uint32_t*>(queue[cur ^ 1]), next, state, support, 0u);
      }
      ++iterations;  // round + 1 runs
      size = round == 1 ? frontier(cur ^ 1, next, 0) : next;
      cur ^= 1;
      if (size == 0) break;  // that round removes nothing
    }
  }
  Grid1D eg = thread_per_item(m);
  launch(dev, get_kernel(dev, "ktruss", "kt_emit_flags_state"), eg.blocks, eg.threads, 0,
         g.offsets.as<const uint32_t>(), g.targets.as<const uint32_t>(), n, m, g.rl.inv.as<const uint32_t>(), off,
         tgt, static_cast<const uint32_t*>(canon), static_cast<const uint8_t*>(state), flags);
  return iterations;
}

// Trussness of every undirected edge of a symmetric graph (peeling on the
// relabelled CSR; kernels kt_truss_* of ktruss.hip). Returns the largest
// trussness (0 without edges); the per-entry values are left in the support
// slots of the canonical entries.
uint32_t truss_peel(Graph& g, uint32_t* support) {
  Device& dev = *g.dev;
  const uint32_t n = g.n, m = g.m;
  const uint32_t* off = g.rl.offsets.as<const uint32_t>();
  const uint32_t* tgt = g.rl.targets.as<const uint32_t>();
  const uint32_t* fwd = g.rl.fwd.as<const uint32_t>();
  ensure_reverse_index(g);
  const uint64_t m_o = g.rl.oriented_entries;
  uint8_t* state = g.buffer<uint8_t>("kt.removed", (static_cast<size_t>(m) + 3) / 4 * 4);
  uint32_t* canon = g.buffer<uint32_t>("kt.canon", sizeof(uint32_t) * m);
  uint32_t* queue[2] = {g.buffer<uint32_t>("kt.queue0", sizeof(uint32_t) * std::max<uint64_t>(m_o, 1)),
                        g.buffer<uint32_t>("kt.queue1", sizeof(uint32_t) * std::max<uint64_t>(m_o, 1))};
  uint32_t* counters = g.buffer<uint32_t>("kt.count", 16);
  // supports of the whole graph: every triangle once, at its middle vertex
  fill_words(dev, state, (static_cast<uint64_t>(m) + 3) / 4, 0);
  fill_words(dev, support, m, 0);
  fill_words(dev, counters, 4, 0);
  support_pass(g, support, counters);
  Grid1D vg = thread_per_item(n);
  launch(dev, get_kernel(dev, "ktruss", "kt_canonical_index"), vg.blocks, vg.threads, 0, off, fwd, n,
         canon, state);
  launch(dev, get_kernel(dev, "ktruss", "kt_canonical_reverse"), vg.blocks, vg.threads, 0, off, fwd,
         g.ktruss_rev.as<const uint32_t>(), n, canon);

  Kernel k_frontier = get_kernel(dev, "ktruss", "kt_truss_frontier");
  Kernel k_peel = get_kernel(dev, "ktruss", "kt_truss_peel");
  Kernel k_commit = get_kernel(dev, "ktruss", "kt_truss_commit");
  const uint32_t frontier_blocks = static_cast<uint32_t>(std::max<uint64_t>(1, (static_cast<uint64_t>(m) + 4095) / 4096));
  const uint32_t peel_grid = std::max(1, dev.cu_count) * 8u;  // 4 waves per workgroup
  uint64_t alive = m_o;
  uint32_t k = 3, max_truss = m_o > 0 ? 2 : 0;
  uint32_t host_counters[4];
  auto read_counters = [&](uint32_t words) {
    copy_to_host(dev, dev.pinned, counters, sizeof(uint32_t) * words);
    std::memcpy(host_counters, dev.pinned, sizeof(uint32_t) * words);
  };
  while (alive > 0) {
    const uint32_t t = k - 2;
    const uint32_t init[2] = {0u, 0xFFFFFFFFu};
    copy_to_device(dev, counters, init, sizeof(init));
    launch(dev, k_frontier, frontier_blocks, 256u, 0, off, tgt, fwd, n, 0u, state,
           static_cast<const uint32_t*>(support), static_cast<uint64_t>(m), t, queue[0], counters);
    read_counters(2);
    uint32_t size = host_counters[0];
    if (size == 0) {
      // no live edge below t: the next level that removes anything is min support + 3
      if (host_counters[1] == 0xFFFFFFFFu) fail_kernel("truss peeling lost track of live edges");
      k = std::max(k + 1, host_counters[1] + 3);
      continue;
    }
    int cur = 0;
    while (size > 0) {
      fill_words(dev, counters + 2, 1, 0);
      launch(dev, k_peel, peel_grid, 256u, 0, off, tgt, static_cast<const uint32_t*>(canon), n,
             static_cast<const uint32_t*>(queue[cur]), size, state, support, t, queue[cur ^ 1],
             counters + 2, peel_width(size, peel_grid));
      copy_to_host(dev, dev.pinned, counters + 2, sizeof(uint32_t));
      uint32_t next = 0;
      std::memcpy(&next, dev.pinned, sizeof(next));
      const uint32_t commit_items = std::max(size, next);
      Grid1D cg = thread_per_item(commit_items);
      launch(dev, k_commit, cg.blocks, cg.threads, 0, static_cast<const uint32_t*>(queue[cur]), size,
             static_cast<const uint32_t*>(queue[cur ^ 1]), next, state, support, k - 1);
      if (size > alive) fail_kernel("truss peeling removed more edges than are live");
      alive -= size;
      size = next;
      cur ^= 1;
    }
    max_truss = k - 1;
    ++k;
  }
  return max_truss;
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

extern "C" {

gc_status gc_truss_numbers(gc_graph* h, uint32_t** edges, uint32_t** truss, size_t* num_edges,
                           uint32_t* max_truss, double* secs) {
  return guard([&] {
    gc::Stopwatch sw;
    gc::GraphCall call(h);
    if (edges == nullptr || truss == nullptr || num_edges == nullptr)
      gc::fail_kernel("edges, truss or num_edges is null");
    gc::Graph& g = *call.g;
    *edges = nullptr;
    *truss = nullptr;
    *num_edges = 0;
    uint32_t top = 0;
    if (g.m == 0) {
      *edges = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * 2));
      *truss = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t)));
      if (*edges == nullptr || *truss == nullptr) {
        std::free(*edges);
        std::free(*truss);
        *edges = nullptr;
        *truss = nullptr;
        throw std::bad_alloc();
      }
    } else {
      gc::prepare(g);
      if (!g.symmetric) gc::fail_kernel("truss numbers need a symmetric CSR");
      uint32_t* support = g.buffer<uint32_t>("kt.support", sizeof(uint32_t) * g.m);
      top = gc::truss_peel(g, support);
      uint8_t* flags = g.buffer<uint8_t>("kt.flags", g.m);
      uint32_t* values = g.buffer<uint32_t>("kt.entry_values", sizeof(uint32_t) * g.m);
      gc::Grid1D eg = gc::thread_per_item(g.m);
      launch(*g.dev, gc::get_kernel(*g.dev, "ktruss", "kt_truss_values"), eg.blocks, eg.threads, 0,
             g.offsets.as<const uint32_t>(), g.targets.as<const uint32_t>(), g.n, g.m,
             g.rl.inv.as<const uint32_t>(), g.rl.offsets.as<const uint32_t>(),
             g.rl.targets.as<const uint32_t>(), static_cast<const uint32_t*>(support), flags, values);
      gc::emit_pairs(g, flags, edges, num_edges, values, truss);
    }
    if (max_truss) *max_truss = top;
    if (secs) *secs = sw.seconds();
  });
}

}  // extern "C"
