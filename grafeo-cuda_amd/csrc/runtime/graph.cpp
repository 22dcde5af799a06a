// Graph upload and derived device layouts.
//
// Upload follows GpuGraph::from_csr_slices / from_csr_arrays
// (crates/grafeo-cuda-core/src/gpu_graph.rs:46-87): validate first (a
// malformed CSR never reaches the device), then copy offsets and targets.
//
// The derived layout (built on first use or by gc_graph_prepare) has no
// reference counterpart: a degree-ordered relabelling of the graph. Vertices
// get new ids sorted by (degree, old id) ascending; each relabelled row is
// sorted, and fwd[i] marks where its entries above i start. The kernels that
// need locality (PageRank, BFS) or the degree-oriented DAG (triangle
// counting, k-truss on symmetric graphs) run on it, and results are mapped
// back to the caller's ids on the device. It is built on the device (radix
// sorts of (degree, id) and (new source, new target) keys, kernels of
// prep.hip); GC_PREPARE=host builds the same arrays on the host instead.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "common.hpp"

namespace gc {

namespace {

unsigned bits_needed(uint64_t max_value) {
  unsigned b = 0;
  while (b < 64 && (max_value >> b) != 0) ++b;
  return b;
}

// (vertex, begin, end) slices of in-lists longer than Relabelled::kHubChunk,
// heaviest vertices first.
std::vector<uint32_t> hub_slices(const std::vector<uint32_t>& r_off, const std::vector<uint32_t>& in_end) {
  std::vector<uint32_t> hubs;
  for (size_t i = in_end.size(); i-- > 0;) {
    if (in_end[i] - r_off[i] > Relabelled::kHubChunk) {
      for (uint32_t s = r_off[i]; s < in_end[i]; s += Relabelled::kHubChunk) {
        hubs.push_back(static_cast<uint32_t>(i));
        hubs.push_back(s);
        hubs.push_back(std::min(s + Relabelled::kHubChunk, in_end[i]));
      }
    }
  }
  return hubs;
}

void upload_hubs(Graph& g, const std::vector<uint32_t>& hubs) {
  g.rl.num_hub_units = static_cast<uint32_t>(hubs.size() / 3);
  g.rl.hub_units.allocate(sizeof(uint32_t) * std::max<size_t>(hubs.size(), 3));
  if (!hubs.empty()) copy_to_device(*g.dev, g.rl.hub_units.get(), hubs.data(), sizeof(uint32_t) * hubs.size());
}

void prepare_device(Graph& g) {
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  const uint64_t m = g.m;
  Relabelled& rl = g.rl;
  const uint32_t* off = g.offsets.as<const uint32_t>();
  const uint32_t* tgt = g.targets.as<const uint32_t>();
  uint32_t max_deg = 0;
  for (uint32_t v = 0; v < n; ++v) max_deg = std::max(max_deg, g.host_offsets[v + 1] - g.host_offsets[v]);
  const unsigned id_bits = std::max(1u, bits_needed(n == 0 ? 0 : n - 1));

  // symmetry and self-loops
  DeviceMemory counts;
  counts.allocate(64);
  fill_words(dev, counts.get(), 2, 0);
  if (m > 0) {
    Grid1D eg = thread_per_item(m);
    launch(dev, get_kernel(dev, "prep", "prep_symmetry"), eg.blocks, eg.threads, 0, off, tgt, n, m,
           counts.as<uint32_t>());
  }
  uint32_t host_counts[2] = {0, 0};
  copy_to_host(dev, host_counts, counts.get(), sizeof(host_counts));
  g.self_loops = host_counts[0];
  g.symmetric = host_counts[1] == 0;

  rl.perm.allocate(sizeof(uint32_t) * std::max<size_t>(n, 1));
  rl.inv.allocate(sizeof(uint32_t) * std::max<size_t>(n, 1));
  rl.offsets.allocate(sizeof(uint32_t) * (static_cast<size_t>(n) + 1));
  rl.targets.allocate(sizeof(uint32_t) * std::max<uint64_t>(m, 1));
  rl.fwd.allocate(sizeof(uint32_t) * std::max<size_t>(n, 1));
  {
    // vertex order: sort (degree, id)
    DeviceMemory ka, kb;
    ka.allocate(sizeof(unsigned long long) * std::max<size_t>(n, 1));
    kb.allocate(sizeof(unsigned long long) * std::max<size_t>(n, 1));
    Grid1D vg = thread_per_item(std::max<uint32_t>(n, 1));
    launch(dev, get_kernel(dev, "prep", "prep_degree_keys"), vg.blocks, vg.threads, 0, off, n, id_bits,
           ka.as<unsigned long long>());
    unsigned long long* sorted = radix_sort(dev, ka.as<unsigned long long>(), kb.as<unsigned long long>(), n,
                                            id_bits + std::max(1u, bits_needed(max_deg)));
    launch(dev, get_kernel(dev, "prep", "prep_perm"), vg.blocks, vg.threads, 0,
           static_cast<const unsigned long long*>(sorted), n, id_bits, rl.perm.as<uint32_t>(),
           rl.inv.as<uint32_t>());
  }
  {
    // relabelled rows: sort (inv[u], inv[v]) over every entry
    DeviceMemory ka, kb;
    ka.allocate(sizeof(unsigned long long) * std::max<uint64_t>(m, 1));
    kb.allocate(sizeof(unsigned long long) * std::max<uint64_t>(m, 1));
    unsigned long long* sorted = ka.as<unsigned long long>();
    if (m > 0) {
      Grid1D eg = thread_per_item(m);
      launch(dev, get_kernel(dev, "prep", "prep_entry_keys"), eg.blocks, eg.threads, 0, off, tgt, n, m,
             static_cast<const uint32_t*>(rl.inv.as<uint32_t>()), static_cast<uint32_t>(id_bits),
             ka.as<unsigned long long>());
      sorted = radix_sort(dev, ka.as<unsigned long long>(), kb.as<unsigned long long>(), m, 2 * id_bits);
    }
    Grid1D cg = thread_per_item(std::max<uint64_t>(m, static_cast<uint64_t>(n) + 1));
    launch(dev, get_kernel(dev, "sort", "csr_from_keys"), cg.blocks, cg.threads, 0,
           static_cast<const unsigned long long*>(sorted), m, n, static_cast<uint32_t>(id_bits),
           rl.targets.as<uint32_t>(), rl.offsets.as<uint32_t>());
  }
  if (n > 0) {
    Grid1D vg = thread_per_item(n);
    launch(dev, get_kernel(dev, "util", "csr_forward_start"), vg.blocks, vg.threads, 0,
           rl.offsets.as<const uint32_t>(), rl.targets.as<const uint32_t>(), n, rl.fwd.as<uint32_t>());
  }
  // statistics and in-list ends
  DeviceMemory stats, in_end;
  stats.allocate(sizeof(unsigned long long) * 8);
  in_end.allocate(sizeof(uint32_t) * std::max<size_t>(n, 1));
  fill_words(dev, stats.get(), 16, 0);
  if (n > 0) {
    Grid1D vg = thread_per_item(n);
    launch(dev, get_kernel(dev, "prep", "prep_stats"), vg.blocks, vg.threads, 0,
           rl.offsets.as<const uint32_t>(), rl.targets.as<const uint32_t>(), rl.fwd.as<const uint32_t>(), n,
           stats.as<unsigned long long>(), in_end.as<uint32_t>());
  }
  unsigned long long s[8];
  copy_to_host(dev, s, stats.get(), sizeof(s));
  rl.host_offsets.resize(static_cast<size_t>(n) + 1);
  rl.host_inv.resize(n);
  std::vector<uint32_t> host_in_end(n);
  copy_to_host(dev, rl.host_offsets.data(), rl.offsets.get(), sizeof(uint32_t) * rl.host_offsets.size());
  if (n > 0) {
    copy_to_host(dev, rl.host_inv.data(), rl.inv.get(), sizeof(uint32_t) * n);
    copy_to_host(dev, host_in_end.data(), in_end.get(), sizeof(uint32_t) * n);
  }
  rl.oriented_entries = s[0];
  rl.intersect_entries = s[1];
  rl.wedges = s[2];
  rl.dangling = static_cast<uint32_t>(s[3]);
  rl.max_degree = static_cast<uint32_t>(s[4]);
  rl.max_out_degree = static_cast<uint32_t>(s[5]);
  upload_hubs(g, hub_slices(rl.host_offsets, host_in_end));
}

void prepare_host(Graph& g) {
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  const uint64_t m = g.m;
  if (g.host_targets.size() != m) {
    g.host_targets.resize(m);
    if (m > 0) copy_to_host(dev, g.host_targets.data(), g.targets.get(), sizeof(uint32_t) * m);
  }
  const uint32_t* off = g.host_offsets.data();
  const uint32_t* tgt = g.host_targets.data();
  Relabelled& rl = g.rl;

  // Symmetry and self-loops of the original CSR.
  std::atomic<bool> symmetric{true};
  std::atomic<uint32_t> loops{0};
  parallel_for(n, [&](uint64_t b, uint64_t e) {
    uint32_t local_loops = 0;
    bool sym = true;
    for (uint64_t u = b; u < e && sym; ++u) {
      for (uint32_t p = off[u]; p < off[u + 1]; ++p) {
        uint32_t v = tgt[p];
        if (v == u) {
          ++local_loops;
          continue;
        }
        if (!std::binary_search(tgt + off[v], tgt + off[v + 1], static_cast<uint32_t>(u))) {
          sym = false;
          break;
        }
      }
    }
    if (!sym) symmetric = false;
    loops += local_loops;
  }, 1u << 12);
  g.symmetric = symmetric.load();
  g.self_loops = loops.load();

  // Stable counting sort by degree.
  std::vector<uint32_t> deg(n);
  uint32_t max_deg = 0;
  for (uint32_t v = 0; v < n; ++v) {
    deg[v] = off[v + 1] - off[v];
    max_deg = std::max(max_deg, deg[v]);
  }
  std::vector<uint32_t> start(static_cast<size_t>(max_deg) + 2, 0);
  for (uint32_t v = 0; v < n; ++v) ++start[deg[v] + 1];
  for (size_t d = 1; d < start.size(); ++d) start[d] += start[d - 1];
  std::vector<uint32_t> perm(n), inv(n);
  for (uint32_t v = 0; v < n; ++v) perm[start[deg[v]]++] = v;
  parallel_for(n, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) inv[perm[i]] = static_cast<uint32_t>(i);
  });
  std::vector<uint32_t> r_off(static_cast<size_t>(n) + 1);
  r_off[0] = 0;
  for (uint32_t i = 0; i < n; ++i) r_off[i + 1] = r_off[i] + deg[perm[i]];
  std::vector<uint32_t> r_tgt(m), fwd(n);
  parallel_for(n, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t old = perm[i];
      uint32_t* row = r_tgt.data() + r_off[i];
      const uint32_t len = r_off[i + 1] - r_off[i];
      const uint32_t* src = tgt + off[old];
      for (uint32_t j = 0; j < len; ++j) row[j] = inv[src[j]];
      std::sort(row, row + len);
      fwd[i] = r_off[i] + static_cast<uint32_t>(std::upper_bound(row, row + len, static_cast<uint32_t>(i)) - row);
    }
  }, 1u << 12);
  uint64_t oriented = 0;
  uint32_t dangling = 0, max_out = 0;
  std::vector<uint32_t> hubs;  // (vertex, begin, end) slices of long in-lists
  for (uint32_t i = n; i-- > 0;) {
    oriented += r_off[i + 1] - fwd[i];
    dangling += r_off[i + 1] == r_off[i];
    max_out = std::max(max_out, r_off[i + 1] - fwd[i]);
    uint32_t in_end = fwd[i];
    if (in_end > r_off[i] && r_tgt[in_end - 1] == i) --in_end;
    if (in_end - r_off[i] > Relabelled::kHubChunk) {
      for (uint32_t s = r_off[i]; s < in_end; s += Relabelled::kHubChunk) {
        hubs.push_back(i);
        hubs.push_back(s);
        hubs.push_back(std::min(s + Relabelled::kHubChunk, in_end));
      }
    }
  }
  std::atomic<uint64_t> intersect{0}, wedges{0};
  parallel_for(n, [&](uint64_t b, uint64_t e) {
    uint64_t local = 0, local_wedges = 0;
    for (uint64_t i = b; i < e; ++i) {
      const uint64_t du = r_off[i + 1] - fwd[i];
      local_wedges += du * (du ? du - 1 : 0) / 2;
      for (uint32_t p = fwd[i]; p < r_off[i + 1]; ++p) {
        const uint32_t j = r_tgt[p];
        local += du + (r_off[j + 1] - fwd[j]);
      }
    }
    intersect += local;
    wedges += local_wedges;
  }, 1u << 12);

  dev.make_current();
  rl.perm.allocate(sizeof(uint32_t) * std::max<size_t>(n, 1));
  rl.inv.allocate(sizeof(uint32_t) * std::max<size_t>(n, 1));
  rl.offsets.allocate(sizeof(uint32_t) * (static_cast<size_t>(n) + 1));
  rl.targets.allocate(sizeof(uint32_t) * std::max<uint64_t>(m, 1));
  rl.fwd.allocate(sizeof(uint32_t) * std::max<size_t>(n, 1));
  copy_to_device(dev, rl.perm.get(), perm.data(), sizeof(uint32_t) * n);
  copy_to_device(dev, rl.inv.get(), inv.data(), sizeof(uint32_t) * n);
  copy_to_device(dev, rl.offsets.get(), r_off.data(), sizeof(uint32_t) * (static_cast<size_t>(n) + 1));
  copy_to_device(dev, rl.targets.get(), r_tgt.data(), sizeof(uint32_t) * m);
  copy_to_device(dev, rl.fwd.get(), fwd.data(), sizeof(uint32_t) * n);
  rl.host_inv = std::move(inv);
  rl.host_offsets = std::move(r_off);
  rl.oriented_entries = oriented;
  rl.intersect_entries = intersect.load();
  rl.wedges = wedges.load();
  rl.dangling = dangling;
  rl.max_degree = max_deg;
  rl.max_out_degree = max_out;
  upload_hubs(g, hubs);
  // The host copy of the original targets is no longer needed.
  std::vector<uint32_t>().swap(g.host_targets);
}

}  // namespace

void prepare(Graph& g) {
  if (g.prepared) return;
  Stopwatch sw;
  const char* env = std::getenv("GC_PREPARE");
  if (env != nullptr && std::strcmp(env, "host") == 0) prepare_host(g);
  else prepare_device(g);
  g.prepared = true;
  g.prepare_seconds = sw.seconds();
}

void ensure_fwd_original(Graph& g) {
  if (g.fwd_original_ready) return;
  Device& dev = *g.dev;
  g.fwd_original.allocate(sizeof(uint32_t) * std::max<size_t>(g.n, 1));
  if (g.n > 0) {
    Kernel k = get_kernel(dev, "util", "csr_forward_start");
    Grid1D grid = thread_per_item(g.n);
    launch(dev, k, grid.blocks, grid.threads, 0, g.offsets.as<const uint32_t>(),
           g.targets.as<const uint32_t>(), g.n, g.fwd_original.as<uint32_t>());
  }
  g.fwd_original_ready = true;
}

static Graph* as_graph(gc_graph* h) {
  if (h == nullptr) fail_kernel("graph is null");
  return reinterpret_cast<Graph*>(h);
}

static void upload(gc_device* dev_handle, const uint32_t* offsets, size_t num_offsets,
                   const uint32_t* targets, size_t num_targets, gc_graph** out) {
  if (out == nullptr) fail_kernel("out is null");
  *out = nullptr;
  if (dev_handle == nullptr) fail_kernel("device is null");
  if ((offsets == nullptr && num_offsets > 0) || (targets == nullptr && num_targets > 0)) {
    fail_kernel("null array with non-zero length");
  }
  validate_csr(offsets, num_offsets, targets, num_targets);
  auto* dev = reinterpret_cast<Device*>(dev_handle);
  std::lock_guard<std::recursive_mutex> lock(dev->mu);
  dev->make_current();
  std::unique_ptr<Graph> g(new Graph());
  g->n = static_cast<uint32_t>(num_offsets - 1);
  g->m = static_cast<uint32_t>(num_targets);
  g->offsets.allocate(sizeof(uint32_t) * num_offsets);
  g->targets.allocate(sizeof(uint32_t) * std::max<size_t>(num_targets, 1));
  copy_to_device(*dev, g->offsets.get(), offsets, sizeof(uint32_t) * num_offsets);
  copy_to_device(*dev, g->targets.get(), targets, sizeof(uint32_t) * num_targets);
  g->host_offsets.assign(offsets, offsets + num_offsets);
  g->dev = reinterpret_cast<Device*>(gc_device_retain(dev_handle));
  *out = reinterpret_cast<gc_graph*>(g.release());
}

}  // namespace gc

using gc::guard;

extern "C" {

gc_status gc_graph_from_csr(gc_device* dev, const uint32_t* offsets, size_t num_offsets,
                            const uint32_t* targets, size_t num_targets, gc_graph** out) {
  return guard([&] { gc::upload(dev, offsets, num_offsets, targets, num_targets, out); });
}

gc_status gc_graph_from_csr_arrays(gc_device* dev, const gc_csr* csr, gc_graph** out) {
  return guard([&] {
    if (csr == nullptr || csr->offsets == nullptr) gc::fail_kernel("csr is null");
    gc::upload(dev, csr->offsets, size_t(csr->num_nodes) + 1, csr->targets, csr->num_edges, out);
  });
}

void gc_graph_free(gc_graph* h) {
  if (h == nullptr) return;
  auto* g = reinterpret_cast<gc::Graph*>(h);
  gc::Device* dev = g->dev;
  gc::Sched* bfs_sched = g->bfs_sched;
  gc::Sched* tc_sched = g->tc_sched;
  {
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    (void)hipSetDevice(dev->ordinal);
    (void)hipStreamSynchronize(dev->stream);
    delete g;
  }
  if (bfs_sched != nullptr) gc::sched_release(bfs_sched);
  if (tc_sched != nullptr) gc::sched_release(tc_sched);
  gc_device_release(reinterpret_cast<gc_device*>(dev));
}

uint32_t gc_graph_num_nodes(const gc_graph* h) {
  return h ? reinterpret_cast<const gc::Graph*>(h)->n : 0;
}

uint32_t gc_graph_num_edges(const gc_graph* h) {
  return h ? reinterpret_cast<const gc::Graph*>(h)->m : 0;
}

const uint32_t* gc_graph_device_offsets(const gc_graph* h) {
  return h ? reinterpret_cast<const gc::Graph*>(h)->offsets.as<const uint32_t>() : nullptr;
}

gc_status gc_graph_relabelled(gc_graph* h, const uint32_t** offsets, const uint32_t** targets,
                              const uint32_t** perm, const uint32_t** fwd) {
  return guard([&] {
    gc::Graph* g = gc::as_graph(h);
    std::lock_guard<std::recursive_mutex> lock(g->dev->mu);
    g->dev->make_current();
    gc::prepare(*g);
    if (offsets) *offsets = g->rl.offsets.as<const uint32_t>();
    if (targets) *targets = g->rl.targets.as<const uint32_t>();
    if (perm) *perm = g->rl.perm.as<const uint32_t>();
    if (fwd) *fwd = g->rl.fwd.as<const uint32_t>();
  });
}

const uint32_t* gc_graph_device_targets(const gc_graph* h) {
  return h ? reinterpret_cast<const gc::Graph*>(h)->targets.as<const uint32_t>() : nullptr;
}

gc_status gc_graph_prepare(gc_graph* h, double* secs) {
  return guard([&] {
    gc::Graph* g = gc::as_graph(h);
    std::lock_guard<std::recursive_mutex> lock(g->dev->mu);
    g->dev->make_current();
    gc::prepare(*g);
    if (secs) *secs = g->prepare_seconds;
  });
}

gc_status gc_graph_dag_stats(gc_graph* h, uint64_t* oriented_entries, uint64_t* intersect_entries,
                             uint64_t* wedges) {
  return guard([&] {
    gc::Graph* g = gc::as_graph(h);
    std::lock_guard<std::recursive_mutex> lock(g->dev->mu);
    g->dev->make_current();
    gc::prepare(*g);
    if (oriented_entries) *oriented_entries = g->rl.oriented_entries;
    if (intersect_entries) *intersect_entries = g->rl.intersect_entries;
    if (wedges) *wedges = g->rl.wedges;
  });
}

gc_status gc_graph_is_symmetric(gc_graph* h, int32_t* symmetric) {
  return guard([&] {
    gc::Graph* g = gc::as_graph(h);
    std::lock_guard<std::recursive_mutex> lock(g->dev->mu);
    g->dev->make_current();
    gc::prepare(*g);
    if (symmetric) *symmetric = g->symmetric ? 1 : 0;
  });
}

}  // extern "C"

namespace gc {
Graph* graph_from_handle(gc_graph* h) { return as_graph(h); }
}  // namespace gc
