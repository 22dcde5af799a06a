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
// symmetric; other CSRs stay top-down. With a learned scheduler attached
// (sched.cpp) the network picks each level among top-down, bottom-up and the
// device-chained levels (captured HIP graphs of many levels, read back once
// per chain), from features computed on the device after each level.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../runtime/common.hpp"

namespace gc {

namespace {

struct BfsCounters {
  uint32_t next_size;
  uint32_t pad;
  unsigned long long next_edges;
};

struct BfsChainState {  // bfs.hip
  uint32_t size;
  int32_t level;
  uint32_t stop;
  uint32_t levels;
  uint32_t next_size;
  uint32_t cap;
  unsigned long long next_edges;
  unsigned long long edges;
  unsigned long long all_edges;
  unsigned long long reached;
  unsigned long long edge_cap;
  unsigned long long avg_cap;
  unsigned long long unexplored;
  double alpha;
};

struct BfsLevelState {  // mlp.hip
  unsigned long long m, m_f, m_u, reached;
  uint32_t n, n_f, prev_n_f, level, bitmap, max_degree, pad0, pad1;
};

struct SchedDecision {  // mlp.hip
  float features[kSchedFeatures];
  float predicted[kSchedChoices];
  uint32_t choice;
};

// Level counters and the scheduler's decision, read back together.
struct BfsReadback {
  BfsCounters counters[2];
  SchedDecision decision;
};

// BfsPublished of bfs.hip: a level's counters (and decision) as a kernel
// writes them to pinned host memory, `seq` last.
struct BfsPublished {
  BfsCounters counters;
  SchedDecision decision;
  uint32_t seq;
  uint32_t pad[3];
};
static_assert(sizeof(BfsPublished) == 96, "BfsPublished layout");

constexpr double kAlpha = 14.0;
constexpr double kBeta = 24.0;
constexpr uint32_t kTopDown = 0, kBottomUp = 1, kChained = 2;
constexpr uint64_t kSmallFrontier = 16384;  // frontiers scanned by one workgroup
// Device-chained levels (choice 2): the largest frontier they take or keep
// going with, in vertices and in edges, and the levels of one captured chain
// (even, so that the frontier is back in the starting queue after a chain).
constexpr uint32_t kChainCap = 1u << 18;
constexpr uint64_t kChainEdges = 1ull << 20;
constexpr uint32_t kChainRunLevels = 1u << 30;
constexpr uint32_t kChainLevels = 16;
// The hand rule's chains also need a frontier of at most this many edges per
// vertex: the chained expansion walks a list with one lane (up to 32 entries)
// or one wave, so a frontier holding a hub would leave one wave walking it.
constexpr uint64_t kChainAvgDegree = 32;

// The fixed rules (GC_BFS_RULE): Beamer's direction switch alone ("beamer"), or the
// same with top-down levels run as device-chained levels while the frontier fits
// the chains ("chain", the default): at most kChainCap vertices, kChainEdges edges
// and kChainAvgDegree edges per vertex; a chain hands back as soon as one of these
// fails or Beamer's rule would switch to bottom-up.
enum class FixedRule { kBeamer, kBeamerChain };
// Bottom-up levels the fixed rule still runs, launched ahead, after Beamer's
// rule would return to top-down (their frontiers are small and each costs
// less than a top-down level with its read-back; the cap bounds path-like
// graphs, where bottom-up rescans the unvisited vertices at every level).
constexpr uint32_t kTailBottomUp = 3;
// Counter slots per set of the bottom-up runs (BFS_SLOTS of bfs.hip): each
// workgroup adds into one, the publication adds them up.
constexpr uint32_t kBuSlots = 64;
// Publication slots in the pinned area, one per sequence-number parity: a
// level launched ahead publishes into the other slot than the one read.
constexpr size_t kPubStride = 128;

// One traversal from one source on the relabelled graph: the frontier is a
// queue (top-down, chained levels) or a bitmap (bottom-up), converted
// on demand; each level's counters are read back before the next choice.
struct Traversal {
  Graph& g;
  Device& dev;
  uint32_t n;
  uint64_t words64;
  int32_t* dist;
  uint32_t* queue[2];
  uint64_t* bitmap[2];
  BfsReadback* rb;
  uint32_t* start;
  uint32_t* scan_buf;
  BfsChainState* chain_state;
  BfsCounters* bu_counters;  // the three counter sets (kBuSlots slots each) that bottom-up runs rotate through
  BfsCounters* lvl;          // the counters (kBuSlots slots) of the levels run one at a time
  const uint32_t* off;
  const uint32_t* tgt;
  BfsPublished* pub;      // pinned host memory the levels' publications land in (two slots)
  BfsPublished* pub_dev;  // the same, as kernels address it

  // The slot of publication `seq`, as the host reads it and as kernels write it.
  BfsPublished* slot(uint32_t seq) const {
    return reinterpret_cast<BfsPublished*>(reinterpret_cast<char*>(pub) + (seq & 1u) * kPubStride);
  }
  BfsPublished* slot_dev(uint32_t seq) const {
    return reinterpret_cast<BfsPublished*>(reinterpret_cast<char*>(pub_dev) + (seq & 1u) * kPubStride);
  }
  Kernel k_td, k_deg, k_scan_small, k_bu, k_q2b, k_b2q, k_chain_expand, k_chain_advance, k_decide, k_publish;
  uint32_t cu;
  // state at the start of the next level
  uint64_t n_f = 1, m_f = 0, m_u = 0, reached = 1, prev_n_f = 1;
  int32_t level = 0;
  bool in_bitmap = false;
  int cur = 0;
  bool done = false;
  bool top_down_dir = true;  // the fixed rule's current direction
  double alpha = kAlpha;     // the fixed rule's top-down -> bottom-up threshold (GC_BFS_ALPHA)
  FixedRule rule = FixedRule::kBeamerChain;
  // In queue form, whether bitmap[cur] holds exactly the queue's vertices and
  // bitmap[cur ^ 1] is empty (top-down levels then keep it so, and a switch to
  // bottom-up needs no conversion). Kept for symmetric graphs only.
  bool synced = false;
  bool first_level = true;   // start[0] = 0 from bfs_init: the first level needs no scan
  uint32_t bu_runs = 0;      // bottom-up runs started (bfs_init clears the counter sets for the first)
  uint32_t tail_levels = kTailBottomUp;  // GC_BFS_TAIL

  Traversal(Graph& graph, uint32_t src)
      : g(graph), dev(*graph.dev), n(graph.n), words64((static_cast<uint64_t>(graph.n) + 63) / 64) {
    dist = g.buffer<int32_t>("bfs.dist", sizeof(int32_t) * n);
    queue[0] = g.buffer<uint32_t>("bfs.q0", sizeof(uint32_t) * n);
    queue[1] = g.buffer<uint32_t>("bfs.q1", sizeof(uint32_t) * n);
    bitmap[0] = g.buffer<uint64_t>("bfs.b0", sizeof(uint64_t) * words64);
    bitmap[1] = g.buffer<uint64_t>("bfs.b1", sizeof(uint64_t) * words64);
    rb = g.buffer<BfsReadback>("bfs.readback", sizeof(BfsReadback));
    start = g.buffer<uint32_t>("bfs.start", sizeof(uint32_t) * n);
    scan_buf = g.buffer<uint32_t>("bfs.scan", sizeof(uint32_t) * scan_scratch_words(n));
    chain_state = g.buffer<BfsChainState>("bfs.chain", sizeof(BfsChainState));
    bu_counters = g.buffer<BfsCounters>("bfs.bu_counters", 3 * kBuSlots * sizeof(BfsCounters));
    lvl = g.buffer<BfsCounters>("bfs.lvl_counters", kBuSlots * sizeof(BfsCounters));
    off = g.rl.offsets.as<const uint32_t>();
    tgt = g.rl.targets.as<const uint32_t>();
    k_td = get_kernel(dev, "bfs", "bfs_topdown_edges");
    k_deg = get_kernel(dev, "bfs", "bfs_frontier_degrees");
    k_scan_small = get_kernel(dev, "bfs", "bfs_frontier_scan_small");
    k_bu = get_kernel(dev, "bfs", "bfs_bottomup");
    k_q2b = get_kernel(dev, "bfs", "bfs_queue_to_bitmap");
    k_b2q = get_kernel(dev, "bfs", "bfs_bitmap_to_queue");
    k_chain_expand = get_kernel(dev, "bfs", "bfs_chain_expand");
    k_chain_advance = get_kernel(dev, "bfs", "bfs_chain_advance");
    k_decide = get_kernel(dev, "mlp", "sched_bfs_decide");
    k_publish = get_kernel(dev, "bfs", "bfs_publish");
    static_assert(sizeof(BfsPublished) <= kPubStride, "publication slot size");
    static_assert(Device::kPinnedKernelOffset + 2 * kPubStride <= Device::kPinnedBytes, "pinned area size");
    pub = reinterpret_cast<BfsPublished*>(static_cast<char*>(dev.pinned) + Device::kPinnedKernelOffset);
    pub_dev = reinterpret_cast<BfsPublished*>(static_cast<char*>(dev.pinned_device) + Device::kPinnedKernelOffset);
    cu = static_cast<uint32_t>(std::max(1, dev.cu_count));
    const std::vector<uint32_t>& hoff = g.rl.host_offsets;
    // four entries per lane and store, a grid-stride loop over at most 16 workgroups per CU
    Grid1D vgrid = thread_per_item(std::max<uint64_t>(n / 4, 1));
    launch(dev, get_kernel(dev, "bfs", "bfs_init"), std::min(vgrid.blocks, cu * 16), vgrid.threads, 0, dist, n,
           src, queue[0], lvl, bu_counters, start, bitmap[0], bitmap[1], words64);
    synced = g.symmetric;
    m_f = hoff[src + 1] - hoff[src];
    if (const char* a = std::getenv("GC_BFS_ALPHA")) alpha = std::max(1.0, std::atof(a));
    if (const char* a = std::getenv("GC_BFS_TAIL")) tail_levels = static_cast<uint32_t>(std::strtoul(a, nullptr, 10));
    if (const char* r = std::getenv("GC_BFS_RULE")) {
      if (std::strcmp(r, "beamer") == 0) rule = FixedRule::kBeamer;
      else if (std::strcmp(r, "chain") == 0) rule = FixedRule::kBeamerChain;
      else fail_kernel(std::string("GC_BFS_RULE must be beamer or chain, not ") + r);
    }
    m_u = static_cast<uint64_t>(g.m) - m_f;
  }

  uint32_t allowed_mask() const {
    uint32_t mask = 1u << kTopDown;
    if (g.symmetric) mask |= 1u << kBottomUp;
    if (n_f <= kChainCap && m_f <= kChainEdges) mask |= 1u << kChained;
    return mask;
  }

  BfsLevelState level_state() const {
    BfsLevelState st{};
    st.m = g.m;
    st.m_f = m_f;
    st.m_u = m_u;
    st.reached = reached;
    st.n = n;
    st.n_f = static_cast<uint32_t>(n_f);
    st.prev_n_f = static_cast<uint32_t>(prev_n_f);
    st.level = static_cast<uint32_t>(level);
    st.bitmap = in_bitmap ? 1u : 0u;
    st.max_degree = g.rl.max_degree;
    return st;
  }

  // The fixed rule: Beamer's direction switch; under the chain rule a
  // top-down level whose frontier fits the chains runs as chained levels.
  uint32_t fixed_choice() {
    if (top_down_dir && g.symmetric && static_cast<double>(m_f) > static_cast<double>(m_u) / alpha) {
      top_down_dir = false;
    } else if (!top_down_dir && static_cast<double>(n_f) < static_cast<double>(n) / kBeta) {
      top_down_dir = true;
    }
    if (!top_down_dir) return kBottomUp;
    if (rule == FixedRule::kBeamerChain && chain_fits()) return kChained;
    return kTopDown;
  }

  // Whether the current frontier fits the hand rule's chains. The first level
  // (the source alone) always runs top-down: one vertex's degree says nothing
  // of the traversal's depth, and on a power-law graph a low-degree source's
  // next frontier holds hubs (measured: chaining it cost 5-9% of an R-MAT BFS).
  bool chain_fits() const {
    return level >= 1 && n_f <= kChainCap && m_f <= kChainEdges && m_f <= kChainAvgDegree * n_f;
  }

  // Launches the decision kernel for the current state (or, with
  // from_counters, for the state after the level whose counters are in rb;
  // with `publish`, it also publishes them and its decision to `pub` under
  // sequence number `seq`).
  void launch_decide(const float* params, bool from_counters, bool publish = false, uint32_t seq = 0) {
    const SchedNorm norm = g.bfs_sched != nullptr ? g.bfs_sched->norm : SchedNorm{};
    // the kernel drops the chained levels itself when the frontier is too large
    const uint32_t mask = (1u << kTopDown) | (g.symmetric ? (1u << kBottomUp) : 0u) | (1u << kChained);
    launch(dev, k_decide, 1u, 64u, 0, params, norm, level_state(), lvl, kBuSlots, from_counters ? 1u : 0u,
           mask, kChainCap, static_cast<unsigned long long>(kChainEdges), &rb->decision,
           publish ? slot_dev(seq) : static_cast<BfsPublished*>(nullptr), seq);
  }

  // Waits until a kernel has published sequence number `seq` to `pub`: the
  // host polls the sequence word instead of waiting for the stream (about 5
  // us less per level), and checks the stream every few thousand polls so
  // that a failed launch is reported instead of waited for.
  void wait_published(uint32_t seq) {
    BfsPublished* p = slot(seq);
    for (uint32_t spins = 1;; ++spins) {
      if (__atomic_load_n(&p->seq, __ATOMIC_ACQUIRE) == seq) return;
      if ((spins & 4095u) == 0) {
        const hipError_t err = hipStreamQuery(dev.stream);
        if (err == hipSuccess) {
          // the stream is idle, so the publishing kernel has ended
          if (__atomic_load_n(&p->seq, __ATOMIC_ACQUIRE) == seq) return;
          fail_kernel("BFS level counters were not published");
        }
        if (err != hipErrorNotReady) hip_check(err, "BFS level");
      }
    }
  }

  void to_queue() {
    if (!in_bitmap) return;
    fill_words(dev, &rb->counters[1], sizeof(BfsCounters) / 4, 0);
    // the other bitmap (an earlier frontier) is cleared, so the queue and its
    // bitmap are in step for the top-down levels that follow
    fill_words(dev, bitmap[cur ^ 1], 2 * words64, 0);
    Grid1D bg = wave_per_item((words64 + 63) / 64);
    launch(dev, k_b2q, std::min(bg.blocks, cu * 16), bg.threads, 0,
           static_cast<const uint64_t*>(bitmap[cur]), words64, queue[cur], &rb->counters[1].next_size);
    in_bitmap = false;
    synced = g.symmetric;
  }

  // The transposed four last entries of every row (built once per graph).
  const uint32_t* top_entries() {
    if (!g.bfs_top_ready) {
      g.bfs_top.allocate(sizeof(uint32_t) * 4 * std::max<uint64_t>(n, 1));
      Grid1D vg = thread_per_item(n);
      launch(dev, get_kernel(dev, "bfs", "bfs_top_build"), vg.blocks, vg.threads, 0, off, tgt, n,
             g.bfs_top.as<uint32_t>());
      g.bfs_top_ready = true;
    }
    return g.bfs_top.as<const uint32_t>();
  }

  void to_bitmap() {
    if (in_bitmap) return;
    if (synced) {
      in_bitmap = true;  // the top-down levels kept bitmap[cur] in step with the queue
      return;
    }
    fill_words(dev, bitmap[cur], 2 * words64, 0);
    Grid1D qg = thread_per_item(n_f);
    launch(dev, k_q2b, qg.blocks, qg.threads, 0, static_cast<const uint32_t*>(queue[cur]),
           static_cast<uint32_t>(n_f), reinterpret_cast<uint32_t*>(bitmap[cur]));
    in_bitmap = true;
  }

  // Launches one top-down or bottom-up level (counters into lvl).
  void launch_level(uint32_t choice) {
    const int32_t next_level = level + 1;
    const bool first = first_level;
    first_level = false;
    if (choice == kTopDown) {
      to_queue();
      // (the level counters are clear: bfs_init and every publication clear them)
      if (first && n_f == 1) {
        // the source alone: bfs_init wrote its degree prefix start[0] = 0
      } else if (n_f <= kSmallFrontier) {
        launch(dev, k_scan_small, 1u, 1024u, 0, off, static_cast<const uint32_t*>(queue[cur]),
               static_cast<uint32_t>(n_f), start, lvl);
      } else {
        Grid1D qg = thread_per_item(n_f);
        launch(dev, k_deg, qg.blocks, qg.threads, 0, off, static_cast<const uint32_t*>(queue[cur]),
               static_cast<uint32_t>(n_f), start);
        exclusive_scan_u32(dev, start, n_f, scan_buf);
      }
      Grid1D tg = wave_per_item((m_f + 255) / 256);
      uint32_t* cur_bits = synced ? reinterpret_cast<uint32_t*>(bitmap[cur]) : nullptr;
      uint32_t* next_bits = synced ? reinterpret_cast<uint32_t*>(bitmap[cur ^ 1]) : nullptr;
      launch(dev, k_td, std::max(1u, std::min(tg.blocks, cu * 32)), tg.threads, 0, off, tgt, dist,
             static_cast<const uint32_t*>(queue[cur]), static_cast<uint32_t>(n_f),
             static_cast<const uint32_t*>(start), static_cast<uint32_t>(m_f), queue[cur ^ 1],
             lvl, next_level, cur_bits, next_bits);
    } else {
      to_bitmap();
      launch_bottom_up(cur, next_level, lvl, kBuSlots);
    }
  }

  // One bottom-up level from the frontier bitmap[in] into bitmap[in ^ 1]; its
  // grid depends on n only, not on the frontier. Its counts go to `counters`;
  // with `pub_slot`, it first publishes the finished counters `pub` of the
  // level before it under sequence number `seq`.
  BfsCounters* bu_set(int set) { return bu_counters + static_cast<size_t>(set) * kBuSlots; }

  void launch_bottom_up(int in, int32_t next_level, BfsCounters* counters, uint32_t slots,
                        BfsCounters* pub = nullptr, BfsPublished* pub_slot = nullptr, uint32_t seq = 0,
                        BfsCounters* clear = nullptr) {
    Grid1D bg = wave_per_item(words64);
    // the rows of degree 0 come first in the degree order: skip their chunks
    const uint32_t first_chunk = g.rl.dangling / 64;
    launch(dev, k_bu, std::min(bg.blocks, cu * 8), bg.threads, 0, off, tgt, top_entries(), dist,
           reinterpret_cast<const uint32_t*>(bitmap[in]), bitmap[in ^ 1], n, first_chunk, counters,
           next_level, pub, pub_slot, seq, clear, slots);
    synced = false;  // the queues no longer hold the frontier
  }

  // Publishes (and clears) the level counters under a new sequence number.
  uint32_t publish_counters() {
    const uint32_t seq = ++dev.publish_seq;
    launch(dev, k_publish, 1u, 64u, 0, lvl, kBuSlots, slot_dev(seq), seq);
    return seq;
  }

  // Bottom-up levels under the fixed rule, each launched before the counters
  // of the level before it are read, so the read-back overlaps a level on the
  // GPU instead of leaving it idle. A level launched ahead is a complete BFS
  // level whatever those counters say: if the traversal has ended, it finds
  // an empty frontier and changes nothing. Continues while Beamer's rule
  // stays bottom-up and then for at most kTailBottomUp levels; returns when
  // the traversal ends or with the next level to run top-down. `on_level` is
  // called after each level's counters are applied.
  // The levels rotate through the three bu_counters sets; each level
  // launched ahead counts into the next set, publishes the set of the level
  // before it (so no publishing launch sits between two levels) and clears
  // the set after that, and the run's last level is published by
  // bfs_publish. A level launched ahead after the traversal has ended sees
  // the empty count of the level before it and returns at once. bfs_init
  // clears the sets for the first run; a later run clears them first.
  template <class OnLevel>
  void run_bottom_up(OnLevel&& on_level) {
    to_bitmap();
    if (bu_runs++ > 0) fill_words(dev, bu_counters, 3 * kBuSlots * sizeof(BfsCounters) / 4, 0);
    int in = cur;
    int32_t lvl = level;  // the level whose frontier is bitmap[in]
    int set = 0;          // the counter set of the oldest level not yet read
    launch_bottom_up(in, lvl + 1, bu_set(set), kBuSlots);
    in ^= 1;
    ++lvl;
    uint32_t seq = ++dev.publish_seq;  // the publication of that level's set
    uint32_t tail = 0;
    bool ahead = true;
    for (;;) {
      if (ahead) {
        launch_bottom_up(in, lvl + 1, bu_set((set + 1) % 3), kBuSlots, bu_set(set), slot_dev(seq), seq,
                         bu_set((set + 2) % 3));
        in ^= 1;
        ++lvl;
      } else {
        launch(dev, k_publish, 1u, 64u, 0, bu_set(set), kBuSlots, slot_dev(seq), seq);
      }
      wait_published(seq);
      BfsCounters counters;
      std::memcpy(&counters, &slot(seq)->counters, sizeof(counters));
      finish_level(kBottomUp, counters);
      on_level();
      if (done || !ahead) break;
      // the level after the one just launched: bottom-up while the rule says so, then the tail
      if (static_cast<double>(n_f) < static_cast<double>(n) / kBeta) ++tail;
      ahead = tail < tail_levels;
      set = (set + 1) % 3;
      seq = ++dev.publish_seq;
    }
    top_down_dir = true;
  }

  // Applies the counters of the level just run (already copied to `c`).
  void finish_level(uint32_t choice, const BfsCounters& c) {
    if (c.next_size == 0) {
      done = true;
      return;
    }
    prev_n_f = n_f;
    level += 1;
    reached += c.next_size;
    m_u -= std::min<uint64_t>(m_u, c.next_edges);
    m_f = c.next_edges;
    n_f = c.next_size;
    in_bitmap = choice == kBottomUp;
    cur ^= 1;
  }

  uint32_t chain_grid() const { return cu * 8; }

  // The captured chain of kChainLevels levels starting from queue[start].
  hipGraphExec_t chain_exec(int start) {
    GraphExecs& ge = g.bfs_chain;
    if (ge.exec[start] != nullptr) return ge.exec[start];
    dev.sync();
    hip_check(hipStreamBeginCapture(dev.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipError_t err = hipSuccess;
    for (uint32_t k = 0; k < kChainLevels && err == hipSuccess; ++k) {
      const uint32_t* in = queue[start ^ static_cast<int>(k & 1)];
      uint32_t* out = queue[start ^ static_cast<int>((k + 1) & 1)];
      void* expand_params[] = {karg(off), karg(tgt), karg(dist), karg(in), karg(out), karg(chain_state)};
      err = hipModuleLaunchKernel(k_chain_expand.fn, chain_grid(), 1, 1, 256, 1, 1, 0, dev.stream,
                                  expand_params, nullptr);
      if (err == hipSuccess) {
        void* advance_params[] = {karg(chain_state)};
        err = hipModuleLaunchKernel(k_chain_advance.fn, 1, 1, 1, 64, 1, 1, 0, dev.stream, advance_params,
                                    nullptr);
      }
    }
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(dev.stream, &graph);
    hip_check(err, "BFS level chain capture");
    hip_check(end, "hipStreamEndCapture");
    ge.graph[start] = graph;
    hip_check(hipGraphInstantiate(&ge.exec[start], graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    return ge.exec[start];
  }

  // Device-side levels, captured chains of kChainLevels levels replayed with
  // one read-back each, until the traversal ends, hands back a large frontier
  // or has built max_levels levels (checked after each chain); with
  // max_levels == 1, one level launched directly. Returns the levels built.
  // With `fixed_rule`, the chain also hands back when the hand rule would stop
  // chaining (a frontier of more than kChainAvgDegree edges per vertex) or
  // switch to bottom-up (Beamer's test on the device).
  uint32_t run_chain(uint32_t max_levels, bool fixed_rule = false) {
    to_queue();
    synced = false;  // the chained levels keep no bitmap
    first_level = false;
    BfsChainState init{};
    init.size = static_cast<uint32_t>(n_f);
    init.level = level;
    init.cap = kChainCap;
    init.edges = m_f;
    init.edge_cap = kChainEdges;
    if (fixed_rule) {
      init.avg_cap = kChainAvgDegree;
      init.unexplored = m_u;
      init.alpha = g.symmetric ? alpha : 0.0;
    }
    copy_to_device(dev, chain_state, &init, sizeof(init));
    BfsChainState st{};
    // The hand rule's first level runs as one direct launch: when the chain
    // hands back after it (a hub next to a low-degree source), no replay of
    // kChainLevels levels pays for the levels that return at once.
    bool direct = max_levels == 1 || fixed_rule;
    for (;;) {
      const int in = cur ^ static_cast<int>(st.levels & 1u);  // the queue holding the frontier
      if (direct) {
        launch(dev, k_chain_expand, chain_grid(), 256u, 0, off, tgt, dist,
               static_cast<const uint32_t*>(queue[in]), queue[in ^ 1], chain_state);
        launch(dev, k_chain_advance, 1u, 64u, 0, chain_state);
      } else {
        hipEvent_t prof = nullptr;
        profile_begin(dev, "bfs_chain", &prof);
        hip_check(hipGraphLaunch(chain_exec(in), dev.stream), "hipGraphLaunch");
        profile_end(dev, "bfs_chain", prof);
      }
      copy_to_host(dev, dev.pinned, chain_state, sizeof(st));
      std::memcpy(&st, dev.pinned, sizeof(st));
      if (st.stop != 0 || max_levels == 1 || st.levels >= max_levels) break;
      direct = false;
      // a full chain has an even number of levels, so `in` changes only after a direct level
    }
    reached += st.reached;
    m_u -= std::min<uint64_t>(m_u, st.all_edges);
    if (st.stop == 1) {
      level = st.level - 1;
      done = true;
      return st.levels;
    }
    level = st.level;
    prev_n_f = n_f;
    n_f = st.size;
    m_f = st.edges;
    if (st.levels & 1u) cur ^= 1;
    in_bitmap = false;
    return st.levels;
  }

  // Runs one decision's worth of work: a level, or a run of chained levels.
  // With `decide_next`, the scheduler's decision for the next level is
  // computed on the device and read back with the counters.
  void step(uint32_t choice, uint32_t chain_levels, bool decide_next, const float* params) {
    if (choice == kChained) {
      run_chain(chain_levels, params == nullptr);
      return;
    }
    launch_level(choice);
    uint32_t seq;
    if (decide_next) {
      seq = ++dev.publish_seq;
      launch_decide(params, true, true, seq);
    } else {
      seq = publish_counters();
    }
    wait_published(seq);
    BfsCounters counters;
    std::memcpy(&counters, &slot(seq)->counters, sizeof(counters));
    if (decide_next) std::memcpy(&last_decision, &slot(seq)->decision, sizeof(last_decision));
    finish_level(choice, counters);
  }

  // The scheduler's decision for the current state (one round trip).
  uint32_t decide_now(const float* params) {
    launch_decide(params, false);
    copy_to_host(dev, dev.pinned, &rb->decision, sizeof(SchedDecision));
    std::memcpy(&last_decision, dev.pinned, sizeof(SchedDecision));
    return last_decision.choice;
  }

  SchedDecision last_decision{};
};

}  // namespace

struct BfsOutcome {
  uint32_t reachable = 0;
  int32_t max_depth = 0;
};

// The fixed rule as the learned rule's fallback: the network's choice stands
// only where it predicts it at least `margin` (in ln microseconds; 0.15, about
// 16%, GC_SCHED_MARGIN) cheaper than the fixed rule's choice for the same
// state, and the fixed rule decides elsewhere. 0.15 was the best of 0, 0.05,
// 0.1, 0.15 and 0.26 on R-MAT scale 22 with seed 1, a graph outside the
// training set (every margin there stayed 5-8% slower than the fixed rule).
static float learned_margin() {
  const char* env = std::getenv("GC_SCHED_MARGIN");
  return env != nullptr ? static_cast<float>(std::atof(env)) : 0.15f;
}

static uint32_t with_fallback(Traversal& t, uint32_t learned, float margin) {
  const uint32_t fixed = t.fixed_choice();  // also advances the rule's direction state
  if (learned == fixed) return learned;
  const float* p = t.last_decision.predicted;
  return p[learned] + margin < p[fixed] ? learned : fixed;
}

// Runs BFS from `source` (caller's id) and writes caller-ordered distances to
// the device array `d_out`. The scheduler attached to the graph, if any,
// chooses each level's method; otherwise the fixed rule does.
BfsOutcome bfs(Graph& g, uint32_t source, int32_t* d_out) {
  prepare(g);
  Device& dev = *g.dev;
  Traversal t(g, g.rl.host_inv[source]);
  const char* trace_env = std::getenv("GC_BFS_TRACE");
  const bool trace = trace_env != nullptr && trace_env[0] == '1';
  const float* params = g.bfs_sched != nullptr ? g.bfs_sched->params.as<const float>() : nullptr;
  g.last_schedule.clear();
  // The first level (the source alone) runs top-down under either rule; the
  // network decides from the next level on, its decision computed on the
  // device after each level and read back with the level's counters (one
  // round trip per level, as for the fixed rule).
  uint32_t choice = params != nullptr ? kTopDown : t.fixed_choice();
  // Bottom-up levels are launched ahead of the read-back (GC_BFS_AHEAD=0: one
  // level per round trip). Under the learned rule, a bottom-up choice starts
  // such a run, which ends as under the fixed rule; the network then decides
  // the next level from the state after it (one round trip).
  const char* ahead_env = std::getenv("GC_BFS_AHEAD");
  const bool ahead = ahead_env == nullptr || ahead_env[0] != '0';
  const float margin = learned_margin();
  Stopwatch level_clock;
  while (!t.done) {
    if (!(t.allowed_mask() >> choice & 1u)) choice = kTopDown;
    if (ahead && choice == kBottomUp) {
      t.run_bottom_up([&] {
        g.last_schedule.push_back(kBottomUp);
        if (trace) {
          if (t.done) {
            std::fprintf(stderr, "bfs level %d bottom-up (launched ahead) -> done (%.3f ms)\n", t.level,
                         level_clock.seconds() * 1e3);
          } else {
            std::fprintf(stderr, "bfs level %d bottom-up (launched ahead) -> level %d frontier %llu (%.3f ms)\n",
                         t.level - 1, t.level, static_cast<unsigned long long>(t.n_f), level_clock.seconds() * 1e3);
          }
          level_clock = Stopwatch();
        }
      });
      if (t.done) break;
      choice = params != nullptr ? with_fallback(t, t.decide_now(params), margin) : kTopDown;
      continue;
    }
    g.last_schedule.push_back(choice);
    const uint64_t n_f = t.n_f;
    const int32_t level = t.level;
    const bool learned_next = params != nullptr && choice != kChained;
    t.step(choice, kChainRunLevels, learned_next, params);
    if (trace) {
      std::fprintf(stderr, "bfs level %d %s frontier %llu -> level %d frontier %llu (%.3f ms)\n", level,
                   choice == kTopDown ? "top-down " : (choice == kBottomUp ? "bottom-up" : "chained  "),
                   static_cast<unsigned long long>(n_f), t.level, static_cast<unsigned long long>(t.n_f),
                   level_clock.seconds() * 1e3);
      level_clock = Stopwatch();
    }
    if (t.done) break;
    if (params == nullptr) choice = t.fixed_choice();
    else if (learned_next) choice = with_fallback(t, t.last_decision.choice, margin);
    else choice = with_fallback(t, t.decide_now(params), margin);
  }
  Grid1D vgrid = thread_per_item((static_cast<uint64_t>(g.n) + 3) / 4);  // four entries per lane
  launch(dev, get_kernel(dev, "util", "gather_u32"), vgrid.blocks, vgrid.threads, 0,
         reinterpret_cast<const uint32_t*>(t.dist), g.rl.inv.as<const uint32_t>(),
         reinterpret_cast<uint32_t*>(d_out), static_cast<uint64_t>(g.n));
  dev.sync();
  BfsOutcome outcome;
  outcome.reachable = static_cast<uint32_t>(t.reached);
  outcome.max_depth = t.level;
  return outcome;
}

// Training data: at every level, the features of the decision and the time of
// each allowed choice from the same state (dist and frontier saved and
// restored around each trial); then the fastest choice is run for one level.
// Each choice is timed the way bfs() runs it when chosen: chained levels as
// one captured chain divided by the levels it built; bottom-up as the run of
// levels launched ahead that the choice starts (unless GC_BFS_AHEAD=0), up to
// the publication of its first level: that level is what the choice adds, as
// the run's later levels are those the fixed rule takes too. (The run's mean
// per level would hide a first level from a small frontier, which scans
// nearly every list: 0.5 ms at R-MAT scale 24 against a 0.12 ms mean.)
uint32_t bfs_level_timings(Graph& g, uint32_t source, uint32_t max_levels, float* features,
                           float* times_us) {
  prepare(g);
  Device& dev = *g.dev;
  Traversal t(g, g.rl.host_inv[source]);
  const char* ahead_env = std::getenv("GC_BFS_AHEAD");
  const bool ahead = ahead_env == nullptr || ahead_env[0] != '0';
  int32_t* save_dist = g.buffer<int32_t>("bfs.save_dist", sizeof(int32_t) * g.n);
  uint64_t* save_front = g.buffer<uint64_t>("bfs.save_front", std::max<uint64_t>(sizeof(uint32_t) * g.n, 16));
  uint32_t levels = 0;
  while (!t.done && levels < max_levels) {
    t.decide_now(nullptr);
    std::memcpy(features + static_cast<uint64_t>(levels) * kSchedFeatures, t.last_decision.features,
                sizeof(float) * kSchedFeatures);
    const Traversal saved = t;
    const size_t front_bytes = saved.in_bitmap ? sizeof(uint64_t) * t.words64 : sizeof(uint32_t) * t.n_f;
    const void* front_src = saved.in_bitmap ? static_cast<const void*>(t.bitmap[t.cur])
                                            : static_cast<const void*>(t.queue[t.cur]);
    hip_check(hipMemcpyAsync(save_dist, t.dist, sizeof(int32_t) * g.n, hipMemcpyDeviceToDevice, dev.stream),
              "hipMemcpyAsync");
    if (front_bytes > 0)
      hip_check(hipMemcpyAsync(save_front, front_src, front_bytes, hipMemcpyDeviceToDevice, dev.stream),
                "hipMemcpyAsync");
    auto restore = [&]() {
      hip_check(hipMemcpyAsync(t.dist, save_dist, sizeof(int32_t) * g.n, hipMemcpyDeviceToDevice, dev.stream),
                "hipMemcpyAsync");
      if (front_bytes > 0)
        hip_check(hipMemcpyAsync(const_cast<void*>(front_src), save_front, front_bytes,
                                 hipMemcpyDeviceToDevice, dev.stream),
                  "hipMemcpyAsync");
      t.n_f = saved.n_f;
      t.m_f = saved.m_f;
      t.m_u = saved.m_u;
      t.reached = saved.reached;
      t.prev_n_f = saved.prev_n_f;
      t.level = saved.level;
      t.in_bitmap = saved.in_bitmap;
      t.cur = saved.cur;
      t.done = false;
      t.synced = false;  // trials change the bitmaps; conversions rebuild them
      t.first_level = false;
      t.top_down_dir = saved.top_down_dir;
      dev.sync();
    };
    const uint32_t allowed = t.allowed_mask();
    float best_time = 0.0f;
    uint32_t best = kTopDown;
    float* row = times_us + static_cast<uint64_t>(levels) * kSchedChoices;
    for (uint32_t c = 0; c < kSchedChoices; ++c) {
      row[c] = -1.0f;
      if (!(allowed >> c & 1u)) continue;
      float fastest = 0.0f;
      for (int rep = 0; rep < 2; ++rep) {
        restore();
        Stopwatch sw;
        float us;
        if (c == kChained) {
          // chained levels cost what one captured chain costs per level it builds
          const uint32_t built = t.run_chain(kChainLevels);
          us = static_cast<float>(sw.seconds() * 1e6) / static_cast<float>(std::max(1u, built));
        } else if (c == kBottomUp && ahead) {
          float first = -1.0f;
          t.run_bottom_up([&] {
            if (first < 0.0f) first = static_cast<float>(sw.seconds() * 1e6);
          });
          us = first >= 0.0f ? first : static_cast<float>(sw.seconds() * 1e6);
        } else {
          t.step(c, 1, false, nullptr);
          us = static_cast<float>(sw.seconds() * 1e6);
        }
        fastest = rep == 0 ? us : std::min(fastest, us);
      }
      row[c] = fastest;
      if (best_time == 0.0f || fastest < best_time) best_time = fastest, best = c;
    }
    restore();
    t.step(best, 1, false, nullptr);
    ++levels;
  }
  dev.sync();
  return levels;
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

gc_status gc_bfs_level_timings(gc_graph* h, uint32_t source, uint32_t max_levels, float* features,
                               float* times_us, uint32_t* levels) {
  return guard([&] {
    gc::GraphCall call(h);
    gc::check_source(*call.g, source);
    if ((features == nullptr || times_us == nullptr) && max_levels > 0)
      gc::fail_kernel("features or times_us is null");
    const uint32_t got = gc::bfs_level_timings(*call.g, source, max_levels, features, times_us);
    if (levels) *levels = got;
  });
}

}  // extern "C"
