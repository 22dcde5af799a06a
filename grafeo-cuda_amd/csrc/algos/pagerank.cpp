// PageRank driver.
//
// Replaces pagerank::pagerank (crates/grafeo-cuda-algos/src/pagerank.rs:46-199):
// r0 = 1/n; per iteration the dangling rank D, the pull update, the L1 change;
// `iterations` counts performed iterations; converged once the change drops
// below `tolerance`; the latest iterate is returned; n == 0 gives
// {[], 0, converged}. Differences: FP64 accumulation and reductions (the
// reference sums in FP32 with FP32 atomics), and the convergence test runs on
// the device, the host reading the state once per batch of iterations.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <stdexcept>

#include "../runtime/common.hpp"
#include "pr_bin_layout.hpp"
#include "pr_slice_layout.hpp"

namespace gc {

namespace {

constexpr uint32_t kPrThreads = 256;     // pr_spmv workgroup
constexpr uint32_t kPrHotThreads = 1024; // pr_spmv_hot workgroup
constexpr uint32_t kPrMaxSegments = 24;
constexpr uint32_t kPrBatch = 8;  // iterations launched between state reads

// Iterations to launch before the next state read. Until two reads are in, a
// whole batch; then the L1 change's decay per iteration, measured between the
// last two reads, predicts when it falls below the tolerance, and the batch
// stops there (at least one iteration, at most kPrBatch), so that few of the
// launches after convergence (which return at once) are made. The read
// schedule changes nothing computed: only how many empty launches follow.
uint32_t pr_next_batch(double prev_diff, uint32_t prev_it, double diff, uint32_t it, double tol) {
  if (prev_it == 0 || it <= prev_it || !(diff > 0.0) || !(prev_diff > diff) || !(tol > 0.0)) return kPrBatch;
  const double rate = std::pow(diff / prev_diff, 1.0 / static_cast<double>(it - prev_it));  // in (0, 1)
  if (!(rate > 0.0) || !(rate < 1.0)) return kPrBatch;
  if (diff < tol) return 1;
  const double need = std::ceil(std::log(tol / diff) / std::log(rate));
  return static_cast<uint32_t>(std::min<double>(kPrBatch, std::max(1.0, need)));
}
constexpr uint32_t kPrExact = 0x80000000u;  // group flag: rows of one exact degree
constexpr uint32_t kPrExactMax = 16;        // exact-degree segments for degrees 0..15
constexpr uint32_t kPrHubLimit = 1024;      // rows of at least this degree are summed in chunks
constexpr uint32_t kPrHubChunk = 4096;      // PR_HUB_CHUNK of pagerank.hip
constexpr uint32_t kPrBlocksPerCu = 8;      // resident 256-lane workgroups per CU (32 waves)

struct PrSegment {
  uint32_t row_begin;
  uint32_t row_end;
  uint32_t group;
  uint32_t block_begin;
  uint32_t entry_begin;
};

struct PrSegments {
  PrSegment seg[kPrMaxSegments];
  uint32_t count;
  uint32_t total_blocks;
};

struct PrState {
  double dangling;
  double diff;
  uint32_t iterations;
  uint32_t done;
};

// Rows per thread of an exact-degree segment (pr_rows_exact's R for degree d).
uint32_t exact_rows_per_thread(uint32_t d) { return d <= 3 ? 4 : (d <= 7 ? 2 : 1); }

// Cuts the degree-sorted rows into segments of one lane-group size each; a
// segment's rows are handed out in tiles of block_threads/group rows (one row
// per tile for the last segment, whose group is the whole workgroup). With
// `exact`, rows of degree below kPrExactMax form one segment per degree.
// With `shift`, the degree thresholds are scaled by 2^shift (for passes that
// see a 2^-shift share of every row).
PrSegments plan_segments(const std::vector<uint32_t>& off, uint32_t n, uint32_t block_threads,
                         bool exact, uint32_t shift = 0, uint32_t rows_per_group = 1) {
  // Degree thresholds and the group size used below each.
  static const uint32_t kLimit[] = {2, 4, 8, 16, 32, 64, 1024, 0xFFFFFFFFu};
  const uint32_t kGroup[] = {1, 2, 4, 8, 16, 32, 64, block_threads};
  const uint32_t kPrThreads = block_threads;
  PrSegments s{};
  uint32_t row = 0, block = 0;
  auto degree = [&](uint32_t i) { return off[i + 1] - off[i]; };
  auto first_with_degree = [&](uint32_t from, uint32_t limit) {
    uint32_t lo = from, hi = n;
    while (lo < hi) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (degree(mid) >= limit) hi = mid;
      else lo = mid + 1;
    }
    return lo;
  };
  if (exact) {
    for (uint32_t d = 0; d < kPrExactMax && row < n; ++d) {
      const uint32_t end = first_with_degree(row, d + 1);
      if (end > row) {
        const uint32_t rows_per_block = kPrThreads * exact_rows_per_thread(d);
        s.seg[s.count++] = PrSegment{row, end, kPrExact | d, block, off[row]};
        block += (end - row + rows_per_block - 1) / rows_per_block;
        row = end;
      }
    }
  }
  for (int t = 0; t < 8 && row < n; ++t) {
    // first row at or after `row` whose degree reaches kLimit[t] (scaled)
    const uint64_t limit = static_cast<uint64_t>(kLimit[t]) << shift;
    uint32_t lo = row, hi = n;
    while (lo < hi) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (degree(mid) >= limit) hi = mid;
      else lo = mid + 1;
    }
    if (t == 7) lo = n;
    if (lo > row) {
      PrSegment seg;
      seg.row_begin = row;
      seg.row_end = lo;
      seg.group = kGroup[t];
      seg.block_begin = block;
      seg.entry_begin = off[row];
      const uint32_t rows_per_block = kGroup[t] >= kPrThreads ? 1 : rows_per_group * (kPrThreads / kGroup[t]);
      block += (lo - row + rows_per_block - 1) / rows_per_block;
      s.seg[s.count++] = seg;
      row = lo;
    }
  }
  s.total_blocks = block;
  return s;
}

}  // namespace

// Which SpMV runs (GC_PR_KERNEL): on a symmetric graph the push-pull
// iterations, by default with the cold entries binned (pr_bin_scatter,
// pr_bin_accum) and the hot stream sliced one row per lane (pr_spmv_slice,
// pr_slice_finish); with GC_PR_KERNEL=bin the same bins and the hot stream in
// 512-entry chunks (pr_spmv_bin, pr_merge_finish); with GC_PR_KERNEL=merge the
// nonzero-split pull that gathers the cold entries (pr_spmv_merge,
// pr_merge_finish), with GC_PR_KERNEL=push row tiles of G lanes per row
// (pr_spmv_push, pr_push_finish); otherwise, or with GC_PR_KERNEL=hot, the
// persistent pull kernel with the hot contributions in LDS (pr_spmv_hot); with
// GC_PR_KERNEL=rows one workgroup per row tile (pr_spmv). Measured on MI355X
// at R-MAT scale 22, per iteration that ran: slice 226 us and bin 266 us
// (scatter, accumulation, pull and finishing together), merge 380 us, push 401
// us (pull alone, both with the default sets), hot 527 us. The row tiles fill
// 70% of their lane slots there (a row of degree d on G lanes takes
// ceil(d / 8G) steps of 8G slots); the chunks of 512 entries fill all.
// (Earlier, 17 iterations at scale 22: hot 11.3 ms against rows 14.0 ms; a
// source-partitioned pass, one part of the sources per XCD with eight FP32
// partial sums per row, halved the L2 requests but was no faster: 12.1
// against 10.2-10.5 ms.)
enum class PrKernel { kPush, kMerge, kBin, kSlice, kHot, kRows };

static PrKernel choose_kernel(const Graph& g) {
  const char* env = std::getenv("GC_PR_KERNEL");
  if (env != nullptr && std::strcmp(env, "rows") == 0) return PrKernel::kRows;
  if (env != nullptr && std::strcmp(env, "hot") == 0) return PrKernel::kHot;
  if (!g.symmetric) return PrKernel::kHot;
  if (env != nullptr && std::strcmp(env, "push") == 0) return PrKernel::kPush;
  if (env != nullptr && std::strcmp(env, "merge") == 0) return PrKernel::kMerge;
  if (env != nullptr && std::strcmp(env, "bin") == 0) return PrKernel::kBin;
  return PrKernel::kSlice;
}

static uint32_t env_u32(const char* name, uint32_t fallback) {
  const char* env = std::getenv(name);
  return env != nullptr ? static_cast<uint32_t>(std::strtoul(env, nullptr, 10)) : fallback;
}

// The binned layout of the cold entries (pr_bin_layout.hpp) for the push set
// [n - push_count, n) and the hot set [n - hot_count, n) of a prepared
// symmetric graph: built on the host from the relabelled CSR (copied back
// once), uploaded, and kept with the graph until a call asks for other sizes.
static PrBinPlan& bin_plan(Graph& g, uint32_t push_count, uint32_t hot_count, uint32_t stage, uint32_t bin_rows,
                           uint32_t bin_values, bool sliced, uint32_t cap_groups, uint32_t waves) {
  if (g.pr_bin && g.pr_bin->push_count == push_count && g.pr_bin->hot_count == hot_count &&
      g.pr_bin->stage == stage && g.pr_bin->bin_rows == bin_rows && g.pr_bin->bin_values == bin_values &&
      g.pr_bin->sliced == sliced && (!sliced || (g.pr_bin->cap_groups == cap_groups && g.pr_bin->waves == waves)))
    return *g.pr_bin;
  g.pr_bin.reset();
  Stopwatch sw;
  Device& dev = *g.dev;
  std::vector<uint32_t> tgt(std::max<uint32_t>(g.m, 1));
  if (g.m > 0) copy_to_host(dev, tgt.data(), g.rl.targets.get(), sizeof(uint32_t) * g.m);
  PrBinLayout h;
  try {
    h = build_pr_bin_layout(g.rl.host_offsets.data(), tgt.data(), g.n, g.rl.dangling, push_count, hot_count,
                            stage, bin_rows, bin_values,
                            [](uint64_t count, const std::function<void(uint64_t, uint64_t)>& fn) {
                              parallel_for(count, fn, 1);
                            });
  } catch (const std::runtime_error& e) {
    fail_kernel(e.what());
  }
  auto plan = std::unique_ptr<PrBinPlan>(new PrBinPlan());
  PrBinPlan& p = *plan;
  p.push_count = push_count;
  p.hot_count = hot_count;
  p.stage = stage;
  p.bin_rows = bin_rows;
  p.bin_values = bin_values;
  p.zero_rows = h.zero_rows;
  p.push_base = h.push_base;
  p.hot_base = h.hot_base;
  p.hot_entries = static_cast<uint32_t>(h.hot16.size());
  p.chunks = static_cast<uint32_t>((static_cast<uint64_t>(p.hot_entries) + 511) / 512);
  p.subs = h.subs();
  p.bins = h.bins();
  p.values = h.bin_val.empty() ? 0 : h.bin_val.back();
  p.cold_entries = h.cold_entries;
  p.slot_rows = 8ull * h.block_group.size();
  auto up = [&](DeviceMemory& m, const void* src, size_t bytes) {
    m.allocate(std::max<size_t>(bytes, 16));
    if (bytes > 0) copy_to_device(dev, m.get(), src, bytes);
  };
  p.sliced = sliced;
  if (sliced) {
    PrSliceLayout sl;
    try {
      sl = build_pr_slice_layout(h.hoff.data(), h.hot16.data(), h.zero_rows, h.push_base, hot_count, h.hub_pieces,
                                 g.rl.host_offsets.data(), tgt.data(), cap_groups, waves,
                                 [](uint64_t count, const std::function<void(uint64_t, uint64_t)>& fn) {
                                   parallel_for(count, fn, 1);
                                 });
    } catch (const std::runtime_error& e) {
      fail_kernel(e.what());
    }
    p.cap_groups = cap_groups;
    p.waves = waves;
    p.slices = sl.slices;
    p.slice_items = sl.item_count();
    p.long_count = sl.long_count();
    p.partial_slots = sl.partial_slots;
    p.slice_entries = sl.slice_off.back();  // without the padding past the last slice
    up(p.slice_data, sl.data.data(), sizeof(uint16_t) * sl.data.size());
    up(p.items, sl.items.data(), sizeof(uint32_t) * sl.items.size());
    up(p.wave_first, sl.wave_first.data(), sizeof(uint32_t) * sl.wave_first.size());
    up(p.long_slices, sl.long_slices.data(), sizeof(uint32_t) * sl.long_slices.size());
    up(p.hub16, sl.hub16.data(), sizeof(uint16_t) * sl.hub16.size());
    p.slice_partial.allocate(sizeof(double) * 64 * std::max<uint32_t>(sl.partial_slots, 1));
  }
  std::vector<uint32_t>().swap(tgt);
  if (!sliced) {
    // the 16-bit stream padded to whole 512-entry chunks with entries that read the zero
    h.hot16.resize(static_cast<size_t>(p.chunks) * 512, static_cast<uint16_t>(hot_count));
    up(p.hot16, h.hot16.data(), sizeof(uint16_t) * h.hot16.size());
    up(p.hoff, h.hoff.data(), sizeof(uint32_t) * h.hoff.size());
  }
  up(p.slots, h.slots.data(), sizeof(uint16_t) * h.slots.size());
  up(p.block_group, h.block_group.data(), sizeof(uint32_t) * h.block_group.size());
  up(p.sub, h.sub.data(), sizeof(uint32_t) * h.sub.size());
  up(p.run_first, h.run_first.data(), sizeof(uint32_t) * h.run_first.size());
  up(p.run_info, h.run_info.data(), sizeof(uint32_t) * h.run_info.size());
  up(p.run_dst, h.run_dst.data(), sizeof(uint32_t) * h.run_dst.size());
  up(p.bin_row, h.bin_row.data(), sizeof(uint32_t) * h.bin_row.size());
  up(p.bin_val, h.bin_val.data(), sizeof(uint32_t) * h.bin_val.size());
  up(p.uloc, h.uloc.data(), sizeof(uint16_t) * h.uloc.size());
  up(p.hub_pieces, h.hub_pieces.data(), sizeof(uint32_t) * h.hub_pieces.size());
  p.pieces = h.pieces();
  p.runs = h.run_info.size();
  for (uint32_t i = 0; i < p.pieces; ++i) {
    const uint32_t row = h.hub_pieces[2 * i], b = h.hub_pieces[2 * i + 1];
    p.piece_entries += std::min(b + 512u, g.rl.host_offsets[row + 1]) - b;
  }
  const uint64_t vals = std::max<uint64_t>(p.values, 8);
  p.vals.allocate(sizeof(float) * vals);
  fill_words(dev, p.vals.get(), vals, 0);  // the padding values stay 0
  const uint32_t pull = p.push_base - p.zero_rows;
  p.cold_sum.allocate(sizeof(double) * std::max<uint32_t>(pull, 1));
  fill_words(dev, p.cold_sum.get(), 2ull * std::max<uint32_t>(pull, 1), 0);  // stays 0 without cold entries
  if (!sliced) {
    const uint64_t words = std::max<uint64_t>(1, static_cast<uint64_t>(p.chunks) * 16);  // whole chunks
    p.starts.allocate(sizeof(uint32_t) * words);
    fill_words(dev, p.starts.get(), words, 0);
    p.chunk_row.allocate(sizeof(uint32_t) * std::max<uint32_t>(p.chunks, 1));
  }
  if (!sliced && pull > 0) {
    const uint32_t* d_hoff = p.hoff.as<const uint32_t>();
    Grid1D sgrid = thread_per_item(pull);
    launch(dev, get_kernel(dev, "pagerank", "pr_merge_starts"), sgrid.blocks, sgrid.threads, 0, d_hoff,
           p.zero_rows, p.push_base, p.starts.as<uint32_t>());
    Grid1D cgrid = thread_per_item(p.chunks);
    launch(dev, get_kernel(dev, "pagerank", "pr_merge_chunks"), cgrid.blocks, cgrid.threads, 0, d_hoff,
           p.zero_rows, p.push_base, p.chunks, p.chunk_row.as<uint32_t>());
  }
  dev.sync();
  p.build_seconds = sw.seconds();
  g.pr_bin = std::move(plan);
  return *g.pr_bin;
}

// Push-pull iterations on a symmetric graph (pr_bin_scatter and pr_bin_accum,
// then pr_spmv_slice and pr_slice_finish or pr_spmv_bin and pr_merge_finish;
// or pr_spmv_merge and pr_merge_finish; or pr_spmv_push and pr_push_finish;
// then pr_finalize; the kernels' comments give the method). One 1024-lane workgroup per CU holds push_count 8-byte
// accumulators and hot_count FP32 contributions in LDS: 6,144 and 26,624 (152
// KB of the 160 KB). At R-MAT scale 22 that leaves 36.5% of the entries cold
// (below the hot set: binned, or gathered from global memory by the merge and
// push kernels; 58% for the pull kernel's 19,456 hot ids) and skips 21% of the
// target stream (the hubs' entries outside the push set). 8,192 and 22,528
// measured 385 us per iteration (merge) and 423 us (push) there. GC_PR_PUSH
// and GC_PR_HOT override the two sizes; GC_PR_STAGE (staged entries per
// scatter workgroup, at least 512), GC_PR_BIN_ROWS and GC_PR_BIN_VALUES (rows
// and values per accumulation bin) the binned layout's.
void pagerank_push(Graph& g, float damping, uint32_t max_iterations, float tolerance, float* d_out,
                   uint32_t* iterations, int32_t* converged, PrKernel kind) {
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  const bool merge = kind == PrKernel::kMerge, bin = kind == PrKernel::kBin, slice = kind == PrKernel::kSlice;
  const bool binned = bin || slice;  // the cold entries binned (pr_bin_scatter, pr_bin_accum)
  const uint32_t cu = static_cast<uint32_t>(std::max(1, dev.cu_count));
  // workgroups per CU (GC_PR_WGS, 1 or 2): each holds its own sets in its share of the LDS
  const uint32_t wgs = std::min(2u, std::max(1u, env_u32("GC_PR_WGS", 1u)));
  // lane groups take this many rows of a tile each, their loads in flight together (1 or 2,
  // GC_PR_ROWS; pr_spmv_push's tiles have this many slices). Measured at R-MAT scale 22,
  // one box: 357 us per launch with 1 row (420 per iteration that ran), 382 with 2.
  const uint32_t rows_per_group = env_u32("GC_PR_ROWS", 1u) >= 2 ? 2u : 1u;
  const size_t lds_limit = std::min<size_t>(dev.max_lds_per_block, 160 * 1024);
  // less the static LDS (pr_spmv_bin: 8 KB of staged row totals besides the partials)
  const size_t lds_cap = lds_limit / wgs - (bin ? 9216 : 1024);
  // at most 1/3 of the LDS for the accumulators, the rest for the contributions (and the
  // binned stream's zero)
  const uint32_t push_cap = static_cast<uint32_t>(lds_cap / 3 / 8);
  const uint32_t push_count = std::min({n, std::max(1u, env_u32("GC_PR_PUSH", 6144u)), push_cap});
  const uint32_t hot_cap = static_cast<uint32_t>((lds_cap - 8ull * push_count) / 4) - 1u;
  const uint32_t hot_count = std::min({n, std::max(push_count, env_u32("GC_PR_HOT", 26624u)), hot_cap});
  const uint32_t push_base = n - push_count;
  const uint32_t shmem = static_cast<uint32_t>(8ull * push_count + 4ull * (hot_count + (binned ? 1u : 0u)));
  const PrSegments segs = plan_segments(g.rl.host_offsets, push_base, kPrHotThreads, false,
                                        env_u32("GC_PR_SHIFT", 2u), rows_per_group);
  const uint32_t push_tiles = (push_count + 15) / 16;  // 16 hub rows (one wave each) per tile
  const uint32_t grid = cu * wgs;
  const uint32_t hub_blocks = (push_count + 15) / 16;
  // the nonzero-split pull: rows [zero_rows, push_base) own the entries [0, entries)
  const uint32_t zero_rows = std::min(g.rl.dangling, push_base);
  if ((merge || binned) && g.rl.host_offsets[zero_rows] != 0)
    fail_kernel("PageRank: the degree-ordered rows do not start with the rows of degree 0");
  PrBinPlan* plan = nullptr;
  if (binned) {
    // Defaults measured at R-MAT scale 22 (per launch, scatter + accumulation): 36,864 staged
    // entries and 65,536 values per bin 136 us; 40,448 and 196,608 124 us (longer runs per bin).
    const uint32_t stage = std::max(512u, std::min(env_u32("GC_PR_STAGE", 40448u),
                                                   static_cast<uint32_t>(lds_limit / 4 - 1)));
    const uint32_t bin_rows = std::max(1u, std::min(env_u32("GC_PR_BIN_ROWS", 16384u),
                                                    static_cast<uint32_t>(lds_limit / 8 - 1)));
    const uint32_t bin_values = std::max(8u, env_u32("GC_PR_BIN_VALUES", 196608u));
    // pr_spmv_slice: slices wider than this many groups of eight entries per lane are cut
    // into pieces (GC_PR_SLICE_CAP)
    const uint32_t cap_groups = std::max(1u, std::min(env_u32("GC_PR_SLICE_CAP", 16u), 4096u));
    plan = &bin_plan(g, push_count, hot_count, stage, bin_rows, bin_values, slice, cap_groups,
                     grid * (kPrHotThreads / 64));
    plan->grid = cu * wgs;
  }
  const uint32_t entries = bin ? plan->hot_entries : (merge ? g.rl.host_offsets[push_base] : 0u);
  const uint32_t chunks = static_cast<uint32_t>((static_cast<uint64_t>(entries) + 511) / 512);
  // finishing blocks after the hubs': 256 chunk tails each, or four wide slices
  const uint32_t fix_blocks = slice ? (plan->long_count + 3) / 4 : (chunks + 255) / 256;
  const uint32_t finish_blocks = hub_blocks + fix_blocks;
  const uint32_t partial_blocks = grid + finish_blocks;
  float* rank[2] = {g.buffer<float>("pr.rank0", sizeof(float) * n),
                    g.buffer<float>("pr.rank1", sizeof(float) * n)};
  float* contrib[2] = {g.buffer<float>("pr.contrib0", sizeof(float) * n),
                       g.buffer<float>("pr.contrib1", sizeof(float) * n)};
  double* partials = g.buffer<double>("pr.partials", sizeof(double) * 2 * partial_blocks);
  PrState* state = g.buffer<PrState>("pr.state", sizeof(PrState));
  uint32_t* work = g.buffer<uint32_t>("pr.work", 16);
  uint32_t* psplit = g.buffer<uint32_t>("pr.psplit", sizeof(uint32_t) * push_count);
  auto* push_partials =
      g.buffer<unsigned long long>("pr.push_partials", sizeof(unsigned long long) * grid * push_count);
  fill_words(dev, work, 4, 0);
  uint32_t* starts = nullptr;
  uint32_t* chunk_row = nullptr;
  double* head_partial = nullptr;
  double* tail_partial = nullptr;
  uint32_t* tail_row = nullptr;
  if (merge || bin) {
    const size_t nc = std::max<uint32_t>(1, chunks);
    head_partial = g.buffer<double>("pr.head_partial", sizeof(double) * nc);
    tail_partial = g.buffer<double>("pr.tail_partial", sizeof(double) * nc);
    tail_row = g.buffer<uint32_t>("pr.tail_row", sizeof(uint32_t) * nc);
  }
  if (merge) {
    const uint64_t words = std::max<uint64_t>(1, (static_cast<uint64_t>(entries) + 31) / 32);
    starts = g.buffer<uint32_t>("pr.starts", sizeof(uint32_t) * words);
    chunk_row = g.buffer<uint32_t>("pr.chunk_row", sizeof(uint32_t) * std::max<uint32_t>(1, chunks));
  }
  if (bin) {
    starts = plan->starts.as<uint32_t>();
    chunk_row = plan->chunk_row.as<uint32_t>();
  }

  const float r0 = 1.0f / static_cast<float>(n);
  const uint32_t* off = g.rl.offsets.as<const uint32_t>();
  const uint32_t* tgt = g.rl.targets.as<const uint32_t>();
  Grid1D igrid = thread_per_item(n);
  launch(dev, get_kernel(dev, "pagerank", "pr_init"), igrid.blocks, igrid.threads, 0, off, n, r0, rank[0],
         contrib[0]);
  if (!binned)  // the binned layout holds the hubs' entries inside the push set as pieces
    launch(dev, get_kernel(dev, "pagerank", "pr_push_split"), (push_count + 255) / 256, 256u, 0, off, tgt, n,
           push_count, psplit);
  // the row-start bitmap and the chunks' first rows depend on the graph and the push set only:
  // built once per graph (78 us per call at R-MAT 22 when rebuilt each time)
  if (merge && chunks > 0 && g.pr_merge_push != push_count) {
    g.pr_merge_push = 0;  // until the launches below have been made
    fill_words(dev, starts, (static_cast<uint64_t>(entries) + 31) / 32, 0);
    Grid1D sgrid = thread_per_item(push_base - zero_rows);
    launch(dev, get_kernel(dev, "pagerank", "pr_merge_starts"), sgrid.blocks, sgrid.threads, 0, off, zero_rows,
           push_base, starts);
    Grid1D cgrid = thread_per_item(chunks);
    launch(dev, get_kernel(dev, "pagerank", "pr_merge_chunks"), cgrid.blocks, cgrid.threads, 0, off, zero_rows,
           push_base, chunks, chunk_row);
    g.pr_merge_push = push_count;
  }
  PrState init{static_cast<double>(g.rl.dangling) * static_cast<double>(r0), 0.0, 0, 0};
  copy_to_device(dev, state, &init, sizeof(init));

  const char* spmv_name = slice ? "pr_spmv_slice" : bin ? "pr_spmv_bin" : (merge ? "pr_spmv_merge" : "pr_spmv_push");
  Kernel spmv = get_kernel(dev, "pagerank", spmv_name);
  Kernel finish = get_kernel(dev, "pagerank",
                             slice ? "pr_slice_finish" : (merge || bin ? "pr_merge_finish" : "pr_push_finish"));
  Kernel finalize = get_kernel(dev, "pagerank", "pr_finalize");
  Kernel scatter{}, accum{};
  if (binned) {
    scatter = get_kernel(dev, "pagerank", "pr_bin_scatter");
    accum = get_kernel(dev, "pagerank", "pr_bin_accum");
  }
  const double d = static_cast<double>(damping);
  const double tol = static_cast<double>(tolerance);
  PrState host_state = init;
  // the state is read back after iteration next_read (and after the last one)
  uint32_t next_read = kPrBatch;
  double prev_diff = 0.0;
  uint32_t prev_it = 0;
  for (uint32_t it = 0; it < max_iterations; ++it) {
    const int in = it & 1, out = in ^ 1;
    hipError_t err;
    if (binned) {
      if (plan->cold_entries > 0) {
        err = try_launch(dev, scatter, plan->subs, kPrHotThreads, 4u * (plan->stage + 1),
                         static_cast<const float*>(contrib[in]), plan->zero_rows, plan->hot_base,
                         plan->slots.as<const uint16_t>(), plan->block_group.as<const uint32_t>(),
                         plan->sub.as<const uint32_t>(), plan->run_first.as<const uint32_t>(),
                         plan->run_info.as<const uint32_t>(), plan->run_dst.as<const uint32_t>(),
                         plan->vals.as<float>(), static_cast<const PrState*>(state));
        if (err != hipSuccess) fail_kernel(std::string("pr_bin_scatter launch failed: ") + hipGetErrorString(err));
        err = try_launch(dev, accum, plan->bins, kPrHotThreads, 8u * (plan->bin_rows + 1),
                         plan->vals.as<const float>(), plan->uloc.as<const uint16_t>(),
                         plan->bin_row.as<const uint32_t>(), plan->bin_val.as<const uint32_t>(), plan->zero_rows,
                         plan->cold_sum.as<double>(), static_cast<const PrState*>(state));
        if (err != hipSuccess) fail_kernel(std::string("pr_bin_accum launch failed: ") + hipGetErrorString(err));
      }
      if (slice) {
        err = try_launch(dev, spmv, grid, kPrHotThreads, shmem, off, plan->hub16.as<const uint16_t>(),
                         plan->slice_data.as<const uint16_t>(),
                         plan->items.as<const uint4>(), plan->wave_first.as<const uint32_t>(),
                         plan->cold_sum.as<const double>(), static_cast<const float*>(contrib[in]),
                         static_cast<const float*>(rank[in]), contrib[out], rank[out], n, d,
                         static_cast<const PrState*>(state), hot_count, push_count, zero_rows,
                         plan->slice_partial.as<double>(), push_partials, partials);
      } else {
      err = try_launch(dev, spmv, grid, kPrHotThreads, shmem, off, tgt, plan->hot16.as<const uint16_t>(),
                       plan->cold_sum.as<const double>(), static_cast<const float*>(contrib[in]),
                       static_cast<const float*>(rank[in]), contrib[out], rank[out], n, d,
                       static_cast<const PrState*>(state), hot_count, push_count,
                       plan->hub_pieces.as<const uint32_t>(), plan->pieces, chunks, zero_rows,
                       reinterpret_cast<const uint8_t*>(starts), static_cast<const uint32_t*>(chunk_row),
                       head_partial, tail_partial, tail_row, push_partials, partials);
      }
    } else if (merge) {
      err = try_launch(dev, spmv, grid, kPrHotThreads, shmem, off, tgt, static_cast<const float*>(contrib[in]),
                       static_cast<const float*>(rank[in]), contrib[out], rank[out], n, d,
                       static_cast<const PrState*>(state), hot_count, push_count,
                       static_cast<const uint32_t*>(psplit), entries, chunks, zero_rows,
                       reinterpret_cast<const uint8_t*>(starts), static_cast<const uint32_t*>(chunk_row),
                       head_partial, tail_partial, tail_row, push_partials, partials);
    } else {
      err = try_launch(dev, spmv, grid, kPrHotThreads, shmem, off, tgt, static_cast<const float*>(contrib[in]),
                       static_cast<const float*>(rank[in]), contrib[out], rank[out], n, d, segs,
                       static_cast<const PrState*>(state), hot_count, push_count,
                       static_cast<const uint32_t*>(psplit), push_tiles, rows_per_group, push_partials, partials);
    }
    if (err != hipSuccess) fail_kernel(std::string(spmv_name) + " launch failed: " + hipGetErrorString(err));
    if (slice) {
      launch(dev, finish, finish_blocks, 256u, 0, off, static_cast<const float*>(rank[in]), contrib[out],
             rank[out], n, d, static_cast<const PrState*>(state), push_count,
             static_cast<const unsigned long long*>(push_partials), grid, hub_blocks,
             plan->long_slices.as<const uint32_t>(), plan->long_count, plan->slice_partial.as<const double>(),
             plan->cold_sum.as<const double>(), zero_rows, partials, grid);
    } else if (merge || bin) {
      launch(dev, finish, finish_blocks, 256u, 0, off, static_cast<const float*>(rank[in]), contrib[out],
             rank[out], n, d, static_cast<const PrState*>(state), push_count,
             static_cast<const unsigned long long*>(push_partials), grid, hub_blocks, chunks,
             static_cast<const double*>(head_partial), static_cast<const double*>(tail_partial),
             static_cast<const uint32_t*>(tail_row), partials, grid);
    } else {
      launch(dev, finish, finish_blocks, 256u, 0, off, static_cast<const float*>(rank[in]), contrib[out],
             rank[out], n, d, static_cast<const PrState*>(state), push_count,
             static_cast<const unsigned long long*>(push_partials), grid, partials, grid);
    }
    launch(dev, finalize, 1u, 1024u, 0, state, static_cast<const double*>(partials), partial_blocks, tol, work);
    if (it + 1 == next_read || it + 1 == max_iterations) {
      copy_to_host(dev, dev.pinned, state, sizeof(PrState));
      std::memcpy(&host_state, dev.pinned, sizeof(PrState));
      if (host_state.done) break;
      next_read = it + 1 + pr_next_batch(prev_diff, prev_it, host_state.diff, host_state.iterations, tol);
      prev_diff = host_state.diff;
      prev_it = host_state.iterations;
    }
  }
  const int latest = host_state.iterations & 1;
  Grid1D ogrid = thread_per_item((static_cast<uint64_t>(n) + 3) / 4);  // four entries per lane
  launch(dev, get_kernel(dev, "util", "gather_u32"), ogrid.blocks, ogrid.threads, 0,
         reinterpret_cast<const uint32_t*>(rank[latest]), g.rl.inv.as<const uint32_t>(),
         reinterpret_cast<uint32_t*>(d_out), static_cast<uint64_t>(n));
  dev.sync();
  *iterations = host_state.iterations;
  *converged = host_state.done ? 1 : 0;
}

void pagerank(Graph& g, float damping, uint32_t max_iterations, float tolerance, float* d_out,
              uint32_t* iterations, int32_t* converged) {
  Device& dev = *g.dev;
  prepare(g);
  const uint32_t n = g.n;
  const PrKernel kind = choose_kernel(g);
  if (kind == PrKernel::kPush || kind == PrKernel::kMerge || kind == PrKernel::kBin || kind == PrKernel::kSlice) {
    pagerank_push(g, damping, max_iterations, tolerance, d_out, iterations, converged, kind);
    return;
  }
  const bool rows_kernel = kind == PrKernel::kRows;
  const uint32_t block_threads = rows_kernel ? kPrThreads : kPrHotThreads;
  // Row kernel: rows of degree >= kPrHubLimit are summed in chunks (hub plan).
  const std::vector<uint32_t>& hoff = g.rl.host_offsets;
  uint32_t hub_begin = n;
  if (rows_kernel) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (hoff[mid + 1] - hoff[mid] >= kPrHubLimit) hi = mid;
      else lo = mid + 1;
    }
    hub_begin = lo;
  }
  // The hot kernel gives a row a quarter of the lanes of the row kernel (its
  // degree thresholds scaled by 4): with the eight-deep unrolled loop and one
  // guarded step for the rest, a lane then has 4-8 entries in flight at once.
  // Measured at R-MAT scale 22 (17 iterations, one box): thresholds x1 12.9
  // ms, x2 11.1, x4 10.2, x8 10.5.
  uint32_t hot_shift = 2;
  if (const char* shift_env = std::getenv("GC_PR_SHIFT"))
    hot_shift = static_cast<uint32_t>(std::strtoul(shift_env, nullptr, 10));
  const PrSegments segs = plan_segments(hoff, hub_begin, block_threads, rows_kernel, rows_kernel ? 0u : hot_shift);
  uint32_t num_hub_chunks = 0;
  uint32_t* hub_chunks = nullptr;
  uint32_t* hub_first = nullptr;
  uint32_t* hub_count = nullptr;
  double* hub_partials = nullptr;
  if (rows_kernel && hub_begin < n) {
    std::vector<uint32_t> chunks, first(n - hub_begin), count(n - hub_begin);
    for (uint32_t r = n; r-- > hub_begin;) {  // heaviest rows first
      first[r - hub_begin] = static_cast<uint32_t>(chunks.size() / 2);
      uint32_t c = 0;
      for (uint32_t e = hoff[r]; e < hoff[r + 1]; e += kPrHubChunk, ++c) {
        chunks.push_back(e);
        chunks.push_back(std::min(e + kPrHubChunk, hoff[r + 1]));
      }
      count[r - hub_begin] = c;
    }
    num_hub_chunks = static_cast<uint32_t>(chunks.size() / 2);
    hub_chunks = g.buffer<uint32_t>("pr.hub_chunks", sizeof(uint32_t) * chunks.size());
    hub_first = g.buffer<uint32_t>("pr.hub_first", sizeof(uint32_t) * first.size());
    hub_count = g.buffer<uint32_t>("pr.hub_count", sizeof(uint32_t) * count.size());
    hub_partials = g.buffer<double>("pr.hub_partials", sizeof(double) * num_hub_chunks);
    copy_to_device(dev, hub_chunks, chunks.data(), sizeof(uint32_t) * chunks.size());
    copy_to_device(dev, hub_first, first.data(), sizeof(uint32_t) * first.size());
    copy_to_device(dev, hub_count, count.data(), sizeof(uint32_t) * count.size());
  }
  const uint32_t finish_blocks = rows_kernel ? (n - hub_begin + kPrThreads - 1) / kPrThreads : 0;
  // Persistent kernel: two 1024-lane workgroups per CU (32 waves, the full
  // occupancy), each holding the hot set in half of the CU's LDS less 2 KB
  // for the reduction scratch: 19,456 ids (76 KB) on a 160 KB CU, which at
  // R-MAT scale 22 serve 42% of the gathers. One workgroup per CU with twice
  // the hot set (16 waves) measured slower: 14.2 ms against 11.3 ms, and
  // still 9.64 against 9.33 ms with sixteen loads in flight per lane.
  // GC_PR_HOT overrides the hot-set size.
  const uint32_t lds = static_cast<uint32_t>(std::min<size_t>(dev.max_lds_per_block, 160 * 1024));
  uint32_t hot_count = lds >= 8 * 1024 ? ((lds / 2 - 2048) / 4) & ~1023u : 1024u;
  if (const char* hot_env = std::getenv("GC_PR_HOT"))
    hot_count = std::max<uint32_t>(1024u, static_cast<uint32_t>(std::strtoul(hot_env, nullptr, 10)));
  hot_count = std::min({hot_count, n, dev.pr_hot_count_limit});
  const uint32_t cu = static_cast<uint32_t>(std::max(1, dev.cu_count));
  auto hot_grid = [&](uint32_t hot) { return cu * (hot * 4u > 80u * 1024u ? 1u : 2u); };
  // Row kernel: one workgroup per virtual block by default (the hardware
  // dispatcher balances them); GC_PR_RESIDENT=1 launches only the resident
  // workgroups, which then take the virtual blocks round-robin.
  const char* resident_env = std::getenv("GC_PR_RESIDENT");
  const bool resident = resident_env != nullptr && std::strcmp(resident_env, "1") == 0;
  const uint32_t virtual_blocks = num_hub_chunks + segs.total_blocks;
  uint32_t grid_blocks = rows_kernel ? std::max(1u, resident ? std::min(virtual_blocks,
                                                                        cu * kPrBlocksPerCu)
                                                             : virtual_blocks)
                                     : hot_grid(hot_count);
  // partial sums: one pair per spmv workgroup, then one per pr_hub_finish workgroup
  uint32_t partial_blocks = rows_kernel ? grid_blocks + finish_blocks : grid_blocks;
  float* rank[2] = {g.buffer<float>("pr.rank0", sizeof(float) * n),
                    g.buffer<float>("pr.rank1", sizeof(float) * n)};
  float* contrib[2] = {g.buffer<float>("pr.contrib0", sizeof(float) * n),
                       g.buffer<float>("pr.contrib1", sizeof(float) * n)};
  double* partials = g.buffer<double>(
      "pr.partials", sizeof(double) * 2 * std::max({1u, partial_blocks, 2 * cu}));
  PrState* state = g.buffer<PrState>("pr.state", sizeof(PrState));
  uint32_t* work = g.buffer<uint32_t>("pr.work", 16);
  fill_words(dev, work, 4, 0);

  const float r0 = 1.0f / static_cast<float>(n);
  const uint32_t* off = g.rl.offsets.as<const uint32_t>();
  const uint32_t* tgt = g.rl.targets.as<const uint32_t>();
  Grid1D grid = thread_per_item(n);
  launch(dev, get_kernel(dev, "pagerank", "pr_init"), grid.blocks, grid.threads, 0, off, n, r0,
         rank[0], contrib[0]);
  PrState init{static_cast<double>(g.rl.dangling) * static_cast<double>(r0), 0.0, 0, 0};
  copy_to_device(dev, state, &init, sizeof(init));

  Kernel spmv = get_kernel(dev, "pagerank", rows_kernel ? "pr_spmv" : "pr_spmv_hot");
  Kernel hub_finish = get_kernel(dev, "pagerank", "pr_hub_finish");
  Kernel finalize = get_kernel(dev, "pagerank", "pr_finalize");
  const double d = static_cast<double>(damping);
  const double tol = static_cast<double>(tolerance);
  PrState host_state = init;
  for (uint32_t it = 0; it < max_iterations; ++it) {
    const int in = it & 1, out = in ^ 1;
    if (rows_kernel) {
      launch(dev, spmv, grid_blocks, kPrThreads, 0, off, tgt, static_cast<const float*>(contrib[in]),
             static_cast<const float*>(rank[in]), contrib[out], rank[out], n, d, segs,
             static_cast<const PrState*>(state), partials, static_cast<const uint32_t*>(hub_chunks),
             num_hub_chunks, hub_partials);
      launch(dev, hub_finish, finish_blocks, kPrThreads, 0, off, static_cast<const float*>(rank[in]),
             contrib[out], rank[out], n, d, static_cast<const PrState*>(state), hub_begin,
             static_cast<const uint32_t*>(hub_first), static_cast<const uint32_t*>(hub_count),
             static_cast<const double*>(hub_partials), partials, grid_blocks);
    } else {
      for (;;) {
        hipError_t err = try_launch(dev, spmv, grid_blocks, kPrHotThreads, hot_count * 4u, off, tgt,
                                    static_cast<const float*>(contrib[in]),
                                    static_cast<const float*>(rank[in]), contrib[out], rank[out], n,
                                    d, segs, static_cast<const PrState*>(state), work, hot_count,
                                    partials);
        if (err == hipSuccess) break;
        // A device may refuse the dynamic LDS size: halve the hot set and retry.
        if (hot_count <= 1024) {
          fail_kernel(std::string("pr_spmv_hot launch failed: ") + hipGetErrorString(err));
        }
        hot_count /= 2;
        dev.pr_hot_count_limit = hot_count;
        grid_blocks = hot_grid(hot_count);
        partial_blocks = grid_blocks;  // one partial pair per workgroup of the new grid
      }
    }
    launch(dev, finalize, 1u, 1024u, 0, state, static_cast<const double*>(partials), partial_blocks, tol, work);
    if ((it + 1) % kPrBatch == 0 || it + 1 == max_iterations) {
      copy_to_host(dev, dev.pinned, state, sizeof(PrState));
      std::memcpy(&host_state, dev.pinned, sizeof(PrState));
      if (host_state.done) break;
    }
  }
  const int latest = host_state.iterations & 1;
  Grid1D ogrid = thread_per_item((static_cast<uint64_t>(n) + 3) / 4);  // four entries per lane
  launch(dev, get_kernel(dev, "util", "gather_u32"), ogrid.blocks, ogrid.threads, 0,
         reinterpret_cast<const uint32_t*>(rank[latest]), g.rl.inv.as<const uint32_t>(),
         reinterpret_cast<uint32_t*>(d_out), static_cast<uint64_t>(n));
  dev.sync();
  *iterations = host_state.iterations;
  *converged = host_state.done ? 1 : 0;
}

}  // namespace gc

using gc::guard;

extern "C" {

gc_status gc_pagerank_device(gc_graph* h, float damping, uint32_t max_iterations, float tolerance,
                             float* d_ranks, uint32_t* iterations, int32_t* converged,
                             double* secs) {
  return guard([&] {
    gc::Stopwatch sw;
    gc::GraphCall call(h);
    uint32_t it = 0;
    int32_t conv = 1;
    if (call.g->n > 0) {
      if (d_ranks == nullptr) gc::fail_kernel("ranks is null");
      gc::pagerank(*call.g, damping, max_iterations, tolerance, d_ranks, &it, &conv);
    }
    if (iterations) *iterations = it;
    if (converged) *converged = conv;
    if (secs) *secs = sw.seconds();
  });
}

gc_status gc_pagerank_layout(gc_graph* h, uint64_t* stats, uint32_t count) {
  return guard([&] {
    gc::GraphCall call(h);
    if (stats == nullptr && count > 0) gc::fail_kernel("stats is null");
    const gc::PrBinPlan* p = call.g->pr_bin.get();
    uint64_t v[25] = {};
    if (p != nullptr) {
      const uint64_t vals[25] = {p->push_count, p->hot_count, p->stage, p->bin_rows, p->bin_values,
                                 p->cold_entries, p->values, p->slot_rows, p->hot_entries, p->chunks,
                                 p->subs, p->bins, p->runs, p->pieces, p->piece_entries,
                                 static_cast<uint64_t>(p->push_base - p->zero_rows),
                                 static_cast<uint64_t>(p->hot_base > p->zero_rows ? p->hot_base - p->zero_rows : 0),
                                 static_cast<uint64_t>(p->build_seconds * 1e9), p->grid,
                                 p->sliced ? 1u : 0u, p->slices, p->slice_entries, p->slice_items,
                                 p->long_count, p->partial_slots};
      std::memcpy(v, vals, sizeof(v));
    }
    for (uint32_t i = 0; i < count && i < 25; ++i) stats[i] = v[i];
  });
}

gc_status gc_pagerank(gc_graph* h, float damping, uint32_t max_iterations, float tolerance,
                      float* ranks, uint32_t* iterations, int32_t* converged, double* secs) {
  return guard([&] {
    gc::Stopwatch sw;
    gc::GraphCall call(h);
    gc::Graph& g = *call.g;
    uint32_t it = 0;
    int32_t conv = 1;
    if (g.n > 0) {
      if (ranks == nullptr) gc::fail_kernel("ranks is null");
      float* d_out = g.buffer<float>("pr.out", sizeof(float) * g.n);
      gc::pagerank(g, damping, max_iterations, tolerance, d_out, &it, &conv);
      gc::copy_to_host(call.dev(), ranks, d_out, sizeof(float) * g.n);
    }
    if (iterations) *iterations = it;
    if (converged) *converged = conv;
    if (secs) *secs = sw.seconds();
  });
}

}  // extern "C"
