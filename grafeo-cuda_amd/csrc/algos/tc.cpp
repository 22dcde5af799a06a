// Triangle counting drivers.
//
// Replace triangle_counting::{triangle_count, triangle_count_oriented}
// (crates/grafeo-cuda-algos/src/triangle_counting.rs:40-210): same results,
// same early return for an edgeless graph, no forward-edge ceiling (the
// reference refuses more than 134,217,727 forward edges, launch.rs:27-41).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../runtime/common.hpp"
#include "../runtime/noalloc.hpp"

namespace gc {

namespace {

constexpr uint32_t kBitmap = 0, kHash = 1, kLists = 2;  // the scheduler's choices
constexpr uint32_t kHotBits = 32768;                    // TCB_HOT_BITS of tc.hip
constexpr uint32_t kMaxLaunches = 64;                   // launches per count
constexpr float kSwitchUs = 20.0f;                      // predicted saving that pays for a launch
constexpr uint32_t kMaxParts = 8;                       // middle-vertex parts of tc_bitmap (one per XCD)
constexpr uint32_t kDefaultParts = 4;
constexpr uint32_t kCounterStride = 64;                 // TCB_COUNTER_STRIDE of tc.hip
constexpr uint32_t kCounterWords = kCounterStride * kMaxParts;  // work counters of one launch

uint64_t read_result(Device& dev, const unsigned long long* result) {
  uint64_t total = 0;
  copy_to_host(dev, dev.pinned, result, sizeof(uint64_t));
  std::memcpy(&total, dev.pinned, sizeof(uint64_t));
  return total;
}

// Middle-vertex parts of tc_bitmap: GC_TC_PARTS (1..8), default 4 (each
// pair of XCDs shares a part). Measured on MI355X, R-MAT scale 22: 40.3 ms
// at 1 part, 21.1 at 8 (first version of the kernel); with the current
// kernel 4 parts beat 8 by 2-4% at scales 20 and 22 (half the per-part
// bitmap builds) and tie at scale 24 (145.6 against 146.0 ms).
uint32_t tc_parts() {
  const char* env = std::getenv("GC_TC_PARTS");
  if (env == nullptr) return kDefaultParts;
  const long p = std::strtol(env, nullptr, 10);
  return static_cast<uint32_t>(std::min<long>(std::max<long>(p, 1), kMaxParts));
}

// Rows handed out one at a time by tc_bitmap: the leading rows of [row_lo,
// row_hi) in descending order whose degree is at least GC_TC_HEAVY_DEGREE
// (default 64), at most 65536; the rest go in chunks of 64.
uint32_t tc_heavy_rows(const Graph& g, uint32_t row_lo, uint32_t row_hi, bool mid = false) {
  const char* env = std::getenv("GC_TC_HEAVY_DEGREE");
  // with the middle side on, 1024 (R-MAT scale 22: 5.10 ms against 5.28 at 64, 256 or 16)
  const uint32_t degree = env != nullptr ? static_cast<uint32_t>(std::strtoul(env, nullptr, 10)) : (mid ? 1024u : 64u);
  const std::vector<uint32_t>& off = g.rl.host_offsets;
  // rows are sorted by degree: the first row of degree >= `degree`
  uint32_t a = row_lo, b = row_hi;
  while (a < b) {
    const uint32_t mid = a + (b - a) / 2;
    if (off[mid + 1] - off[mid] >= degree) b = mid;
    else a = mid + 1;
  }
  return std::min<uint32_t>(row_hi - a, 65536);
}

// tc_bitmap16 (16-bit hot lists) unless GC_TC_HOT16=0 (tc_bitmap).
bool tc_hot16() {
  const char* env = std::getenv("GC_TC_HOT16");
  return env == nullptr || env[0] != '0';
}

// tc_bitmap16 walks the hot parts of at most GC_TC_QMAX octs (default 32) a
// quarter wave each, GC_TC_QUARTERS octs per lane and step (1 to 4, default
// 2; 0: every list a wave step, as the longer ones). Measured at R-MAT scale
// 22 per launch: a wave step per list 7.13 ms; quarters with 2 octs per lane
// up to 16 / 24 / 32 / 40 / 63 octs 7.04 / 6.89 / 6.86 / 6.89 / 7.06 ms; 1
// oct per lane up to 63 octs 7.65 ms (one load in flight per lane against
// two for the lockstep pairs), 4 up to 63 6.92 ms.
// Waves per workgroup of the 16-bit bitmap kernel: GC_TC_WG = 1 (tc_bitmap16,
// the bitmap at LDS address 0; the default) or 4 (tc_bitmap16_w4).
uint32_t tc_wg_waves() {
  const char* env = std::getenv("GC_TC_WG");
  return env != nullptr && env[0] == '4' ? 4u : 1u;
}

uint32_t tc_quarters() {
  const char* env = std::getenv("GC_TC_QUARTERS");
  const uint32_t d = env != nullptr ? static_cast<uint32_t>(std::strtoul(env, nullptr, 10)) : 2u;
  const char* qmax = std::getenv("GC_TC_QMAX");
  const uint32_t m = qmax != nullptr ? static_cast<uint32_t>(std::strtoul(qmax, nullptr, 10)) : 32u;
  // the cold parts' walk in tc_bitmap16: GC_TC_COLD_GROUP lanes per part (4, 8 or 16; at
  // R-MAT scale 22 16 measured best: 4.79 ms against 4.85 with 8 and 4.89 with 4)
  const char* cg = std::getenv("GC_TC_COLD_GROUP");
  const uint32_t g = cg != nullptr ? static_cast<uint32_t>(std::strtoul(cg, nullptr, 10)) : 16u;
  return std::min(d, 4u) | std::min(m, 4096u) << 8 | std::min(g, 255u) << 24;
}

// Masks of the top ids in the 16-bit lists (tc_bitmap16): GC_TC_TOP ids (a
// power of two from 128 to 8192; 0: no masks), default 4096, i.e. 32 octs
// (512 bytes) per mask. At R-MAT scale 22 the top 1024 ids alone are 51% of
// the entries the kernel streams (29.4 GB as offsets per count) and the top
// 2048 ids 69%. Measured tc_bitmap16 per launch with masks of 0 / 1024 /
// 2048 / 4096 / 8192 ids: R-MAT 22 8.55 / 7.16 / 6.74 / 6.69 / 7.17 ms, R-MAT
// 24 83.5 / 78.7 / 74.9 / 74.0 / 75.5 ms, R-MAT 20 1.85 / 1.67 / 1.62 / 1.63 /
// 1.76 ms.
uint32_t tc_top_octs() {
  const char* env = std::getenv("GC_TC_TOP");
  const uint32_t bits = env != nullptr ? static_cast<uint32_t>(std::strtoul(env, nullptr, 10)) : 4096u;
  if (bits == 0) return 0;
  if (bits < 128 || bits > 8192 || (bits & (bits - 1)) != 0)
    fail_kernel("GC_TC_TOP must be 0 or a power of two from 128 to 8192");
  return bits / 128;
}

// First id of the masked top ids: the top `octs` * 128 ids, rounded up to a
// multiple of 32 (so its bitmap words start at a word boundary), and never
// below hot_base.
uint32_t tc_top_base(uint32_t n, uint32_t hot_base, uint32_t octs) {
  const uint32_t bits = octs * 128u;
  const uint32_t t = n > bits ? ((n - bits) + 31u) & ~31u : 0u;
  return std::max(t, hot_base);
}

// The 16-bit hot lists of tc_bitmap16, built once per graph (again if the
// parts or the masks change): at most (m + 7n) / 8 + 1 octs of 16 bytes (a
// mask replaces more octs than it takes), and 12 bytes per vertex.
void ensure_hot16(Graph& g, uint32_t hot_base, uint32_t parts, uint32_t top_octs) {
  if (g.tc_hot16_ready && g.tc_hot16_parts == parts && g.tc_hot16_top == top_octs) return;
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  if (n == 0) return;  // no rows, nothing to count
  const uint64_t m = g.rl.host_offsets[n];
  const uint64_t max_quads = (m + 7ull * n) / 8 + 1;  // 16-byte octs of eight offsets
  if (max_quads > 0xFFFFFFFFull) fail_kernel("graph too large for the 16-bit triangle-count lists");
  const uint64_t rows = std::max<uint64_t>(n, 1);
  g.tc_split.allocate(sizeof(uint32_t) * rows);
  g.tc_hw.allocate(sizeof(uint32_t) * 2 * rows);
  g.tc_words.allocate(sizeof(uint16_t) * 8 * max_quads);
  DeviceMemory qoff, scan, tsplit;
  qoff.allocate(sizeof(uint32_t) * rows);
  scan.allocate(sizeof(uint32_t) * scan_scratch_words(rows));
  tsplit.allocate(sizeof(uint32_t) * (top_octs != 0 ? rows : 1));
  const uint32_t top_base = tc_top_base(n, hot_base, top_octs);
  const uint32_t* offsets = g.rl.offsets.as<const uint32_t>();
  const uint32_t* tgt = g.rl.targets.as<const uint32_t>();
  Grid1D vg = thread_per_item(rows);
  launch(dev, get_kernel(dev, "tc", "tc_hot16_rows"), vg.blocks, vg.threads, 0, g.rl.fwd.as<const uint32_t>(),
         offsets + 1, tgt, n, hot_base, parts, g.tc_split.as<uint32_t>(), qoff.as<uint32_t>(), top_base, top_octs,
         tsplit.as<uint32_t>());
  exclusive_scan_u32(dev, qoff.as<uint32_t>(), n, scan.as<uint32_t>());
  if (top_octs != 0) fill_words(dev, g.tc_words.as<uint32_t>(), 4 * max_quads, 0);  // the masks are ORed in
  Grid1D eg = thread_per_item(std::max<uint64_t>(m, n));
  launch(dev, get_kernel(dev, "tc", "tc_hot16_fill"), eg.blocks, eg.threads, 0, offsets, tgt, n, m, hot_base,
         parts, static_cast<const uint32_t*>(g.tc_split.as<uint32_t>()), static_cast<const uint32_t*>(qoff.as<uint32_t>()),
         g.tc_hw.as<uint32_t>() /* uint2[n] */, g.tc_words.as<uint16_t>(), top_base, top_octs,
         static_cast<const uint32_t*>(tsplit.as<uint32_t>()));
  dev.sync();  // qoff, scan and tsplit are released on return
  g.tc_hot16_ready = true;
  g.tc_hot16_parts = parts;
  g.tc_hot16_top = top_octs;
}

// First relabelled row of degree >= d (rows are sorted by degree), or n.
uint32_t first_row_of_degree(const Graph& g, uint64_t d) {
  const std::vector<uint32_t>& off = g.rl.host_offsets;
  uint32_t lo = 0, hi = g.n;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (off[mid + 1] - off[mid] >= d) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// The middle-vertex side of tc_bitmap16 (tc_mid16) unless GC_TC_MID=0. An edge
// u->v goes to it when the octs of its walk plus GC_TC_MID_COST (default 2) are
// fewer than the octs tc_bitmap16 streams for L(v); GC_TC_MID_UNIT (default
// 8192) entries of a hot row's in-list make one work unit. Measured at R-MAT
// scale 22 (rows split at 1024 entries): 5.97 ms without the middle side, 5.30
// with it; costs 0, 2, 8 and 32 octs within 5%; units of 512, 2048, 8192 and
// 32768 entries 5.96, 5.56, 5.30 and 5.30 ms (2048 and 512 at split 256).
bool tc_mid() {
  const char* env = std::getenv("GC_TC_MID");
  return env == nullptr || std::strcmp(env, "0") != 0;
}

// The 32-bit middle tier (tc_mid32: middle vertices among the top
// GC_TC_MID32_IDS ids, default 131072, but below the hot ids) with
// GC_TC_MID32=1. Off by default: at R-MAT scale 22 it takes 7.5 M edges (1.8e9
// streamed entries) from tc_bitmap16, which then runs 1.0 ms faster, but its
// own walks take 1.39 ms (short walks a quarter wave each, a global round trip
// per step, and the searches of L(v)'s cold prefix): 5.84 against 5.29 ms.
bool tc_mid32() {
  const char* env = std::getenv("GC_TC_MID32");
  return env != nullptr && std::strcmp(env, "0") != 0;
}

uint32_t tc_mid32_ids() {
  const char* env = std::getenv("GC_TC_MID32_IDS");
  return env == nullptr ? 131072u : static_cast<uint32_t>(std::strtoul(env, nullptr, 10));
}

uint32_t tc_mid_cost() {
  const char* env = std::getenv("GC_TC_MID_COST");
  return env == nullptr ? 2u : static_cast<uint32_t>(std::strtoul(env, nullptr, 10));
}

// Work units a wave of tc_mid16 / tc_mid32 takes per atomic on their shared
// counter: GC_TC_MID_GRAB (default 1). The units come heaviest first, so a wave
// that takes several at once holds the tail: at R-MAT scale 22, 1, 8 and 32
// units per atomic took 5.29, 8.2 and 24.4 ms.
uint32_t tc_mid_grab() {
  const char* env = std::getenv("GC_TC_MID_GRAB");
  return env == nullptr ? 1u : std::max(1u, static_cast<uint32_t>(std::strtoul(env, nullptr, 10)));
}

// Lanes per walk of at most GC_TC_QMAX octs in tc_mid16: GC_TC_MID_GROUP = 4,
// 8 or 16 (two octs per lane and step), or 0 for tc_bitmap16's quarter walk.
// Half the walks hold at most 7 octs: at R-MAT scale 22, 4.79 ms with 8 lanes
// against 4.81 (4), 4.87 (16) and 4.91 (the quarter walk).
uint32_t tc_mid_group() {
  const char* env = std::getenv("GC_TC_MID_GROUP");
  return env == nullptr ? 8u : static_cast<uint32_t>(std::strtoul(env, nullptr, 10));
}

uint32_t tc_mid_unit() {
  const char* env = std::getenv("GC_TC_MID_UNIT");
  const uint32_t u = env == nullptr ? 8192u : static_cast<uint32_t>(std::strtoul(env, nullptr, 10));
  return std::max(64u, u);
}

// The middle side's lists, records, edge bits and units, built once per graph
// after the 16-bit lists (again if the parts, the masks or the per-edge cost
// change, since they decide which edges go where). At R-MAT scale 22: 16-byte
// octs of the DAG rows' hot parts (about 160 MB), 8 bytes per entry of the top
// 32768 rows, one bit per CSR entry.
void ensure_mid(Graph& g, uint32_t hot_base, uint32_t parts, uint32_t top_octs, uint32_t cost) {
  const uint32_t unit = tc_mid_unit();
  const uint32_t tier32 = tc_mid32() ? std::max(1u, tc_mid32_ids()) : 0u;
  if (g.tc_mid_ready && g.tc_mid_key[0] == parts && g.tc_mid_key[1] == top_octs && g.tc_mid_key[2] == cost &&
      g.tc_mid_key[3] == unit && g.tc_mid_key[4] == tier32)
    return;
  g.tc_mid_ready = false;
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  const std::vector<uint32_t>& off = g.rl.host_offsets;
  const uint64_t m = off[n];
  const uint32_t* offsets = g.rl.offsets.as<const uint32_t>();
  const uint32_t* tgt = g.rl.targets.as<const uint32_t>();
  const uint64_t rows = std::max<uint64_t>(n, 1);
  DeviceMemory hq, scan;
  hq.allocate(sizeof(uint32_t) * (rows + 1));
  scan.allocate(sizeof(uint32_t) * scan_scratch_words(rows + 1));
  Grid1D vg = thread_per_item(rows);
  launch(dev, get_kernel(dev, "tc", "tc_mid_rows"), vg.blocks, vg.threads, 0, g.tc_split.as<const uint32_t>(),
         offsets + 1, n, hq.as<uint32_t>());
  fill_words(dev, hq.as<uint32_t>() + n, 1, 0);
  exclusive_scan_u32(dev, hq.as<uint32_t>(), static_cast<uint64_t>(n) + 1, scan.as<uint32_t>());
  uint32_t total_octs = 0;
  copy_to_host(dev, &total_octs, hq.as<uint32_t>() + n, sizeof(uint32_t));
  g.tc_mid_octs.ensure(16ull * (static_cast<uint64_t>(total_octs) + 1));
  Grid1D eg = thread_per_item(std::max<uint64_t>(m, n));
  launch(dev, get_kernel(dev, "tc", "tc_mid_fill"), eg.blocks, eg.threads, 0, offsets, tgt, n, m, hot_base,
         g.tc_split.as<const uint32_t>(), static_cast<const uint32_t*>(hq.as<uint32_t>()),
         g.tc_mid_octs.as<uint16_t>());
  const uint32_t rec_base = off[hot_base];
  const uint64_t recs = m - rec_base;
  g.tc_mid_rec.ensure(sizeof(uint32_t) * 2 * std::max<uint64_t>(recs, 1));
  const uint64_t bit_words = (m + 31) / 32 + 1;
  g.tc_midbits.ensure(sizeof(uint32_t) * bit_words);
  fill_words(dev, g.tc_midbits.as<uint32_t>(), bit_words, 0);
  // the rows (of the hot ones, then of the 32-bit tier's) with a non-empty walk: only they get units
  DeviceMemory walks;
  std::vector<uint8_t> has_walk(n - hot_base, 0);
  const uint64_t walk_words = (std::max<uint64_t>(n, 1) + 3) / 4;
  walks.allocate(4 * walk_words);
  fill_words(dev, walks.as<uint32_t>(), walk_words, 0);
  if (recs > 0) {
    Grid1D rg = thread_per_item(recs);
    launch(dev, get_kernel(dev, "tc", "tc_mid_records"), rg.blocks, rg.threads, 0, offsets, tgt,
           g.rl.fwd.as<const uint32_t>(), n, hot_base, g.tc_split.as<const uint32_t>(),
           g.tc_hw.as<const uint32_t>() /* uint2[n] */, static_cast<const uint32_t*>(hq.as<uint32_t>()), cost,
           g.tc_mid_rec.as<uint32_t>() /* uint2 */, g.tc_midbits.as<uint32_t>(), walks.as<uint8_t>());
    copy_to_host(dev, has_walk.data(), walks.as<uint8_t>(), has_walk.size());
  }
  // the 32-bit tier: rows [mid_base, hot_base), when the graph has ids below the hot ones
  const uint32_t mid_base = n > tier32 ? n - tier32 : 0u;
  g.tc_mid32_num_units = 0;
  g.tc_mid32_base = mid_base;
  g.tc_mid32_rec_base = off[mid_base];
  std::vector<uint32_t> units32;
  if (tier32 && mid_base < hot_base && off[hot_base] > off[mid_base]) {
    const uint64_t recs32 = off[hot_base] - off[mid_base];
    g.tc_mid32_rec.ensure(sizeof(uint32_t) * 2 * recs32);
    Grid1D rg = thread_per_item(recs32);
    fill_words(dev, walks.as<uint32_t>(), walk_words, 0);
    launch(dev, get_kernel(dev, "tc", "tc_mid32_records"), rg.blocks, rg.threads, 0, offsets, tgt,
           g.rl.fwd.as<const uint32_t>(), mid_base, hot_base, g.tc_split.as<const uint32_t>(),
           g.tc_hw.as<const uint32_t>() /* uint2[n] */, cost, g.tc_mid32_rec.as<uint32_t>() /* uint2 */,
           g.tc_midbits.as<uint32_t>(), walks.as<uint8_t>());
    std::vector<uint8_t> has_walk32(hot_base - mid_base, 0);
    copy_to_host(dev, has_walk32.data(), walks.as<uint8_t>(), has_walk32.size());
    for (uint32_t v = hot_base; v-- > mid_base;) {
      if (!has_walk32[v - mid_base]) continue;
      for (uint32_t b = off[v]; b < off[v + 1]; b += unit) {
        units32.insert(units32.end(), {v, b, std::min<uint32_t>(off[v + 1], b + unit), 0u});
      }
    }
    g.tc_mid32_num_units = static_cast<uint32_t>(units32.size() / 4);
    g.tc_mid32_units.ensure(sizeof(uint32_t) * std::max<size_t>(units32.size(), 4));
    copy_to_device(dev, g.tc_mid32_units.as<uint32_t>(), units32.data(), sizeof(uint32_t) * units32.size());
  }
  g.tc_rowwork.ensure(rows);
  launch(dev, get_kernel(dev, "tc", "tc_mid_rowwork"), vg.blocks, vg.threads, 0, g.rl.fwd.as<const uint32_t>(),
         offsets + 1, n, static_cast<const uint32_t*>(g.tc_midbits.as<uint32_t>()), g.tc_rowwork.as<uint8_t>());
  // units: the hot rows' entry ranges in pieces of at most tc_mid_unit() entries, top row first
  std::vector<uint32_t> units;
  for (uint32_t v = n; v-- > hot_base;) {
    if (!has_walk[v - hot_base]) continue;
    for (uint32_t b = off[v]; b < off[v + 1]; b += unit) {
      units.insert(units.end(), {v, b, std::min<uint32_t>(off[v + 1], b + unit), 0u});
    }
  }
  g.tc_mid_num_units = static_cast<uint32_t>(units.size() / 4);
  g.tc_mid_units.ensure(sizeof(uint32_t) * std::max<size_t>(units.size(), 4));
  if (!units.empty()) copy_to_device(dev, g.tc_mid_units.as<uint32_t>(), units.data(), sizeof(uint32_t) * units.size());
  dev.sync();  // hq and scan are released on return
  g.tc_mid_rec_base = rec_base;
  g.tc_mid_key[0] = parts;
  g.tc_mid_key[1] = top_octs;
  g.tc_mid_key[2] = cost;
  g.tc_mid_key[3] = unit;
  g.tc_mid_key[4] = tier32;
  g.tc_mid_ready = true;
}

// Shortest L(u) that tc_bitmap splits over the parts (shorter rows run whole,
// once): GC_TC_SPLIT, default 256. With the 16-bit lists, R-MAT scale 22
// counted in 8.92 ms at 256 against 9.13 at 64 (scale 24: 86.0 against
// 87.3 ms; scale 20 unchanged): a split row builds its bitmap once per part.
// With the middle side on (mid), whose walks take most of the long lists, the
// default is 1024: 5.30 ms at R-MAT scale 22 against 5.56 at 256 (64: 5.99).
uint32_t tc_split_min(bool mid = false) {
  const char* env = std::getenv("GC_TC_SPLIT");
  if (env == nullptr) return mid ? 1024 : 256;
  return static_cast<uint32_t>(std::strtoul(env, nullptr, 10));
}

// Runs tc_count over the lists L(x) = targets[lo[x] .. hi[x]) for rows
// [row_lo, row_hi), adding into `result`; `counters` are two zeroed words.
void launch_tc_count(Graph& g, const uint32_t* lo, const uint32_t* hi, const uint32_t* targets,
                     uint32_t row_lo, uint32_t row_hi, bool descending, bool after_only,
                     uint32_t* counters, unsigned long long* result) {
  Device& dev = *g.dev;
  const uint32_t rows = row_hi - row_lo;
  const uint32_t heavy = descending ? std::min<uint32_t>(rows, 65536) : 0;
  const uint32_t light_chunk = descending ? 64 : 16;
  const uint32_t waves = std::max(1, dev.cu_count) * 32u;
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(waves, rows));
  uint32_t descending_u = descending ? 1 : 0, after_u = after_only ? 1 : 0;
  launch(dev, get_kernel(dev, "tc", "tc_count"), grid, kWaveSize, 0, lo, hi, targets, row_hi, row_lo,
         descending_u, after_u, heavy, light_chunk, counters, result);
}

// One symmetric-graph method over rows [row_lo, row_hi) of the relabelled DAG;
// `counters` are kCounterWords zeroed words. With `no_alloc` (bitmap method
// only) nothing is built, allocated or timed: tc_bitmap16 runs if its 16-bit
// lists already exist for the current parts, tc_bitmap (which needs none)
// otherwise; both count the same triangles.
void launch_method(Graph& g, uint32_t method, uint32_t row_lo, uint32_t row_hi, uint32_t* counters,
                   unsigned long long* result, bool no_alloc = false) {
  Device& dev = *g.dev;
  const uint32_t n = g.n;
  if (row_hi <= row_lo) return;
  const uint32_t* fwd = g.rl.fwd.as<const uint32_t>();
  const uint32_t* end = g.rl.offsets.as<const uint32_t>() + 1;
  const uint32_t* tgt = g.rl.targets.as<const uint32_t>();
  const uint32_t rows = row_hi - row_lo;
  const uint32_t heavy = std::min<uint32_t>(rows, 65536);
  const uint32_t light_chunk = 64;
  const uint32_t full_grid = std::max(1, dev.cu_count) * 8u;
  // Workgroups of four waves, no more than the tasks (single heavy rows and
  // light chunks, per part) can keep busy: a small row range does not pay
  // for thousands of waves that only find the work counters drained.
  auto grid_for = [&](uint64_t tasks) {
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(full_grid, (tasks + 3) / 4)));
  };
  if (method == kLists) {
    launch_tc_count(g, fwd, end, tgt, row_lo, row_hi, true, true, counters, result);
  } else if (method == kHash) {
    const uint64_t tasks = heavy + (static_cast<uint64_t>(rows - heavy) + light_chunk - 1) / light_chunk;
    launch(dev, get_kernel(dev, "tc", "tc_hash"), grid_for(tasks), 4 * kWaveSize, 0, fwd, end, tgt, row_hi,
           row_lo, heavy, light_chunk, counters, result);
  } else {
    // hot range [hot_base, n): at most kHotBits ids, hot_base a multiple of 32
    const uint32_t hot_base = n > kHotBits ? ((n - kHotBits) + 31u) & ~31u : 0u;
    const uint32_t parts = tc_parts();
    uint32_t single = tc_heavy_rows(g, row_lo, row_hi);
    uint64_t tasks = parts * (single + (static_cast<uint64_t>(rows - single) + light_chunk - 1) / light_chunk);
    // workgroups of `waves` waves, as many waves in all as grid_for's four-wave workgroups
    auto go = [&](const char* name, uint32_t waves, auto... args) {
      const Kernel k = get_kernel(dev, "tc", name);
      const uint32_t blocks = grid_for(tasks) * (4 / waves);
      if (no_alloc) launch_untimed(dev, k, blocks, waves * kWaveSize, 0, args...);
      else launch(dev, k, blocks, waves * kWaveSize, 0, args...);
    };
    // the no-alloc call uses the lists as they were built (their masks included)
    const uint32_t top_octs = no_alloc ? g.tc_hot16_top : tc_top_octs();
    const bool hot16 = tc_hot16() && (!no_alloc || (g.tc_hot16_ready && g.tc_hot16_parts == parts));
    if (hot16) {
      if (!no_alloc) ensure_hot16(g, hot_base, parts, top_octs);
      const uint32_t top_word0 = (tc_top_base(n, hot_base, top_octs) - hot_base) / 32u;
      const bool w1 = tc_wg_waves() == 1;
      // the middle side takes its edges when the launch covers every row that can close a
      // triangle (the rows below the first of degree 2 have at most one DAG entry)
      bool mid = !no_alloc && tc_mid() && row_hi == n && row_lo <= first_row_of_degree(g, 2);
      if (mid) ensure_mid(g, hot_base, parts, top_octs, tc_mid_cost());
      // no walk on either middle kernel (e.g. a grid, where every L(v) is short): the edge bits would
      // only skip edges that close no triangle, at a load per edge (6.36 against 6.02 ms, 4096 x 4096 grid)
      mid = mid && g.tc_mid_num_units + g.tc_mid32_num_units > 0;
      const uint32_t* midbits = mid ? g.tc_midbits.as<const uint32_t>() : nullptr;
      const uint8_t* rowwork = mid ? g.tc_rowwork.as<const uint8_t>() : nullptr;
      if (mid) {
        single = tc_heavy_rows(g, row_lo, row_hi, true);
        tasks = parts * (single + (static_cast<uint64_t>(rows - single) + light_chunk - 1) / light_chunk);
      }
      go(w1 ? "tc_bitmap16" : "tc_bitmap16_w4", w1 ? 1u : 4u, fwd, end, tgt, g.tc_split.as<const uint32_t>(),
         g.tc_hw.as<const uint32_t>() /* uint2[n] */, g.tc_words.as<const uint32_t>() /* uint4 octs */, row_hi,
         row_lo, hot_base, single, light_chunk, parts, tc_split_min(mid), counters, result, top_word0, top_octs,
         tc_quarters(), midbits, rowwork);
      if (mid && g.tc_mid_num_units > 0) {
        // one-wave workgroups, as many as tc_bitmap16's, on work counter word 2 (unused by the parts)
        const uint32_t blocks = std::min<uint32_t>(grid_for(g.tc_mid_num_units * 4ull) * 4u, g.tc_mid_num_units);
        launch(dev, get_kernel(dev, "tc", "tc_mid16"), blocks, kWaveSize, 0, g.rl.offsets.as<const uint32_t>(), tgt,
               fwd, hot_base, g.tc_mid_rec.as<const uint32_t>() /* uint2 */, g.tc_mid_rec_base,
               g.tc_mid_octs.as<const uint32_t>() /* uint4 */, g.tc_mid_units.as<const uint32_t>() /* uint4 */,
               g.tc_mid_num_units, tc_quarters(), tc_mid_grab(), tc_mid_group(), counters + 2, result);
      }
      if (mid && g.tc_mid32_num_units > 0) {
        // one-wave workgroups, as tc_mid16's, on work counter word 3
        const uint32_t blocks = std::min<uint32_t>(grid_for(g.tc_mid32_num_units * 4ull) * 4u, g.tc_mid32_num_units);
        launch(dev, get_kernel(dev, "tc", "tc_mid32"), blocks, kWaveSize, 0, g.rl.offsets.as<const uint32_t>(), tgt,
               fwd, g.tc_split.as<const uint32_t>(), hot_base, g.tc_mid32_rec.as<const uint32_t>() /* uint2 */,
               g.tc_mid32_rec_base, g.tc_mid32_units.as<const uint32_t>() /* uint4 */, g.tc_mid32_num_units,
               tc_mid_grab(), counters + 3, result);
      }
    } else {
      go("tc_bitmap", 4u, fwd, end, tgt, row_hi, row_lo, hot_base, single, light_chunk, parts, tc_split_min(),
         counters, result);
    }
  }
}

// Scratch: [0..1] u64 result, then kCounterWords work counters per launch.
uint32_t* tc_scratch(Graph& g) {
  constexpr uint64_t words = 2 + uint64_t(kCounterWords) * kMaxLaunches;
  uint32_t* scratch = g.buffer<uint32_t>("tc.scratch", sizeof(uint32_t) * words);
  fill_words(*g.dev, scratch, words, 0);
  return scratch;
}

// Degree buckets of the relabelled rows: bucket b holds the rows of degree in
// [2^b, 2^(b+1)) (contiguous, rows being sorted by degree); rows of degree 0
// or 1 close no triangle and are left out. bounds has buckets + 1 entries.
std::vector<uint32_t> bucket_bounds(const Graph& g) {
  const std::vector<uint32_t>& off = g.rl.host_offsets;
  const uint32_t n = g.n;
  auto first_with_degree = [&](uint64_t d) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (off[mid + 1] - off[mid] >= d) hi = mid;
      else lo = mid + 1;
    }
    return lo;
  };
  std::vector<uint32_t> bounds{first_with_degree(2)};
  for (uint32_t b = 1; bounds.back() < n && b < 32; ++b) bounds.push_back(first_with_degree(2ull << b));
  if (bounds.back() < n) bounds.push_back(n);
  // drop empty buckets
  std::vector<uint32_t> out{bounds[0]};
  for (size_t i = 1; i < bounds.size(); ++i)
    if (bounds[i] > out.back()) out.push_back(bounds[i]);
  return out;
}

using BucketInfo = Graph::TcBucket;

// Work of every bucket (cached per graph: the graph never changes).
const std::vector<BucketInfo>& bucket_info(Graph& g) {
  if (g.tc_buckets_ready) return g.tc_buckets;
  Device& dev = *g.dev;
  const std::vector<uint32_t> bounds = bucket_bounds(g);
  std::vector<BucketInfo> info;
  if (bounds.size() >= 2) {
    const uint32_t buckets = static_cast<uint32_t>(bounds.size() - 1);
    uint32_t* d_bounds = g.buffer<uint32_t>("tc.bucket_bounds", sizeof(uint32_t) * bounds.size());
    auto* d_stats = g.buffer<unsigned long long>("tc.bucket_stats", sizeof(unsigned long long) * 5 * buckets);
    copy_to_device(dev, d_bounds, bounds.data(), sizeof(uint32_t) * bounds.size());
    fill_words(dev, d_stats, 10ull * buckets, 0);
    const uint32_t grid = std::max(1, dev.cu_count) * 8u;
    const uint32_t n = g.n;
    const uint32_t hot_base = n > kHotBits ? ((n - kHotBits) + 31u) & ~31u : 0u;  // as the bitmap kernels
    launch(dev, get_kernel(dev, "tc", "tc_bucket_stats"), grid, 256u, 0, g.rl.fwd.as<const uint32_t>(),
           g.rl.offsets.as<const uint32_t>() + 1, g.rl.targets.as<const uint32_t>(),
           static_cast<const uint32_t*>(d_bounds), buckets, hot_base, d_stats);
    std::vector<unsigned long long> stats(5 * buckets);
    copy_to_host(dev, stats.data(), d_stats, sizeof(unsigned long long) * stats.size());
    for (uint32_t b = 0; b < buckets; ++b) {
      info.push_back({bounds[b], bounds[b + 1], stats[5 * b], stats[5 * b + 1], stats[5 * b + 2],
                      stats[5 * b + 3], stats[5 * b + 4]});
    }
  }
  g.tc_buckets.assign(info.begin(), info.end());
  g.tc_buckets_ready = true;
  return g.tc_buckets;
}

float lg2(double x) { return static_cast<float>(std::log2(x + 1.0)); }

// The 12 features of one bucket's decision.
void bucket_features(const Graph& g, const std::vector<BucketInfo>& all, size_t b, float* f) {
  const BucketInfo& k = all[b];
  uint64_t total_probes = 0;
  for (const BucketInfo& x : all) total_probes += x.probes;
  const uint32_t degree_lo = g.rl.host_offsets[k.row_lo + 1] - g.rl.host_offsets[k.row_lo];
  f[0] = lg2(g.n);
  f[1] = lg2(g.m);
  f[2] = lg2(static_cast<double>(k.rows));
  f[3] = lg2(static_cast<double>(k.entries));
  f[4] = lg2(static_cast<double>(k.probes));
  f[5] = lg2(static_cast<double>(k.max_out));
  f[6] = lg2(static_cast<double>(k.entries) / std::max<uint64_t>(k.rows, 1));
  f[7] = lg2(degree_lo);
  f[8] = lg2(g.n - k.row_lo);
  f[9] = lg2(static_cast<double>(k.probes) / std::max<uint64_t>(k.entries, 1));
  // the share of the bucket's probes among the bitmap kernels' hot ids, whose lists the
  // bitmap kernels stream from 16-bit words (the rest are binary-searched): near 1 on
  // R-MAT graphs, near 0 on large grids, where the binary-search lists win
  f[10] = static_cast<float>(static_cast<double>(k.hot_probes) / static_cast<double>(std::max<uint64_t>(k.probes, 1)));
  f[11] = lg2(static_cast<double>(total_probes));
}

// Which symmetric-graph method the fixed rule runs (GC_TC_METHOD): "bitmap"
// (default, tc_bitmap16 or tc_bitmap), "hash" (tc_hash), "lists" (binary-search
// intersection, tc_count) or "wedge" (enumeration at the middle vertex,
// tc_wedge). All four count the same triangles.
const char* tc_method() {
  const char* env = std::getenv("GC_TC_METHOD");
  return env != nullptr ? env : "bitmap";
}

}  // namespace

void tc_bitmap_schedule(const Graph& g, uint32_t* hot_base, uint32_t* heavy, uint32_t* parts, uint32_t* split_min) {
  const uint32_t n = g.n;
  *hot_base = n > kHotBits ? ((n - kHotBits) + 31u) & ~31u : 0u;
  *heavy = tc_heavy_rows(g, 0, n);
  *parts = tc_parts();
  *split_min = tc_split_min();
}

// Triangles of a symmetric graph by enumeration at the middle vertex of the
// relabelled DAG (tc_wedge).
uint64_t run_tc_wedge(Graph& g) {
  Device& dev = *g.dev;
  uint32_t* scratch = tc_scratch(g);
  const uint32_t bits = wedge_table_bits(g.rl);
  const uint32_t grid = std::max(1, dev.cu_count) * 8u;
  auto* result = reinterpret_cast<unsigned long long*>(scratch);
  const uint32_t hub_chunk = Relabelled::kHubChunk;
  launch(dev, get_kernel(dev, "tc", "tc_wedge"), grid, kWaveSize, (1u << bits) * 4u,
         g.rl.offsets.as<const uint32_t>(), g.rl.targets.as<const uint32_t>(),
         g.rl.fwd.as<const uint32_t>(), g.n, g.rl.hub_units.as<const uint32_t>(), g.rl.num_hub_units,
         hub_chunk, bits, scratch + 2, result);
  return read_result(dev, result);
}

uint64_t triangle_count_symmetric(Graph& g) {
  prepare(g);
  Device& dev = *g.dev;
  g.last_schedule.clear();
  if (g.symmetric) {
    if (g.tc_sched == nullptr) {
      const char* method = tc_method();
      if (std::strcmp(method, "wedge") == 0) return run_tc_wedge(g);
      const uint32_t m = std::strcmp(method, "hash") == 0    ? kHash
                         : std::strcmp(method, "lists") == 0 ? kLists
                                                             : kBitmap;
      uint32_t* scratch = tc_scratch(g);
      auto* result = reinterpret_cast<unsigned long long*>(scratch);
      // rows of degree 0 or 1 (sorted first) close no triangle
      launch_method(g, m, bucket_bounds(g).front(), g.n, scratch + 2, result);
      return read_result(dev, result);
    }
    // learned: one decision per degree bucket, adjacent buckets with the same
    // method merged into one launch. The decisions depend on the graph and
    // the network only, so they are computed once per network version.
    const std::vector<BucketInfo>& buckets = bucket_info(g);
    uint32_t* scratch = tc_scratch(g);
    auto* result = reinterpret_cast<unsigned long long*>(scratch);
    if (buckets.empty()) return 0;
    if (g.tc_choice_sched != g.tc_sched || g.tc_choice_version != g.tc_sched->version ||
        g.tc_choice.size() != buckets.size()) {
      std::vector<float> x(buckets.size() * kSchedFeatures), y(buckets.size() * kSchedChoices);
      for (size_t b = 0; b < buckets.size(); ++b) bucket_features(g, buckets, b, &x[b * kSchedFeatures]);
      sched_predict(*g.tc_sched, x.data(), static_cast<uint32_t>(buckets.size()), y.data());
      g.tc_choice.assign(buckets.size(), kBitmap);
      for (size_t b = 0; b < buckets.size(); ++b) {
        const float* yb = &y[b * kSchedChoices];
        uint32_t best = 0;
        for (uint32_t c = 1; c < kSchedChoices; ++c)
          if (yb[c] < yb[best]) best = c;
        // Leaving the bitmap for a bucket costs a launch of its own (adjacent
        // buckets of one method share a launch): only worth it when the
        // predicted saving exceeds kSwitchUs.
        if (best != kBitmap && std::exp(yb[kBitmap]) - std::exp(yb[best]) < kSwitchUs) best = kBitmap;
        g.tc_choice[b] = best;
      }
      g.tc_choice_sched = g.tc_sched;
      g.tc_choice_version = g.tc_sched->version;
    }
    const std::vector<uint32_t>& choice = g.tc_choice;
    g.last_schedule = choice;
    uint32_t launches = 0;
    for (size_t b = 0; b < buckets.size();) {
      size_t e = b + 1;
      while (e < buckets.size() && choice[e] == choice[b]) ++e;
      if (launches >= kMaxLaunches) fail_kernel("too many triangle-count launches");
      launch_method(g, choice[b], buckets[b].row_lo, buckets[e - 1].row_hi, scratch + 2 + kCounterWords * launches,
                    result);
      ++launches;
      b = e;
    }
    return read_result(dev, result);
  }
  // Any other CSR: the suffix of each list above its row, in the caller's ids.
  ensure_fwd_original(g);
  uint32_t* scratch = tc_scratch(g);
  auto* result = reinterpret_cast<unsigned long long*>(scratch);
  launch_tc_count(g, g.fwd_original.as<const uint32_t>(), g.offsets.as<const uint32_t>() + 1,
                  g.targets.as<const uint32_t>(), 0, g.n, false, true, scratch + 2, result);
  return read_result(dev, result);
}

// Calls of the allocation-free count on one GPU share its static scratch:
// they are serialized per device ordinal (also across gc_device handles).
std::mutex& noalloc_mutex(int ordinal) {
  static std::mutex mutexes[64];
  return mutexes[static_cast<unsigned>(ordinal) % 64u];
}

// gc_triangle_count_noalloc: the fixed rule's bitmap count of a prepared
// symmetric graph with its u64 result and one launch's work counters in the
// device's static scratch (zeroed first by the fill kernel), so the call
// allocates, builds and loads nothing.
uint64_t triangle_count_noalloc(Graph& g) {
  Device& dev = *g.dev;
  if (dev.noalloc_scratch == nullptr)
    fail(GC_ERR_DRIVER, "the allocation-free scratch is not available on this device: " + dev.noalloc_error);
  if (!g.prepared)
    fail_kernel("gc_triangle_count_noalloc needs a prepared graph: call gc_graph_prepare first");
  if (!g.symmetric) fail_kernel("gc_triangle_count_noalloc needs a symmetric CSR");
  constexpr uint64_t words = 2 + uint64_t(kCounterWords);  // u64 result, then the work counters
  if (words * sizeof(uint32_t) > kNoAllocScratchBytes) {
    fail_kernel("the triangle count needs " + std::to_string(words * sizeof(uint32_t)) +
                " bytes of scratch, more than the " + std::to_string(kNoAllocScratchBytes) +
                "-byte static scratch");
  }
  const std::lock_guard<std::mutex> lock(noalloc_mutex(dev.ordinal));
  uint32_t* scratch = static_cast<uint32_t*>(dev.noalloc_scratch);
  const Grid1D fg = thread_per_item(words);
  launch_untimed(dev, get_kernel(dev, "util", "fill_u32"), fg.blocks, fg.threads, 0, scratch, words, 0u);
  auto* result = reinterpret_cast<unsigned long long*>(scratch);
  // rows of degree 0 or 1 (sorted first) close no triangle
  launch_method(g, kBitmap, first_row_of_degree(g, 2), g.n, scratch + 2, result, true);
  return read_result(dev, result);
}

uint64_t triangle_count_oriented(Graph& g) {
  prepare(g);
  // sum over entries u->v of |N(u) intersect N(v)| is invariant under the
  // relabelling; the degree order lets the heaviest rows start first.
  uint32_t* scratch = tc_scratch(g);
  auto* result = reinterpret_cast<unsigned long long*>(scratch);
  launch_tc_count(g, g.rl.offsets.as<const uint32_t>(), g.rl.offsets.as<const uint32_t>() + 1,
                  g.rl.targets.as<const uint32_t>(), 0, g.n, true, false, scratch + 2, result);
  return read_result(*g.dev, result);
}

// Training data: every bucket's features and the time of each method run on
// that bucket alone; the methods' counts must agree.
uint32_t tc_bucket_timings(Graph& g, uint32_t max_buckets, float* features, float* times_us) {
  prepare(g);
  if (!g.symmetric) fail_kernel("triangle-count bucket timings need a symmetric CSR");
  Device& dev = *g.dev;
  const std::vector<BucketInfo>& buckets = bucket_info(g);
  uint32_t count = 0;
  for (size_t b = 0; b < buckets.size() && count < max_buckets; ++b, ++count) {
    bucket_features(g, buckets, b, features + static_cast<uint64_t>(count) * kSchedFeatures);
    uint64_t reference = 0;
    for (uint32_t c = 0; c < kSchedChoices; ++c) {
      float fastest = 0.0f;
      for (int rep = 0; rep < 2; ++rep) {
        uint32_t* scratch = tc_scratch(g);
        auto* result = reinterpret_cast<unsigned long long*>(scratch);
        dev.sync();
        Stopwatch sw;
        launch_method(g, c, buckets[b].row_lo, buckets[b].row_hi, scratch + 2, result);
        const uint64_t got = read_result(dev, result);
        const float us = static_cast<float>(sw.seconds() * 1e6);
        fastest = rep == 0 ? us : std::min(fastest, us);
        if (c == 0 && rep == 0) reference = got;
        else if (got != reference) fail_kernel("triangle-count methods disagree on a bucket");
      }
      times_us[static_cast<uint64_t>(count) * kSchedChoices + c] = fastest;
    }
  }
  return count;
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

gc_status gc_triangle_count_noalloc(gc_graph* h, uint64_t* total_triangles) {
  return guard([&] {
    gc::GraphCall call(h);
    if (total_triangles == nullptr) gc::fail_kernel("total_triangles is null");
    *total_triangles = call.g->m == 0 ? 0 : gc::triangle_count_noalloc(*call.g);
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

gc_status gc_tc_bucket_timings(gc_graph* h, uint32_t max_buckets, float* features, float* times_us,
                               uint32_t* buckets) {
  return guard([&] {
    gc::GraphCall call(h);
    if ((features == nullptr || times_us == nullptr) && max_buckets > 0)
      gc::fail_kernel("features or times_us is null");
    uint32_t got = 0;
    if (call.g->m > 0) got = gc::tc_bucket_timings(*call.g, max_buckets, features, times_us);
    if (buckets) *buckets = got;
  });
}

}  // extern "C"
