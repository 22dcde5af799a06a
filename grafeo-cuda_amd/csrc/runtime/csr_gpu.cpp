// CSR construction on the device (kernels of sort.hip): the same arrays as
// from_edge_pairs (crates/grafeo-cuda-core/src/csr_export.rs:202-239) -- the
// reverse of every non-loop pair added when symmetric, sorted by (source,
// target), duplicates removed -- built by a stable LSD radix sort of packed
// keys on the GPU instead of the host.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"

namespace gc {

namespace {

constexpr uint64_t kTile = 4096;  // RS_TILE of sort.hip and the compaction tile of util.hip

unsigned bits_of(uint64_t max_value) {
  unsigned b = 0;
  while (b < 64 && (max_value >> b) != 0) ++b;
  return b;
}

}  // namespace

unsigned long long* radix_sort(Device& dev, unsigned long long* keys, unsigned long long* tmp, uint64_t n,
                               uint32_t bits) {
  if (n <= 1) return keys;
  const uint64_t blocks = (n + kTile - 1) / kTile;
  if (blocks * 256 > 0xFFFFFFFFull) fail_kernel("too many keys for the device sort");
  DeviceMemory hist, scan;
  hist.allocate(sizeof(uint32_t) * 256 * blocks);
  scan.allocate(sizeof(uint32_t) * scan_scratch_words(256 * blocks));
  Kernel k_hist = get_kernel(dev, "sort", "rs_histogram");
  Kernel k_scatter = get_kernel(dev, "sort", "rs_scatter");
  const uint32_t grid = static_cast<uint32_t>(blocks);
  for (uint32_t shift = 0; shift < bits; shift += 8) {
    launch(dev, k_hist, grid, 256u, 0, static_cast<const unsigned long long*>(keys), n, shift,
           hist.as<uint32_t>());
    exclusive_scan_u32(dev, hist.as<uint32_t>(), 256 * blocks, scan.as<uint32_t>());
    launch(dev, k_scatter, grid, 256u, 0, static_cast<const unsigned long long*>(keys),
           static_cast<const uint32_t*>(nullptr), n, shift, static_cast<const uint32_t*>(hist.as<uint32_t>()),
           tmp, static_cast<uint32_t*>(nullptr));
    std::swap(keys, tmp);
  }
  return keys;
}

// As radix_sort, carrying a u32 value with each key (stable). Returns the
// buffers holding the result, keys and values.
std::pair<unsigned long long*, uint32_t*> radix_sort_pairs(Device& dev, unsigned long long* keys, uint32_t* vals,
                                                           unsigned long long* keys_tmp, uint32_t* vals_tmp,
                                                           uint64_t n, uint32_t bits) {
  if (n <= 1) return {keys, vals};
  const uint64_t blocks = (n + kTile - 1) / kTile;
  if (blocks * 256 > 0xFFFFFFFFull) fail_kernel("too many keys for the device sort");
  DeviceMemory hist, scan;
  hist.allocate(sizeof(uint32_t) * 256 * blocks);
  scan.allocate(sizeof(uint32_t) * scan_scratch_words(256 * blocks));
  Kernel k_hist = get_kernel(dev, "sort", "rs_histogram");
  Kernel k_scatter = get_kernel(dev, "sort", "rs_scatter");
  const uint32_t grid = static_cast<uint32_t>(blocks);
  for (uint32_t shift = 0; shift < bits; shift += 8) {
    launch(dev, k_hist, grid, 256u, 0, static_cast<const unsigned long long*>(keys), n, shift, hist.as<uint32_t>());
    exclusive_scan_u32(dev, hist.as<uint32_t>(), 256 * blocks, scan.as<uint32_t>());
    launch(dev, k_scatter, grid, 256u, 0, static_cast<const unsigned long long*>(keys),
           static_cast<const uint32_t*>(vals), n, shift, static_cast<const uint32_t*>(hist.as<uint32_t>()), keys_tmp,
           vals_tmp);
    std::swap(keys, keys_tmp);
    std::swap(vals, vals_tmp);
  }
  return {keys, vals};
}

struct DeviceCsr {
  DeviceMemory offsets;  // u32[n + 1]
  DeviceMemory targets;  // u32[m]
  uint32_t n = 0;
  uint32_t m = 0;
};

// Builds the CSR of `num_pairs` pairs on the device: `pairs` in host memory,
// or with `d_pairs_in` already on the device (then `pairs` may be null).
// Pairs naming a node >= n are an InvalidCsr error with the host builder's
// message (the first such pair in input order, when the host pairs are given).
void build_csr_on_device(Device& dev, const uint32_t* pairs, size_t num_pairs, uint32_t n, bool symmetric,
                         DeviceCsr& out, const uint32_t* d_pairs_in = nullptr) {
  const unsigned id_bits = std::max(1u, bits_of(n == 0 ? 0 : n - 1));
  const uint64_t num_keys = symmetric ? 2ull * num_pairs : num_pairs;
  DeviceMemory d_pairs, keys_a, keys_b, bad;
  keys_a.allocate(sizeof(unsigned long long) * std::max<uint64_t>(num_keys, 1));
  keys_b.allocate(sizeof(unsigned long long) * std::max<uint64_t>(num_keys, 1));
  bad.allocate(16);
  fill_words(dev, bad.get(), 1, 0);
  const uint32_t* src_pairs = d_pairs_in;
  if (src_pairs == nullptr) {
    d_pairs.allocate(sizeof(uint32_t) * 2 * std::max<size_t>(num_pairs, 1));
    if (num_pairs > 0) copy_to_device(dev, d_pairs.get(), pairs, sizeof(uint32_t) * 2 * num_pairs);
    src_pairs = d_pairs.as<uint32_t>();
  }
  Grid1D pg = thread_per_item(std::max<uint64_t>(num_pairs, 1));
  launch(dev, get_kernel(dev, "sort", "csr_make_keys"), pg.blocks, pg.threads, 0, src_pairs,
         static_cast<uint64_t>(num_pairs), n, id_bits, symmetric ? 1u : 0u, keys_a.as<unsigned long long>(),
         bad.as<uint32_t>());
  uint32_t bad_count = 0;
  copy_to_host(dev, &bad_count, bad.get(), sizeof(uint32_t));
  if (bad_count > 0) {
    if (pairs == nullptr) fail_csr("an edge has a node id >= num_nodes " + std::to_string(n));
    // the first offending pair in input order, with the host builder's message
    for (size_t i = 0; i < num_pairs; ++i) {
      const uint32_t s = pairs[2 * i], d = pairs[2 * i + 1];
      if (s >= n || d >= n) {
        fail_csr("edge (" + std::to_string(s) + ", " + std::to_string(d) +
                 ") has a node id >= num_nodes " + std::to_string(n));
      }
    }
  }
  d_pairs.release();
  unsigned long long* sorted =
      radix_sort(dev, keys_a.as<unsigned long long>(), keys_b.as<unsigned long long>(), num_keys, 2 * id_bits);
  unsigned long long* spare = sorted == keys_a.as<unsigned long long>() ? keys_b.as<unsigned long long>()
                                                                        : keys_a.as<unsigned long long>();
  // first copies of each key, compacted in order into `spare`
  uint64_t unique = 0;
  if (num_keys > 0) {
    DeviceMemory flags, tile_counts;
    flags.allocate(std::max<uint64_t>(num_keys, 4));
    const uint64_t tiles = (num_keys + kTile - 1) / kTile;
    tile_counts.allocate(sizeof(uint32_t) * (tiles + 1));
    Grid1D kg = thread_per_item(num_keys);
    launch(dev, get_kernel(dev, "sort", "csr_first_copies"), kg.blocks, kg.threads, 0,
           static_cast<const unsigned long long*>(sorted), num_keys, flags.as<uint8_t>());
    launch(dev, get_kernel(dev, "util", "compact_count"), static_cast<uint32_t>(tiles), 256u, 0,
           static_cast<const uint8_t*>(flags.as<uint8_t>()), num_keys, tile_counts.as<uint32_t>());
    launch(dev, get_kernel(dev, "util", "exclusive_scan_small"), 1u, 1024u, 0, tile_counts.as<uint32_t>(),
           static_cast<uint32_t>(tiles), tile_counts.as<uint32_t>() + tiles);
    uint32_t total = 0;
    copy_to_host(dev, &total, tile_counts.as<uint32_t>() + tiles, sizeof(uint32_t));
    unique = total;
    launch(dev, get_kernel(dev, "sort", "csr_compact_keys"), static_cast<uint32_t>(tiles), 256u, 0,
           static_cast<const unsigned long long*>(sorted), static_cast<const uint8_t*>(flags.as<uint8_t>()),
           num_keys, static_cast<const uint32_t*>(tile_counts.as<uint32_t>()), spare);
  }
  if (unique > 0xFFFFFFFFull) {
    fail_csr("graph too large for u32 indices: " + std::to_string(unique) + " targets");
  }
  out.n = n;
  out.m = static_cast<uint32_t>(unique);
  out.offsets.allocate(sizeof(uint32_t) * (static_cast<size_t>(n) + 1));
  out.targets.allocate(sizeof(uint32_t) * std::max<uint64_t>(unique, 1));
  Grid1D cg = thread_per_item(std::max<uint64_t>(unique, static_cast<uint64_t>(n) + 1));
  launch(dev, get_kernel(dev, "sort", "csr_from_keys"), cg.blocks, cg.threads, 0,
         static_cast<const unsigned long long*>(spare), unique, n, static_cast<uint32_t>(id_bits),
         out.targets.as<uint32_t>(), out.offsets.as<uint32_t>());
  dev.sync();
}

// Compacts the flagged keys of sorted[0, count) in order into `out` and
// returns how many there are (tiles of kTile entries: compact_count and
// exclusive_scan_small of util.hip, csr_compact_keys of sort.hip).
uint64_t compact_flagged(Device& dev, const unsigned long long* sorted, uint64_t count, const uint8_t* flags,
                         unsigned long long* out) {
  if (count == 0) return 0;
  const uint64_t tiles = (count + kTile - 1) / kTile;
  DeviceMemory tile_counts;
  tile_counts.allocate(sizeof(uint32_t) * (tiles + 1));
  launch(dev, get_kernel(dev, "util", "compact_count"), static_cast<uint32_t>(tiles), 256u, 0, flags, count,
         tile_counts.as<uint32_t>());
  launch(dev, get_kernel(dev, "util", "exclusive_scan_small"), 1u, 1024u, 0, tile_counts.as<uint32_t>(),
         static_cast<uint32_t>(tiles), tile_counts.as<uint32_t>() + tiles);
  uint32_t total = 0;
  copy_to_host(dev, &total, tile_counts.as<uint32_t>() + tiles, sizeof(uint32_t));
  launch(dev, get_kernel(dev, "sort", "csr_compact_keys"), static_cast<uint32_t>(tiles), 256u, 0, sorted, flags,
         count, static_cast<const uint32_t*>(tile_counts.as<uint32_t>()), out);
  return total;
}

// Builds the oriented CSR of from_edge_pairs_oriented
// (crates/grafeo-cuda-core/src/csr_export.rs:299-350) on the device: canonical
// (min, max) keys, sorted, first copies of the non-loop keys kept, undirected
// degrees, each edge oriented from the lower degree to the higher (ties from
// the lower id), sorted again, offsets by binary search.
void build_oriented_csr_on_device(Device& dev, const uint32_t* pairs, size_t num_pairs, uint32_t n,
                                  DeviceCsr& out) {
  const unsigned id_bits = std::max(1u, bits_of(n == 0 ? 0 : n - 1));
  const uint64_t slots = std::max<uint64_t>(num_pairs, 1);
  DeviceMemory d_pairs, keys_a, keys_b, bad;
  d_pairs.allocate(sizeof(uint32_t) * 2 * slots);
  keys_a.allocate(sizeof(unsigned long long) * slots);
  keys_b.allocate(sizeof(unsigned long long) * slots);
  bad.allocate(16);
  fill_words(dev, bad.get(), 1, 0);
  if (num_pairs > 0) copy_to_device(dev, d_pairs.get(), pairs, sizeof(uint32_t) * 2 * num_pairs);
  Grid1D pg = thread_per_item(slots);
  launch(dev, get_kernel(dev, "sort", "csr_undirected_keys"), pg.blocks, pg.threads, 0,
         static_cast<const uint32_t*>(d_pairs.as<uint32_t>()), static_cast<uint64_t>(num_pairs), n, id_bits,
         keys_a.as<unsigned long long>(), bad.as<uint32_t>());
  uint32_t bad_count = 0;
  copy_to_host(dev, &bad_count, bad.get(), sizeof(uint32_t));
  if (bad_count > 0) {
    for (size_t i = 0; i < num_pairs; ++i) {
      const uint32_t s = pairs[2 * i], d = pairs[2 * i + 1];
      if (s >= n || d >= n) {
        fail_csr("edge (" + std::to_string(s) + ", " + std::to_string(d) + ") has a node id >= num_nodes " +
                 std::to_string(n));
      }
    }
  }
  d_pairs.release();
  unsigned long long* sorted =
      radix_sort(dev, keys_a.as<unsigned long long>(), keys_b.as<unsigned long long>(), num_pairs, 2 * id_bits);
  unsigned long long* spare = sorted == keys_a.as<unsigned long long>() ? keys_b.as<unsigned long long>()
                                                                        : keys_a.as<unsigned long long>();
  uint64_t unique = 0;
  if (num_pairs > 0) {
    DeviceMemory flags;
    flags.allocate(std::max<uint64_t>(num_pairs, 4));
    Grid1D kg = thread_per_item(num_pairs);
    launch(dev, get_kernel(dev, "sort", "csr_unique_edges"), kg.blocks, kg.threads, 0,
           static_cast<const unsigned long long*>(sorted), static_cast<uint64_t>(num_pairs), id_bits,
           flags.as<uint8_t>());
    unique = compact_flagged(dev, sorted, num_pairs, flags.as<uint8_t>(), spare);
  }
  if (unique > 0xFFFFFFFFull) fail_csr("graph too large for u32 indices: " + std::to_string(unique) + " targets");
  unsigned long long* oriented = sorted;  // free again: the unique keys are in `spare`
  if (unique > 0) {
    DeviceMemory degree;
    degree.allocate(sizeof(uint32_t) * std::max<uint64_t>(n, 1));
    fill_words(dev, degree.get(), std::max<uint64_t>(n, 1), 0);
    Grid1D ug = thread_per_item(unique);
    launch(dev, get_kernel(dev, "sort", "csr_edge_degrees"), ug.blocks, ug.threads, 0,
           static_cast<const unsigned long long*>(spare), unique, id_bits, degree.as<uint32_t>());
    launch(dev, get_kernel(dev, "sort", "csr_orient_keys"), ug.blocks, ug.threads, 0,
           static_cast<const unsigned long long*>(spare), unique, id_bits,
           static_cast<const uint32_t*>(degree.as<uint32_t>()), oriented);
    oriented = radix_sort(dev, oriented, spare, unique, 2 * id_bits);
  }
  out.n = n;
  out.m = static_cast<uint32_t>(unique);
  out.offsets.allocate(sizeof(uint32_t) * (static_cast<size_t>(n) + 1));
  out.targets.allocate(sizeof(uint32_t) * std::max<uint64_t>(unique, 1));
  Grid1D cg = thread_per_item(std::max<uint64_t>(unique, static_cast<uint64_t>(n) + 1));
  launch(dev, get_kernel(dev, "sort", "csr_from_keys"), cg.blocks, cg.threads, 0,
         static_cast<const unsigned long long*>(oriented), unique, n, static_cast<uint32_t>(id_bits),
         out.targets.as<uint32_t>(), out.offsets.as<uint32_t>());
  dev.sync();
}

// parse_tsv_to_edges (crates/grafeo-cuda-core/src/csr_export.rs:246-287) on
// the device: the file's bytes are uploaded once; line breaks are found per
// 64-byte chunk and scanned into line starts; one lane parses each line; the
// edge lines are compacted; the 2E id occurrences, in order of appearance,
// are stably radix-sorted by id, so each id's first occurrence heads its run;
// the runs sorted by that first position number the ids in order of first
// appearance. `pairs` receives 2E u32 ids on the device.
void parse_tsv_on_device(Device& dev, const char* path, DeviceMemory& pairs, uint64_t& num_pairs,
                         uint32_t& num_nodes) {
  FILE* f = std::fopen(path, "rb");
  if (f == nullptr) fail(GC_ERR_EXPORT, std::string("cannot open ") + path + ": " + std::strerror(errno));
  std::vector<unsigned char> host;
  {
    std::vector<unsigned char> buf(1 << 22);
    size_t got;
    while ((got = std::fread(buf.data(), 1, buf.size(), f)) > 0) host.insert(host.end(), buf.begin(), buf.begin() + got);
    const bool failed = std::ferror(f) != 0;
    std::fclose(f);
    if (failed) fail(GC_ERR_EXPORT, std::string("cannot read ") + path);
  }
  const uint64_t size = host.size();
  if (size >= 0xFFFFFFFFull) fail_kernel("TSV files of 4 GiB or more are not supported on the device");
  constexpr uint64_t kChunk = 64;  // TSV_CHUNK of sort.hip
  const uint64_t chunks = (size + kChunk - 1) / kChunk;
  DeviceMemory bytes, counts, scan;
  bytes.allocate((chunks + 1) * kChunk);
  fill_words(dev, bytes.get(), (chunks + 1) * kChunk / 4, 0);
  if (size > 0) copy_to_device(dev, bytes.get(), host.data(), size);
  host.clear();
  host.shrink_to_fit();
  counts.allocate(sizeof(uint32_t) * std::max<uint64_t>(chunks, 1));
  scan.allocate(sizeof(uint32_t) * scan_scratch_words(std::max<uint64_t>(chunks, 1)));
  uint32_t breaks = 0;
  if (chunks > 0) {
    Grid1D cg = thread_per_item(chunks);
    launch(dev, get_kernel(dev, "sort", "tsv_count_breaks"), cg.blocks, cg.threads, 0,
           static_cast<const uint32_t*>(bytes.as<uint32_t>()) /* uint4 */, chunks, counts.as<uint32_t>());
    exclusive_scan_u32(dev, counts.as<uint32_t>(), chunks, scan.as<uint32_t>());
    copy_to_host(dev, &breaks, scan.as<uint32_t>() + scan_scratch_words(chunks) - 1, sizeof(uint32_t));
  }
  const uint32_t lines = breaks + 1;
  DeviceMemory starts;
  starts.allocate(sizeof(uint32_t) * lines);
  fill_words(dev, starts.get(), 1, 0);
  if (chunks > 0 && breaks > 0) {
    Grid1D cg = thread_per_item(chunks);
    launch(dev, get_kernel(dev, "sort", "tsv_line_starts"), cg.blocks, cg.threads, 0,
           static_cast<const uint8_t*>(bytes.as<uint8_t>()), chunks, static_cast<const uint32_t*>(counts.as<uint32_t>()),
           starts.as<uint32_t>());
  }
  counts.release();
  scan.release();
  // one lane per line
  DeviceMemory src, dst, valid;
  src.allocate(sizeof(unsigned long long) * lines);
  dst.allocate(sizeof(unsigned long long) * lines);
  valid.allocate(std::max<uint64_t>(lines, 4));
  DeviceMemory bad_utf8;
  bad_utf8.allocate(sizeof(uint32_t));
  fill_words(dev, bad_utf8.get(), 1, 0);
  Grid1D lg = thread_per_item(lines);
  launch(dev, get_kernel(dev, "sort", "tsv_parse_lines"), lg.blocks, lg.threads, 0,
         static_cast<const uint8_t*>(bytes.as<uint8_t>()), static_cast<uint32_t>(size),
         static_cast<const uint32_t*>(starts.as<uint32_t>()), lines, src.as<unsigned long long>(),
         dst.as<unsigned long long>(), valid.as<uint8_t>(), bad_utf8.as<uint32_t>());
  uint32_t bad = 0;
  copy_to_host(dev, dev.pinned, bad_utf8.get(), sizeof(uint32_t));
  std::memcpy(&bad, dev.pinned, sizeof(uint32_t));
  if (bad != 0) fail(GC_ERR_EXPORT, kInvalidUtf8);
  bytes.release();
  starts.release();
  // the edge lines, in order
  DeviceMemory src_e, dst_e;
  src_e.allocate(sizeof(unsigned long long) * lines);
  dst_e.allocate(sizeof(unsigned long long) * lines);
  const uint64_t edges =
      compact_flagged(dev, src.as<unsigned long long>(), lines, valid.as<uint8_t>(), src_e.as<unsigned long long>());
  compact_flagged(dev, dst.as<unsigned long long>(), lines, valid.as<uint8_t>(), dst_e.as<unsigned long long>());
  src.release();
  dst.release();
  valid.release();
  if (2 * edges >= 0xFFFFFFFFull) fail_kernel("too many edges for the device TSV parse");
  num_pairs = edges;
  pairs.allocate(sizeof(uint32_t) * std::max<uint64_t>(2 * edges, 2));
  if (edges == 0) {
    num_nodes = 0;
    return;
  }
  // ids numbered in order of first appearance
  const uint64_t occ = 2 * edges;
  DeviceMemory keys_a, keys_b, pos_a, pos_b, max_id;
  keys_a.allocate(sizeof(unsigned long long) * occ);
  keys_b.allocate(sizeof(unsigned long long) * occ);
  pos_a.allocate(sizeof(uint32_t) * occ);
  pos_b.allocate(sizeof(uint32_t) * occ);
  max_id.allocate(sizeof(unsigned long long));
  fill_words(dev, max_id.get(), 2, 0);
  Grid1D eg = thread_per_item(edges);
  launch(dev, get_kernel(dev, "sort", "tsv_occurrences"), eg.blocks, eg.threads, 0,
         static_cast<const unsigned long long*>(src_e.as<unsigned long long>()),
         static_cast<const unsigned long long*>(dst_e.as<unsigned long long>()), edges, keys_a.as<unsigned long long>(),
         pos_a.as<uint32_t>(), max_id.as<unsigned long long>());
  src_e.release();
  dst_e.release();
  unsigned long long largest = 0;
  copy_to_host(dev, &largest, max_id.get(), sizeof(largest));
  const auto by_id = radix_sort_pairs(dev, keys_a.as<unsigned long long>(), pos_a.as<uint32_t>(),
                                      keys_b.as<unsigned long long>(), pos_b.as<uint32_t>(), occ,
                                      std::max(1u, bits_of(largest)));
  DeviceMemory heads, run_scan;
  heads.allocate(sizeof(uint32_t) * occ);
  run_scan.allocate(sizeof(uint32_t) * (occ + scan_scratch_words(occ)));
  Grid1D og = thread_per_item(occ);
  launch(dev, get_kernel(dev, "sort", "tsv_heads"), og.blocks, og.threads, 0,
         static_cast<const unsigned long long*>(by_id.first), occ, heads.as<uint32_t>());
  hip_check(hipMemcpyAsync(run_scan.get(), heads.get(), sizeof(uint32_t) * occ, hipMemcpyDeviceToDevice, dev.stream),
            "hipMemcpyAsync");
  uint32_t* scanned = run_scan.as<uint32_t>();
  uint32_t* scan_scratch = scanned + occ;
  exclusive_scan_u32(dev, scanned, occ, scan_scratch);
  uint32_t distinct = 0;
  copy_to_host(dev, &distinct, scan_scratch + scan_scratch_words(occ) - 1, sizeof(uint32_t));
  // the runs' first positions, then the runs sorted by them (the key and
  // position buffers of the first sort are free again)
  unsigned long long* first = by_id.first == keys_a.as<unsigned long long>() ? keys_b.as<unsigned long long>()
                                                                              : keys_a.as<unsigned long long>();
  uint32_t* run_ids = by_id.second == pos_a.as<uint32_t>() ? pos_b.as<uint32_t>() : pos_a.as<uint32_t>();
  DeviceMemory first_tmp, runs_tmp, new_id;
  first_tmp.allocate(sizeof(unsigned long long) * distinct);
  runs_tmp.allocate(sizeof(uint32_t) * distinct);
  new_id.allocate(sizeof(uint32_t) * distinct);
  launch(dev, get_kernel(dev, "sort", "tsv_run_first"), og.blocks, og.threads, 0,
         static_cast<const uint32_t*>(scanned), static_cast<const uint32_t*>(heads.as<uint32_t>()),
         static_cast<const uint32_t*>(by_id.second), occ, first);
  Grid1D dg = thread_per_item(distinct);
  launch(dev, get_kernel(dev, "sort", "tsv_iota"), dg.blocks, dg.threads, 0, run_ids, static_cast<uint64_t>(distinct));
  const auto by_first = radix_sort_pairs(dev, first, run_ids, first_tmp.as<unsigned long long>(),
                                         runs_tmp.as<uint32_t>(), distinct, std::max(1u, bits_of(occ - 1)));
  launch(dev, get_kernel(dev, "sort", "tsv_rank_runs"), dg.blocks, dg.threads, 0,
         static_cast<const uint32_t*>(by_first.second), distinct, new_id.as<uint32_t>());
  launch(dev, get_kernel(dev, "sort", "tsv_write_pairs"), og.blocks, og.threads, 0,
         static_cast<const uint32_t*>(scanned), static_cast<const uint32_t*>(heads.as<uint32_t>()),
         static_cast<const uint32_t*>(by_id.second), occ, static_cast<const uint32_t*>(new_id.as<uint32_t>()),
         pairs.as<uint32_t>());
  dev.sync();
  num_nodes = distinct;
}

}  // namespace gc

using gc::guard;

extern "C" {

gc_status gc_csr_from_edge_pairs_device(gc_device* dev_handle, const uint32_t* pairs, size_t num_pairs,
                                        uint32_t num_nodes, int32_t symmetric, gc_csr* out) {
  return guard([&] {
    if (out == nullptr) gc::fail_kernel("out is null");
    std::memset(out, 0, sizeof(*out));
    if (dev_handle == nullptr) gc::fail_kernel("device is null");
    if (pairs == nullptr && num_pairs > 0) gc::fail_kernel("pairs is null");
    auto* dev = reinterpret_cast<gc::Device*>(dev_handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    gc::DeviceCsr csr;
    gc::build_csr_on_device(*dev, pairs, num_pairs, num_nodes, symmetric != 0, csr);
    auto* offsets = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (size_t(num_nodes) + 1)));
    auto* targets = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * std::max<size_t>(1, csr.m)));
    if (!offsets || !targets) {
      std::free(offsets);
      std::free(targets);
      throw std::bad_alloc();
    }
    gc::copy_to_host(*dev, offsets, csr.offsets.get(), sizeof(uint32_t) * (size_t(num_nodes) + 1));
    if (csr.m > 0) gc::copy_to_host(*dev, targets, csr.targets.get(), sizeof(uint32_t) * csr.m);
    out->offsets = offsets;
    out->targets = targets;
    out->num_nodes = num_nodes;
    out->num_edges = csr.m;
  });
}

gc_status gc_csr_from_edge_pairs_oriented_device(gc_device* dev_handle, const uint32_t* pairs, size_t num_pairs,
                                                 uint32_t num_nodes, gc_csr* out) {
  return guard([&] {
    if (out == nullptr) gc::fail_kernel("out is null");
    std::memset(out, 0, sizeof(*out));
    if (dev_handle == nullptr) gc::fail_kernel("device is null");
    if (pairs == nullptr && num_pairs > 0) gc::fail_kernel("pairs is null");
    auto* dev = reinterpret_cast<gc::Device*>(dev_handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    gc::DeviceCsr csr;
    gc::build_oriented_csr_on_device(*dev, pairs, num_pairs, num_nodes, csr);
    auto* offsets = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (size_t(num_nodes) + 1)));
    auto* targets = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * std::max<size_t>(1, csr.m)));
    if (!offsets || !targets) {
      std::free(offsets);
      std::free(targets);
      throw std::bad_alloc();
    }
    gc::copy_to_host(*dev, offsets, csr.offsets.get(), sizeof(uint32_t) * (size_t(num_nodes) + 1));
    if (csr.m > 0) gc::copy_to_host(*dev, targets, csr.targets.get(), sizeof(uint32_t) * csr.m);
    out->offsets = offsets;
    out->targets = targets;
    out->num_nodes = num_nodes;
    out->num_edges = csr.m;
  });
}

gc_status gc_parse_tsv_to_edges_device(gc_device* dev_handle, const char* path, uint32_t** pairs,
                                       size_t* num_pairs, uint32_t* num_nodes) {
  return guard([&] {
    if (path == nullptr || pairs == nullptr || num_pairs == nullptr || num_nodes == nullptr) {
      gc::fail_kernel("null argument");
    }
    *pairs = nullptr;
    *num_pairs = 0;
    *num_nodes = 0;
    if (dev_handle == nullptr) gc::fail_kernel("device is null");
    auto* dev = reinterpret_cast<gc::Device*>(dev_handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    gc::DeviceMemory d_pairs;
    uint64_t count = 0;
    uint32_t nodes = 0;
    gc::parse_tsv_on_device(*dev, path, d_pairs, count, nodes);
    auto* out = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * std::max<uint64_t>(2 * count, 1)));
    if (out == nullptr) throw std::bad_alloc();
    if (count > 0) gc::copy_to_host(*dev, out, d_pairs.get(), sizeof(uint32_t) * 2 * count);
    *pairs = out;
    *num_pairs = count;
    *num_nodes = nodes;
  });
}

gc_status gc_graph_from_tsv(gc_device* dev_handle, const char* path, int32_t symmetric, gc_graph** out,
                            size_t* num_pairs) {
  return guard([&] {
    if (out == nullptr || path == nullptr) gc::fail_kernel("null argument");
    *out = nullptr;
    if (dev_handle == nullptr) gc::fail_kernel("device is null");
    auto* dev = reinterpret_cast<gc::Device*>(dev_handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    gc::DeviceMemory d_pairs;
    uint64_t count = 0;
    uint32_t nodes = 0;
    gc::parse_tsv_on_device(*dev, path, d_pairs, count, nodes);
    gc::DeviceCsr csr;
    gc::build_csr_on_device(*dev, nullptr, count, nodes, symmetric != 0, csr, d_pairs.as<uint32_t>());
    d_pairs.release();
    std::unique_ptr<gc::Graph> g(new gc::Graph());
    g->n = csr.n;
    g->m = csr.m;
    g->offsets = std::move(csr.offsets);
    g->targets = std::move(csr.targets);
    g->host_offsets.resize(size_t(g->n) + 1);
    gc::copy_to_host(*dev, g->host_offsets.data(), g->offsets.get(), sizeof(uint32_t) * g->host_offsets.size());
    g->dev = reinterpret_cast<gc::Device*>(gc_device_retain(dev_handle));
    if (num_pairs) *num_pairs = count;
    *out = reinterpret_cast<gc_graph*>(g.release());
  });
}

gc_status gc_graph_from_edge_pairs(gc_device* dev_handle, const uint32_t* pairs, size_t num_pairs,
                                   uint32_t num_nodes, int32_t symmetric, gc_graph** out) {
  return guard([&] {
    if (out == nullptr) gc::fail_kernel("out is null");
    *out = nullptr;
    if (dev_handle == nullptr) gc::fail_kernel("device is null");
    if (pairs == nullptr && num_pairs > 0) gc::fail_kernel("pairs is null");
    auto* dev = reinterpret_cast<gc::Device*>(dev_handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    gc::DeviceCsr csr;
    gc::build_csr_on_device(*dev, pairs, num_pairs, num_nodes, symmetric != 0, csr);
    std::unique_ptr<gc::Graph> g(new gc::Graph());
    g->n = csr.n;
    g->m = csr.m;
    g->offsets = std::move(csr.offsets);
    g->targets = std::move(csr.targets);
    // the offsets on the host, for planning (targets stay on the device)
    g->host_offsets.resize(size_t(g->n) + 1);
    gc::copy_to_host(*dev, g->host_offsets.data(), g->offsets.get(), sizeof(uint32_t) * g->host_offsets.size());
    g->dev = reinterpret_cast<gc::Device*>(gc_device_retain(dev_handle));
    *out = reinterpret_cast<gc_graph*>(g.release());
  });
}

}  // extern "C"
