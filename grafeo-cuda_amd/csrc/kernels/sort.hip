// Device CSR construction: a stable LSD radix sort of 64-bit keys and the
// kernels that turn sorted (row, target) keys into a CSR.
//
// Radix sort, one pass per 8-bit digit (the caller runs ceil(bits / 8)
// passes over the significant bits only):
//   rs_histogram  each workgroup counts the digits of its tile of RS_TILE keys
//                 and writes them digit-major, hist[digit * blocks + block],
//                 so that one exclusive scan of the whole array gives every
//                 (digit, tile) pair its first output slot, tiles in order;
//   rs_scatter    each workgroup ranks its keys stably (in input order) and
//                 writes key and value to hist[digit * blocks + block] + rank.
// Stable ranking: the tile is walked in steps of 256 keys, one per lane, in
// input order. Lanes of a wave holding the same digit are found with 8
// ballots (one per digit bit); the rank is the number of such lanes below,
// plus the same digit's count in the lower waves of the step, plus the
// digit's running count over the earlier steps of the tile.
#include "device.hpp"

using namespace gcd;

#define RS_THREADS 256
#define RS_ITEMS 16
#define RS_TILE (RS_THREADS * RS_ITEMS)
#define RS_DIGITS 256
#define RS_WAVES (RS_THREADS / GC_WAVE)

extern "C" __global__ __launch_bounds__(RS_THREADS) void rs_histogram(
    const unsigned long long* __restrict__ keys, uint64_t n, uint32_t shift,
    uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_count[RS_DIGITS];
  s_count[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * RS_TILE;
  for (uint32_t k = 0; k < RS_ITEMS; ++k) {
    const uint64_t i = base + static_cast<uint64_t>(k) * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&s_count[(keys[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[static_cast<uint64_t>(threadIdx.x) * gridDim.x + blockIdx.x] = s_count[threadIdx.x];
}

// Lanes of the wave whose digit equals this lane's (active lanes only).
__device__ __forceinline__ unsigned long long same_digit_lanes(uint32_t digit, bool active) {
  unsigned long long peers = __ballot(active);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const unsigned long long set = __ballot(active && ((digit >> b) & 1u));
    peers &= ((digit >> b) & 1u) ? set : ~set;
  }
  return peers;
}

extern "C" __global__ __launch_bounds__(RS_THREADS) void rs_scatter(
    const unsigned long long* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint64_t n, uint32_t shift, const uint32_t* __restrict__ scanned,
    unsigned long long* __restrict__ keys_out, uint32_t* __restrict__ vals_out) {
  __shared__ uint32_t s_base[RS_DIGITS];               // output slot of the digit's next key
  __shared__ uint32_t s_wave_count[RS_WAVES][RS_DIGITS];
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / GC_WAVE;
  s_base[threadIdx.x] = scanned[static_cast<uint64_t>(threadIdx.x) * gridDim.x + blockIdx.x];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * RS_TILE;
  for (uint32_t k = 0; k < RS_ITEMS; ++k) {
    for (uint32_t w = 0; w < RS_WAVES; ++w) s_wave_count[w][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t i = base + static_cast<uint64_t>(k) * RS_THREADS + threadIdx.x;
    const bool active = i < n;
    unsigned long long key = 0;
    uint32_t val = 0, digit = 0;
    if (active) {
      key = keys_in[i];
      if (vals_in != nullptr) val = vals_in[i];
      digit = static_cast<uint32_t>((key >> shift) & 0xFFu);
    }
    const unsigned long long peers = same_digit_lanes(digit, active);
    const uint32_t below = lanes_below(peers);
    // the lowest lane of each digit group publishes the group's size
    if (active && below == 0) s_wave_count[wave][digit] = static_cast<uint32_t>(__popcll(peers));
    __syncthreads();
    if (active) {
      uint32_t rank = s_base[digit] + below;
      for (uint32_t w = 0; w < wave; ++w) rank += s_wave_count[w][digit];
      keys_out[rank] = key;
      if (vals_in != nullptr) vals_out[rank] = val;
    }
    __syncthreads();
    // advance each digit's base by its count in this step
    uint32_t step = 0;
    for (uint32_t w = 0; w < RS_WAVES; ++w) step += s_wave_count[w][threadIdx.x];
    s_base[threadIdx.x] += step;
    __syncthreads();
  }
}

// ---- CSR from edge pairs --------------------------------------------------------

// keys from (src, dst) pairs: src << id_bits | dst, and, when symmetric,
// dst << id_bits | src for every non-loop pair (slot 2i and 2i+1; a loop's
// second slot repeats its first, removed by the dedup). `bad` counts pairs
// with an endpoint >= n.
extern "C" __global__ __launch_bounds__(256) void csr_make_keys(
    const uint32_t* __restrict__ pairs, uint64_t num_pairs, uint32_t n, uint32_t id_bits,
    uint32_t symmetric, unsigned long long* __restrict__ keys, uint32_t* __restrict__ bad) {
  uint32_t local_bad = 0;
  for (uint64_t p = global_thread(); p < num_pairs; p += grid_threads()) {
    const uint32_t s = pairs[2 * p], d = pairs[2 * p + 1];
    if (s >= n || d >= n) ++local_bad;
    const unsigned long long fwd = (static_cast<unsigned long long>(s) << id_bits) | d;
    if (symmetric) {
      keys[2 * p] = fwd;
      keys[2 * p + 1] = (static_cast<unsigned long long>(d) << id_bits) | s;
    } else {
      keys[p] = fwd;
    }
  }
  local_bad = wave_sum(local_bad);
  if (lane_id() == 0 && local_bad) atomicAdd(bad, local_bad);
}

// ---- oriented CSR from edge pairs -------------------------------------------------
// (the steps of from_edge_pairs_oriented, crates/grafeo-cuda-core/src/csr_export.rs:299-350)

// Canonical undirected keys min << id_bits | max, one per pair (a loop keeps
// its key min == max and is dropped by csr_unique_edges); `bad` counts pairs
// with an endpoint >= n.
extern "C" __global__ __launch_bounds__(256) void csr_undirected_keys(
    const uint32_t* __restrict__ pairs, uint64_t num_pairs, uint32_t n, uint32_t id_bits,
    unsigned long long* __restrict__ keys, uint32_t* __restrict__ bad) {
  uint32_t local_bad = 0;
  for (uint64_t p = global_thread(); p < num_pairs; p += grid_threads()) {
    const uint32_t s = pairs[2 * p], d = pairs[2 * p + 1];
    if (s >= n || d >= n) ++local_bad;
    const unsigned long long lo = min(s, d), hi = max(s, d);
    keys[p] = (lo << id_bits) | hi;
  }
  local_bad = wave_sum(local_bad);
  if (lane_id() == 0 && local_bad) atomicAdd(bad, local_bad);
}

// flags[i] = 1 for the first copy of each sorted undirected key that is not a
// loop (its two ids differ).
extern "C" __global__ __launch_bounds__(256) void csr_unique_edges(
    const unsigned long long* __restrict__ keys, uint64_t count, uint32_t id_bits, uint8_t* __restrict__ flags) {
  const unsigned long long mask = (1ull << id_bits) - 1ull;
  for (uint64_t i = global_thread(); i < count; i += grid_threads()) {
    const unsigned long long k = keys[i];
    const bool first = i == 0 || k != keys[i - 1];
    flags[i] = (first && (k >> id_bits) != (k & mask)) ? 1 : 0;
  }
}

// degree[x] += the unique undirected edges at x (degree cleared by the
// caller). The keys are sorted, so the lower ids come in runs: each run adds
// its length with one atomic (a hub's run would otherwise serialize on one
// address); the higher ids add one each.
extern "C" __global__ __launch_bounds__(256) void csr_edge_degrees(
    const unsigned long long* __restrict__ keys, uint64_t count, uint32_t id_bits,
    uint32_t* __restrict__ degree) {
  const unsigned long long mask = (1ull << id_bits) - 1ull;
  const uint32_t lane = lane_id();
  const uint64_t waves = grid_threads() / GC_WAVE;
  for (uint64_t base = (global_thread() / GC_WAVE) * GC_WAVE; base < count; base += waves * GC_WAVE) {
    const uint64_t i = base + lane;
    const bool valid = i < count;
    const unsigned long long k = valid ? keys[i] : 0ull;
    const uint32_t a = static_cast<uint32_t>(k >> id_bits);
    const uint32_t b = static_cast<uint32_t>(k & mask);
    const uint32_t prev_a = __shfl_up(a, 1);
    // heads of the runs of equal lower ids within the wave
    const bool head = valid && (lane == 0 || prev_a != a);
    const unsigned long long heads = __ballot(head);
    if (head) {
      // the run ends at the next head or at the last valid lane
      const unsigned long long later = lane == GC_WAVE - 1 ? 0ull : heads >> (lane + 1);
      const uint64_t valid_lanes = count - base < GC_WAVE ? count - base : GC_WAVE;
      const uint32_t end = later != 0 ? lane + 1 + static_cast<uint32_t>(__ffsll(static_cast<long long>(later)) - 1)
                                      : static_cast<uint32_t>(valid_lanes);
      atomicAdd(&degree[a], end - lane);
    }
    if (valid) atomicAdd(&degree[b], 1u);
  }
}

// Each unique undirected edge (a, b), a < b, oriented from the endpoint of
// lower degree to the higher, ties from the lower id: a << id_bits | b or
// b << id_bits | a.
extern "C" __global__ __launch_bounds__(256) void csr_orient_keys(
    const unsigned long long* __restrict__ keys, uint64_t count, uint32_t id_bits,
    const uint32_t* __restrict__ degree, unsigned long long* __restrict__ out) {
  const unsigned long long mask = (1ull << id_bits) - 1ull;
  for (uint64_t i = global_thread(); i < count; i += grid_threads()) {
    const unsigned long long k = keys[i];
    const unsigned long long a = k >> id_bits, b = k & mask;
    const uint32_t da = degree[a], db = degree[b];
    out[i] = (da < db || (da == db && a < b)) ? k : ((b << id_bits) | a);
  }
}

// flags[i] = 1 when sorted key i differs from key i-1 (first copies of each key).
extern "C" __global__ __launch_bounds__(256) void csr_first_copies(
    const unsigned long long* __restrict__ keys, uint64_t count, uint8_t* __restrict__ flags) {
  for (uint64_t i = global_thread(); i < count; i += grid_threads())
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// Compacts the flagged (first-copy) sorted keys in order; the per-tile counts
// (tiles of 4096 entries) are scanned by the caller with compact_count and
// exclusive_scan_small of util.hip, which use the same tile.
extern "C" __global__ __launch_bounds__(256) void csr_compact_keys(
    const unsigned long long* __restrict__ keys, const uint8_t* __restrict__ flags, uint64_t count,
    const uint32_t* __restrict__ tile_offsets, unsigned long long* __restrict__ out) {
  __shared__ uint32_t s_wave[256 / GC_WAVE];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * 4096 + static_cast<uint64_t>(threadIdx.x) * 16;
  uint32_t c = 0;
  for (uint32_t k = 0; k < 16; ++k) {
    const uint64_t i = base + k;
    if (i < count) c += flags[i] != 0;
  }
  const uint32_t incl = wave_inclusive_scan(c);
  const uint32_t wave = threadIdx.x / GC_WAVE;
  if (lane_id() == GC_WAVE - 1) s_wave[wave] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wave; ++w) before += s_wave[w];
  uint32_t pos = tile_offsets[blockIdx.x] + before + incl - c;
  for (uint32_t k = 0; k < 16; ++k) {
    const uint64_t i = base + k;
    if (i >= count) break;
    if (flags[i] != 0) out[pos++] = keys[i];
  }
}

// CSR arrays from the unique sorted keys (row << id_bits | target):
// targets[i] = low id_bits of key i; offsets[u] = number of keys whose row is
// below u (a binary search per row), for u = 0..n.
extern "C" __global__ __launch_bounds__(256) void csr_from_keys(
    const unsigned long long* __restrict__ keys, uint64_t count, uint32_t n, uint32_t id_bits,
    uint32_t* __restrict__ targets, uint32_t* __restrict__ offsets) {
  const unsigned long long mask = (1ull << id_bits) - 1ull;
  const uint64_t total = count > static_cast<uint64_t>(n) + 1 ? count : static_cast<uint64_t>(n) + 1;
  for (uint64_t i = global_thread(); i < total; i += grid_threads()) {
    if (i < count) targets[i] = static_cast<uint32_t>(keys[i] & mask);
    if (i <= n) {
      const unsigned long long first = static_cast<unsigned long long>(i) << id_bits;
      uint64_t lo = 0, hi = count;
      while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        if (keys[mid] < first) lo = mid + 1;
        else hi = mid;
      }
      offsets[i] = static_cast<uint32_t>(lo);
    }
  }
}

// ---- TSV edge lists on the device -----------------------------------------------
// The parse of parse_tsv_to_edges (crates/grafeo-cuda-core/src/csr_export.rs:
// 246-287): lines split at '\n', trimmed of ASCII whitespace; blank lines and
// lines starting with '#' or '%' skipped; the first two whitespace-separated
// fields parsed as u64 the way Rust's u64::from_str does (an optional '+',
// then decimal digits, no overflow), the line skipped when either fails. The
// file bytes sit in a buffer padded with zeros to a multiple of TSV_CHUNK
// plus TSV_CHUNK, so every chunk's loads stay inside it.

#define TSV_CHUNK 64

__device__ __forceinline__ uint32_t zero_bytes(uint32_t x) {
  // bytes of x that are zero, counted exactly (no carries between bytes)
  uint32_t k = 0;
#pragma unroll
  for (int s = 0; s < 32; s += 8) k += ((x >> s) & 0xFFu) == 0u;
  return k;
}

// counts[c] = line breaks in bytes [c * TSV_CHUNK, (c + 1) * TSV_CHUNK).
extern "C" __global__ __launch_bounds__(256) void tsv_count_breaks(const uint4* __restrict__ bytes, uint64_t chunks,
                                                                   uint32_t* __restrict__ counts) {
  for (uint64_t c = global_thread(); c < chunks; c += grid_threads()) {
    uint32_t k = 0;
#pragma unroll
    for (int q = 0; q < TSV_CHUNK / 16; ++q) {
      const uint4 w = bytes[c * (TSV_CHUNK / 16) + q];
      k += zero_bytes(w.x ^ 0x0A0A0A0Au) + zero_bytes(w.y ^ 0x0A0A0A0Au) + zero_bytes(w.z ^ 0x0A0A0A0Au) +
           zero_bytes(w.w ^ 0x0A0A0A0Au);
    }
    counts[c] = k;
  }
}

// starts[first[c] + j + 1] = the byte after the j-th line break of chunk c
// (first = the exclusive scan of the counts; starts[0] = 0 is set apart).
extern "C" __global__ __launch_bounds__(256) void tsv_line_starts(const uint8_t* __restrict__ bytes, uint64_t chunks,
                                                                  const uint32_t* __restrict__ first,
                                                                  uint32_t* __restrict__ starts) {
  for (uint64_t c = global_thread(); c < chunks; c += grid_threads()) {
    uint32_t at = first[c];
    const uint64_t b = c * TSV_CHUNK;
    for (uint32_t i = 0; i < TSV_CHUNK; ++i)
      if (bytes[b + i] == '\n') starts[++at] = static_cast<uint32_t>(b + i + 1);
  }
}

// Byte length of the White_Space code point (Unicode's property, which Rust's
// str::trim and split_whitespace use) starting at bytes[p], p < e, or 0:
// U+0009..000D and U+0020 (one byte), U+0085 and U+00A0 (two), U+1680,
// U+2000..200A, U+2028, U+2029, U+202F, U+205F and U+3000 (three).
__device__ __forceinline__ uint32_t tsv_space_at(const uint8_t* bytes, uint32_t p, uint32_t e) {
  const uint32_t c = bytes[p];
  if (c == ' ' || (c >= 0x09 && c <= 0x0D)) return 1;
  if (c < 0xC2 || c > 0xE3 || p + 1 >= e) return 0;
  const uint32_t c1 = bytes[p + 1];
  if (c == 0xC2) return c1 == 0x85 || c1 == 0xA0 ? 2 : 0;
  if (p + 2 >= e) return 0;
  const uint32_t c2 = bytes[p + 2];
  if (c == 0xE1) return c1 == 0x9A && c2 == 0x80 ? 3 : 0;
  if (c == 0xE2) {
    if (c1 == 0x80) return (c2 <= 0x8A && c2 >= 0x80) || c2 == 0xA8 || c2 == 0xA9 || c2 == 0xAF ? 3 : 0;
    return c1 == 0x81 && c2 == 0x9F ? 3 : 0;
  }
  if (c == 0xE3) return c1 == 0x80 && c2 == 0x80 ? 3 : 0;
  return 0;
}

// Byte length of a White_Space code point ending at bytes[e - 1], b < e, or 0
// (the line is valid UTF-8, so a matching suffix is that code point).
__device__ __forceinline__ uint32_t tsv_space_before(const uint8_t* bytes, uint32_t b, uint32_t e) {
  if (tsv_space_at(bytes, e - 1, e) == 1) return 1;
  if (e - b >= 2 && tsv_space_at(bytes, e - 2, e) == 2) return 2;
  if (e - b >= 3 && tsv_space_at(bytes, e - 3, e) == 3) return 3;
  return 0;
}

// Whether bytes [b, e) are well-formed UTF-8 (no overlong forms, no
// surrogates, nothing above U+10FFFF), the check of Rust's BufRead::lines.
__device__ __forceinline__ bool tsv_utf8_ok(const uint8_t* bytes, uint32_t b, uint32_t e) {
  for (uint32_t i = b; i < e;) {
    const uint32_t c = bytes[i];
    if (c < 0x80) {
      ++i;
      continue;
    }
    uint32_t len, lo = 0x80, hi = 0xBF;
    if (c >= 0xC2 && c <= 0xDF) len = 2;
    else if (c >= 0xE0 && c <= 0xEF) {
      len = 3;
      if (c == 0xE0) lo = 0xA0;
      if (c == 0xED) hi = 0x9F;
    } else if (c >= 0xF0 && c <= 0xF4) {
      len = 4;
      if (c == 0xF0) lo = 0x90;
      if (c == 0xF4) hi = 0x8F;
    } else {
      return false;
    }
    if (e - i < len) return false;
    const uint32_t c1 = bytes[i + 1];
    if (c1 < lo || c1 > hi) return false;
    for (uint32_t k = 2; k < len; ++k)
      if ((bytes[i + k] & 0xC0u) != 0x80u) return false;
    i += len;
  }
  return true;
}

__device__ __forceinline__ bool tsv_parse_u64(const uint8_t* bytes, uint32_t b, uint32_t e, unsigned long long* out) {
  if (b < e && bytes[b] == '+') ++b;
  if (b == e) return false;
  unsigned long long v = 0;
  for (uint32_t p = b; p < e; ++p) {
    const uint32_t c = bytes[p];
    if (c < '0' || c > '9') return false;
    const unsigned long long digit = c - '0';
    if (v > (0xFFFFFFFFFFFFFFFFull - digit) / 10ull) return false;
    v = v * 10ull + digit;
  }
  *out = v;
  return true;
}

// Line i = bytes [starts[i], starts[i + 1] - 1), the last one up to `size`:
// its two ids, and whether it holds an edge. As Rust's lines / trim /
// split_whitespace: the line is trimmed of White_Space at both ends, fields
// are separated by White_Space runs, and a line that is not valid UTF-8 sets
// *bad_utf8 (the whole parse then fails, as BufRead::lines does).
extern "C" __global__ __launch_bounds__(256) void tsv_parse_lines(
    const uint8_t* __restrict__ bytes, uint32_t size, const uint32_t* __restrict__ starts, uint32_t lines,
    unsigned long long* __restrict__ src, unsigned long long* __restrict__ dst, uint8_t* __restrict__ valid,
    uint32_t* __restrict__ bad_utf8) {
  for (uint64_t li = global_thread(); li < lines; li += grid_threads()) {
    uint32_t b = starts[li];
    uint32_t e = li + 1 < lines ? starts[li + 1] - 1 : size;
    if (!tsv_utf8_ok(bytes, b, e)) atomicOr(bad_utf8, 1u);
    uint32_t k;
    while (b < e && (k = tsv_space_at(bytes, b, e)) != 0) b += k;
    while (e > b && (k = tsv_space_before(bytes, b, e)) != 0) e -= k;
    bool ok = b < e && bytes[b] != '#' && bytes[b] != '%';
    unsigned long long s = 0, d = 0;
    if (ok) {
      uint32_t p = b;
      while (p < e && tsv_space_at(bytes, p, e) == 0) ++p;
      const uint32_t f1e = p;
      while (p < e && (k = tsv_space_at(bytes, p, e)) != 0) p += k;
      const uint32_t f2b = p;
      while (p < e && tsv_space_at(bytes, p, e) == 0) ++p;
      ok = tsv_parse_u64(bytes, b, f1e, &s) && tsv_parse_u64(bytes, f2b, p, &d);
    }
    src[li] = s;
    dst[li] = d;
    valid[li] = ok ? 1 : 0;
  }
}

// The id occurrences in order of appearance: keys[2e] = src[e], keys[2e + 1] =
// dst[e], pos[p] = p; and the largest id (for the sort's bit count).
extern "C" __global__ __launch_bounds__(256) void tsv_occurrences(
    const unsigned long long* __restrict__ src, const unsigned long long* __restrict__ dst, uint64_t edges,
    unsigned long long* __restrict__ keys, uint32_t* __restrict__ pos, unsigned long long* __restrict__ max_id) {
  unsigned long long m = 0;
  for (uint64_t e = global_thread(); e < edges; e += grid_threads()) {
    const unsigned long long s = src[e], d = dst[e];
    keys[2 * e] = s;
    keys[2 * e + 1] = d;
    pos[2 * e] = static_cast<uint32_t>(2 * e);
    pos[2 * e + 1] = static_cast<uint32_t>(2 * e + 1);
    m = max(m, max(s, d));
  }
#pragma unroll
  for (int o = GC_WAVE / 2; o > 0; o >>= 1) m = max(m, static_cast<unsigned long long>(__shfl_xor(m, o)));
  if (lane_id() == 0 && m != 0) atomicMax(max_id, m);
}

// heads[i] = 1 where the sorted ids change (the first occurrence of an id, the
// earliest in the stable sort).
extern "C" __global__ __launch_bounds__(256) void tsv_heads(const unsigned long long* __restrict__ keys,
                                                            uint64_t count, uint32_t* __restrict__ heads) {
  for (uint64_t i = global_thread(); i < count; i += grid_threads())
    heads[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// With run[i] = (exclusive scan of heads)[i] + heads[i] - 1, the id run of
// sorted occurrence i: first[run] = the position of each run's head.
extern "C" __global__ __launch_bounds__(256) void tsv_run_first(const uint32_t* __restrict__ scanned,
                                                                const uint32_t* __restrict__ heads_in,
                                                                const uint32_t* __restrict__ pos, uint64_t count,
                                                                unsigned long long* __restrict__ first) {
  for (uint64_t i = global_thread(); i < count; i += grid_threads())
    if (heads_in[i]) first[scanned[i]] = pos[i];
}

// new_id[run[r]] = r over the runs sorted by first position.
extern "C" __global__ __launch_bounds__(256) void tsv_rank_runs(const uint32_t* __restrict__ runs_by_first,
                                                                uint32_t count, uint32_t* __restrict__ new_id) {
  for (uint64_t r = global_thread(); r < count; r += grid_threads()) new_id[runs_by_first[r]] = static_cast<uint32_t>(r);
}

// pairs[pos[i]] = new_id[run of i]: the pairs, in line order, with ids
// numbered in order of first appearance.
extern "C" __global__ __launch_bounds__(256) void tsv_write_pairs(const uint32_t* __restrict__ scanned,
                                                                  const uint32_t* __restrict__ heads_in,
                                                                  const uint32_t* __restrict__ pos, uint64_t count,
                                                                  const uint32_t* __restrict__ new_id,
                                                                  uint32_t* __restrict__ pairs) {
  for (uint64_t i = global_thread(); i < count; i += grid_threads())
    pairs[pos[i]] = new_id[scanned[i] + heads_in[i] - 1];
}

// out[i] = i.
extern "C" __global__ __launch_bounds__(256) void tsv_iota(uint32_t* __restrict__ out, uint64_t count) {
  for (uint64_t i = global_thread(); i < count; i += grid_threads()) out[i] = static_cast<uint32_t>(i);
}
