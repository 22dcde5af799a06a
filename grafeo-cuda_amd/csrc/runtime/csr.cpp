// CSR model: validation, edge-pair builders and the TSV edge-list parser.
//
// Follows crates/grafeo-cuda-core/src/csr_export.rs:
//   validate_csr              :71-120   (same checks, same order, same messages)
//   CsrArrays::validate       :37-54
//   from_edge_pairs           :202-239  (symmetric adds the reverse of each non-loop pair; sort; dedup)
//   from_edge_pairs_oriented  :299-350  (drop loops; dedup {min,max}; orient low degree -> high, ties low id)
//   parse_tsv_to_edges        :246-287  (skip blank/#/% lines and malformed lines; remap ids by first appearance)
// The sorts are a parallel LSD radix sort over packed (src << 32 | dst) keys,
// which orders pairs exactly as the reference's tuple sort does.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <unordered_map>

#include "common.hpp"

namespace gc {

void validate_csr(const uint32_t* offsets, size_t num_offsets, const uint32_t* targets,
                  size_t num_targets) {
  if (num_offsets == 0) {
    fail_csr("offsets must have num_nodes + 1 entries, got an empty array");
  }
  if (num_offsets > 0xFFFFFFFFull || num_targets > 0xFFFFFFFFull) {
    fail_csr("graph too large for u32 indices: " + std::to_string(num_offsets) + " offsets, " +
             std::to_string(num_targets) + " targets");
  }
  if (offsets[0] != 0) fail_csr("offsets[0] must be 0, got " + std::to_string(offsets[0]));
  const size_t n = num_offsets - 1;
  if (static_cast<size_t>(offsets[n]) != num_targets) {
    fail_csr("offsets[" + std::to_string(n) + "] is " + std::to_string(offsets[n]) +
             " but targets has " + std::to_string(num_targets) + " entries");
  }
  // The first failing node in id order decides the message, as in the
  // sequential reference scan; chunks record their first failure.
  size_t first_bad = SIZE_MAX;
  std::string first_message;
  std::mutex mu;
  parallel_for(
      n,
      [&](uint64_t b, uint64_t e) {
        for (uint64_t node = b; node < e; ++node) {
          uint32_t start = offsets[node], end = offsets[node + 1];
          std::string msg;
          if (end < start) {
            msg = "offsets decrease at node " + std::to_string(node) + ": " + std::to_string(start) +
                  " then " + std::to_string(end);
          } else if (end > num_targets) {
            // Only reachable when a later offset decreases; report that node.
            continue;
          } else {
            for (uint32_t p = start; p < end; ++p) {
              uint32_t target = targets[p];
              if (target >= n) {
                msg = "node " + std::to_string(node) + " has neighbor " + std::to_string(target) +
                      ", but there are only " + std::to_string(n) + " nodes";
                break;
              }
              if (p > start && targets[p - 1] >= target) {
                msg = "neighbor list of node " + std::to_string(node) +
                      " is not strictly increasing (" + std::to_string(targets[p - 1]) + " then " +
                      std::to_string(target) + ")";
                break;
              }
            }
          }
          if (!msg.empty()) {
            std::lock_guard<std::mutex> lock(mu);
            if (node < first_bad) {
              first_bad = node;
              first_message = msg;
            }
            return;
          }
        }
      },
      1u << 14);
  if (first_bad != SIZE_MAX) fail_csr(first_message);
  // A node whose end exceeded num_targets is always followed by a decrease
  // (offsets[n] == num_targets), which the scan above reports.
}

// ------------------------------------------------------------ radix sort --

namespace {

// Stable LSD radix sort of `keys` on the low `key_bits` bits.
void radix_sort_u64(std::vector<uint64_t>& keys, unsigned key_bits) {
  const size_t n = keys.size();
  if (n < (1u << 16)) {
    std::sort(keys.begin(), keys.end());
    return;
  }
  constexpr unsigned kDigit = 11;
  constexpr size_t kBuckets = size_t(1) << kDigit;
  std::vector<uint64_t> tmp(n);
  unsigned t = host_threads();
  const size_t chunks = std::max<size_t>(1, std::min<size_t>(t, n / 4096));
  std::vector<size_t> hist(chunks * kBuckets);
  uint64_t* src = keys.data();
  uint64_t* dst = tmp.data();
  unsigned passes = (key_bits + kDigit - 1) / kDigit;
  if (passes == 0) passes = 1;
  for (unsigned pass = 0; pass < passes; ++pass) {
    const unsigned shift = pass * kDigit;
    std::fill(hist.begin(), hist.end(), 0);
    std::vector<std::thread> pool;
    for (size_t c = 0; c < chunks; ++c) {
      pool.emplace_back([&, c] {
        size_t b = n * c / chunks, e = n * (c + 1) / chunks;
        size_t* h = &hist[c * kBuckets];
        for (size_t i = b; i < e; ++i) ++h[(src[i] >> shift) & (kBuckets - 1)];
      });
    }
    for (auto& th : pool) th.join();
    pool.clear();
    size_t running = 0;
    for (size_t d = 0; d < kBuckets; ++d) {
      for (size_t c = 0; c < chunks; ++c) {
        size_t v = hist[c * kBuckets + d];
        hist[c * kBuckets + d] = running;
        running += v;
      }
    }
    for (size_t c = 0; c < chunks; ++c) {
      pool.emplace_back([&, c] {
        size_t b = n * c / chunks, e = n * (c + 1) / chunks;
        size_t* h = &hist[c * kBuckets];
        for (size_t i = b; i < e; ++i) {
          uint64_t k = src[i];
          dst[h[(k >> shift) & (kBuckets - 1)]++] = k;
        }
      });
    }
    for (auto& th : pool) th.join();
    std::swap(src, dst);
  }
  if (src != keys.data()) std::memcpy(keys.data(), src, n * sizeof(uint64_t));
}

unsigned bits_for(uint64_t max_value) {
  unsigned b = 0;
  while (b < 64 && (max_value >> b) != 0) ++b;
  return b;
}

// Sorts and deduplicates packed keys (src << 32 | dst) whose ids are < n.
void sort_unique(std::vector<uint64_t>& keys, uint32_t n) {
  unsigned id_bits = std::max(1u, bits_for(n == 0 ? 0 : n - 1));
  // Compact (src, dst) into 2*id_bits bits, sort, expand back.
  if (id_bits < 32) {
    for (auto& k : keys) k = ((k >> 32) << id_bits) | (k & 0xFFFFFFFFull);
  }
  radix_sort_u64(keys, 2 * id_bits);
  if (id_bits < 32) {
    const uint64_t mask = (uint64_t(1) << id_bits) - 1;
    for (auto& k : keys) k = ((k >> id_bits) << 32) | (k & mask);
  }
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
}

// CSR from sorted unique packed keys.
void keys_to_csr(const std::vector<uint64_t>& keys, uint32_t n, gc_csr* out) {
  if (keys.size() > 0xFFFFFFFFull) {
    fail_csr("graph too large for u32 indices: " + std::to_string(keys.size()) + " targets");
  }
  auto* offsets = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (size_t(n) + 1)));
  auto* targets = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * std::max<size_t>(1, keys.size())));
  if (!offsets || !targets) {
    std::free(offsets);
    std::free(targets);
    throw std::bad_alloc();
  }
  std::fill(offsets, offsets + size_t(n) + 1, 0u);
  for (uint64_t k : keys) ++offsets[(k >> 32) + 1];
  for (size_t i = 1; i <= n; ++i) offsets[i] += offsets[i - 1];
  parallel_for(keys.size(), [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) targets[i] = static_cast<uint32_t>(keys[i] & 0xFFFFFFFFull);
  });
  out->offsets = offsets;
  out->targets = targets;
  out->num_nodes = n;
  out->num_edges = static_cast<uint32_t>(keys.size());
}

void check_pairs(const uint32_t* pairs, size_t num_pairs, uint32_t n) {
  for (size_t i = 0; i < num_pairs; ++i) {
    uint32_t s = pairs[2 * i], d = pairs[2 * i + 1];
    if (s >= n || d >= n) {
      fail_csr("edge (" + std::to_string(s) + ", " + std::to_string(d) + ") has a node id >= num_nodes " +
               std::to_string(n));
    }
  }
}

}  // namespace

}  // namespace gc

using gc::guard;

extern "C" {

void gc_csr_free(gc_csr* csr) {
  if (csr == nullptr) return;
  std::free(csr->offsets);
  std::free(csr->targets);
  csr->offsets = nullptr;
  csr->targets = nullptr;
  csr->num_nodes = 0;
  csr->num_edges = 0;
}

gc_status gc_validate_csr(const uint32_t* offsets, size_t num_offsets, const uint32_t* targets,
                          size_t num_targets) {
  return guard([&] {
    if ((offsets == nullptr && num_offsets > 0) || (targets == nullptr && num_targets > 0)) {
      gc::fail_kernel("null array with non-zero length");
    }
    gc::validate_csr(offsets, num_offsets, targets, num_targets);
  });
}

gc_status gc_csr_validate(const gc_csr* csr) {
  return guard([&] {
    if (csr == nullptr || csr->offsets == nullptr) gc::fail_kernel("csr is null");
    // A gc_csr carries no array lengths besides its declared counts, so the
    // counts are the lengths: offsets[num_nodes] must equal num_edges.
    gc::validate_csr(csr->offsets, size_t(csr->num_nodes) + 1, csr->targets, csr->num_edges);
  });
}

gc_status gc_csr_from_edge_pairs(const uint32_t* pairs, size_t num_pairs, uint32_t num_nodes,
                                 int symmetric, gc_csr* out) {
  return guard([&] {
    if (out == nullptr || (pairs == nullptr && num_pairs > 0)) gc::fail_kernel("null argument");
    gc::check_pairs(pairs, num_pairs, num_nodes);
    std::vector<uint64_t> keys;
    keys.reserve(symmetric ? 2 * num_pairs : num_pairs);
    for (size_t i = 0; i < num_pairs; ++i) {
      uint64_t s = pairs[2 * i], d = pairs[2 * i + 1];
      keys.push_back((s << 32) | d);
      if (symmetric && s != d) keys.push_back((d << 32) | s);
    }
    gc::sort_unique(keys, num_nodes);
    gc::keys_to_csr(keys, num_nodes, out);
  });
}

gc_status gc_csr_from_edge_pairs_oriented(const uint32_t* pairs, size_t num_pairs,
                                          uint32_t num_nodes, gc_csr* out) {
  return guard([&] {
    if (out == nullptr || (pairs == nullptr && num_pairs > 0)) gc::fail_kernel("null argument");
    gc::check_pairs(pairs, num_pairs, num_nodes);
    std::vector<uint64_t> keys;
    keys.reserve(num_pairs);
    for (size_t i = 0; i < num_pairs; ++i) {
      uint64_t s = pairs[2 * i], d = pairs[2 * i + 1];
      if (s == d) continue;
      keys.push_back((std::min(s, d) << 32) | std::max(s, d));
    }
    gc::sort_unique(keys, num_nodes);
    std::vector<uint32_t> degree(num_nodes, 0);
    for (uint64_t k : keys) {
      ++degree[k >> 32];
      ++degree[k & 0xFFFFFFFFull];
    }
    gc::parallel_for(keys.size(), [&](uint64_t b, uint64_t e) {
      for (uint64_t i = b; i < e; ++i) {
        uint64_t u = keys[i] >> 32, v = keys[i] & 0xFFFFFFFFull;
        uint32_t du = degree[u], dv = degree[v];
        keys[i] = (du < dv || (du == dv && u < v)) ? ((u << 32) | v) : ((v << 32) | u);
      }
    });
    gc::sort_unique(keys, num_nodes);  // no duplicates remain; sorts by (src, dst)
    gc::keys_to_csr(keys, num_nodes, out);
  });
}

namespace {

// Byte length of the White_Space code point (the property Rust's str::trim and
// split_whitespace use) starting at p < e, or 0: U+0009..000D and U+0020 (one
// byte), U+0085 and U+00A0 (two), U+1680, U+2000..200A, U+2028, U+2029,
// U+202F, U+205F and U+3000 (three).
size_t space_at(const unsigned char* p, const unsigned char* e) {
  const unsigned c = p[0];
  if (c == ' ' || (c >= 0x09 && c <= 0x0D)) return 1;
  if (c < 0xC2 || c > 0xE3 || e - p < 2) return 0;
  const unsigned c1 = p[1];
  if (c == 0xC2) return c1 == 0x85 || c1 == 0xA0 ? 2 : 0;
  if (e - p < 3) return 0;
  const unsigned c2 = p[2];
  if (c == 0xE1) return c1 == 0x9A && c2 == 0x80 ? 3 : 0;
  if (c == 0xE2) {
    if (c1 == 0x80) return (c2 >= 0x80 && c2 <= 0x8A) || c2 == 0xA8 || c2 == 0xA9 || c2 == 0xAF ? 3 : 0;
    return c1 == 0x81 && c2 == 0x9F ? 3 : 0;
  }
  if (c == 0xE3) return c1 == 0x80 && c2 == 0x80 ? 3 : 0;
  return 0;
}

// Byte length of a White_Space code point ending at e[-1], b < e, or 0 (the
// line is valid UTF-8, so a matching suffix is that code point).
size_t space_before(const unsigned char* b, const unsigned char* e) {
  if (space_at(e - 1, e) == 1) return 1;
  if (e - b >= 2 && space_at(e - 2, e) == 2) return 2;
  if (e - b >= 3 && space_at(e - 3, e) == 3) return 3;
  return 0;
}

// Whether [b, e) is well-formed UTF-8 (no overlong forms, no surrogates,
// nothing above U+10FFFF), the check of Rust's BufRead::lines.
bool utf8_ok(const unsigned char* b, const unsigned char* e) {
  for (const unsigned char* p = b; p < e;) {
    const unsigned c = *p;
    if (c < 0x80) {
      ++p;
      continue;
    }
    ptrdiff_t len;
    unsigned lo = 0x80, hi = 0xBF;
    if (c >= 0xC2 && c <= 0xDF) {
      len = 2;
    } else if (c >= 0xE0 && c <= 0xEF) {
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
    if (e - p < len || p[1] < lo || p[1] > hi) return false;
    for (ptrdiff_t k = 2; k < len; ++k)
      if ((p[k] & 0xC0u) != 0x80u) return false;
    p += len;
  }
  return true;
}

}  // namespace

gc_status gc_parse_tsv_to_edges(const char* path, uint32_t** pairs, size_t* num_pairs,
                                uint32_t* num_nodes) {
  return guard([&] {
    if (path == nullptr || pairs == nullptr || num_pairs == nullptr || num_nodes == nullptr) {
      gc::fail_kernel("null argument");
    }
    FILE* f = std::fopen(path, "rb");
    if (f == nullptr) {
      gc::fail(GC_ERR_EXPORT, std::string("cannot open ") + path + ": " + std::strerror(errno));
    }
    std::unique_ptr<FILE, int (*)(FILE*)> file(f, &std::fclose);  // closed on every path
    std::unordered_map<uint64_t, uint32_t> ids;
    std::vector<uint32_t> out;
    uint32_t next_id = 0;
    auto id_of = [&](uint64_t raw) {
      auto it = ids.find(raw);
      if (it != ids.end()) return it->second;
      ids.emplace(raw, next_id);
      return next_id++;
    };
    // Parses a decimal u64 field the way Rust's u64::from_str does: an
    // optional '+', then one or more digits, no overflow.
    auto parse_u64 = [](const char* b, const char* e, uint64_t* value) {
      if (b < e && *b == '+') ++b;
      if (b == e) return false;
      uint64_t v = 0;
      for (const char* p = b; p < e; ++p) {
        if (*p < '0' || *p > '9') return false;
        uint64_t digit = static_cast<uint64_t>(*p - '0');
        if (v > (UINT64_MAX - digit) / 10) return false;
        v = v * 10 + digit;
      }
      *value = v;
      return true;
    };
    std::vector<char> buf(1 << 20);
    std::string line;
    // As Rust's lines / trim / split_whitespace: a line that is not UTF-8
    // fails the parse; White_Space is trimmed and separates the fields.
    auto handle_line = [&](const char* cb, const char* ce) {
      const auto* b = reinterpret_cast<const unsigned char*>(cb);
      const auto* e = reinterpret_cast<const unsigned char*>(ce);
      if (!utf8_ok(b, e)) gc::fail(GC_ERR_EXPORT, gc::kInvalidUtf8);
      size_t k;
      while (b < e && (k = space_at(b, e)) != 0) b += k;
      while (e > b && (k = space_before(b, e)) != 0) e -= k;
      if (b == e || *b == '#' || *b == '%') return;
      const char* f1b = reinterpret_cast<const char*>(b);
      const unsigned char* p = b;
      while (p < e && space_at(p, e) == 0) ++p;
      const char* f1e = reinterpret_cast<const char*>(p);
      while (p < e && (k = space_at(p, e)) != 0) p += k;
      const char* f2b = reinterpret_cast<const char*>(p);
      while (p < e && space_at(p, e) == 0) ++p;
      const char* f2e = reinterpret_cast<const char*>(p);
      uint64_t s, d;
      if (!parse_u64(f1b, f1e, &s)) return;
      if (!parse_u64(f2b, f2e, &d)) return;
      uint32_t si = id_of(s);
      uint32_t di = id_of(d);
      out.push_back(si);
      out.push_back(di);
    };
    size_t got;
    while ((got = std::fread(buf.data(), 1, buf.size(), f)) > 0) {
      size_t start = 0;
      for (size_t i = 0; i < got; ++i) {
        if (buf[i] == '\n') {
          if (line.empty()) {
            handle_line(buf.data() + start, buf.data() + i);
          } else {
            line.append(buf.data() + start, i - start);
            handle_line(line.data(), line.data() + line.size());
            line.clear();
          }
          start = i + 1;
        }
      }
      line.append(buf.data() + start, got - start);
    }
    bool read_error = std::ferror(f) != 0;
    file.reset();
    if (read_error) gc::fail(GC_ERR_EXPORT, std::string("cannot read ") + path);
    if (!line.empty()) handle_line(line.data(), line.data() + line.size());
    auto* result = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * std::max<size_t>(1, out.size())));
    if (!result) throw std::bad_alloc();
    std::memcpy(result, out.data(), sizeof(uint32_t) * out.size());
    *pairs = result;
    *num_pairs = out.size() / 2;
    *num_nodes = next_id;
  });
}

}  // extern "C"
