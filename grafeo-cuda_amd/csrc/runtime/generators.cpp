// Deterministic synthetic graphs for benchmarks and tests.
//
// The reference benchmarks load TSV files (crates/grafeo-cuda-algos/src/harness.rs:84-160);
// these generators make the same kind of input from (kind, size, seed) alone.
//
// Random numbers: SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) used with
// random access, U_j(seed) = mix64(seed + (j + 1) * 0x9E3779B97F4A7C15), so
// draw i of any generator depends only on (seed, i) and generation runs in
// parallel with the same output on any thread count.
//   Erdos-Renyi draw i: u = hi32(U_{2i}) * n >> 32, v = hi32(U_{2i+1}) * n >> 32.
//   R-MAT draw i, level l (0 = most significant bit): a 32-bit uniform r taken
//     from U_{i*ceil(scale/2) + l/2} (high half for even l, low half for odd
//     l); quadrant (0,0) if r < A, (0,1) if r < A+B, (1,0) if r < A+B+C, else
//     (1,1), with A = floor(a * 2^32) etc.
//   Vertex permutation: Fisher-Yates from i = N-1 down to 1 with
//     j = (U_k(seed ^ 0x5851F42D4C957F2D) * (i + 1)) >> 64, k = N-1-i.
// Self-loops are dropped after permutation; pairs keep draw order.
#include <algorithm>
#include <cstdlib>

#include "common.hpp"

namespace gc {
namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t splitmix_at(uint64_t seed, uint64_t j) { return mix64(seed + (j + 1) * kGamma); }

inline uint32_t scale_u32(uint64_t r, uint32_t n) {
  return static_cast<uint32_t>(((r >> 32) * static_cast<uint64_t>(n)) >> 32);
}

uint32_t* alloc_pairs(size_t num_pairs) {
  auto* p = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * 2 * std::max<size_t>(1, num_pairs)));
  if (p == nullptr) throw std::bad_alloc();
  return p;
}

// Drops self-loops from `raw` (2*count entries) into a malloc'd array.
void compact_loops(const std::vector<uint32_t>& raw, size_t count, uint32_t** pairs,
                   size_t* num_pairs) {
  size_t kept = 0;
  for (size_t i = 0; i < count; ++i) kept += raw[2 * i] != raw[2 * i + 1];
  uint32_t* out = alloc_pairs(kept);
  size_t w = 0;
  for (size_t i = 0; i < count; ++i) {
    if (raw[2 * i] != raw[2 * i + 1]) {
      out[2 * w] = raw[2 * i];
      out[2 * w + 1] = raw[2 * i + 1];
      ++w;
    }
  }
  *pairs = out;
  *num_pairs = kept;
}

}  // namespace
}  // namespace gc

using gc::guard;

extern "C" {

gc_status gc_gen_erdos_renyi(uint32_t n, uint64_t num_draws, uint64_t seed, uint32_t** pairs,
                             size_t* num_pairs) {
  return guard([&] {
    if (pairs == nullptr || num_pairs == nullptr) gc::fail_kernel("null argument");
    if (n == 0 && num_draws > 0) gc::fail_kernel("cannot draw edges over 0 nodes");
    std::vector<uint32_t> raw(2 * num_draws);
    gc::parallel_for(num_draws, [&](uint64_t b, uint64_t e) {
      for (uint64_t i = b; i < e; ++i) {
        raw[2 * i] = gc::scale_u32(gc::splitmix_at(seed, 2 * i), n);
        raw[2 * i + 1] = gc::scale_u32(gc::splitmix_at(seed, 2 * i + 1), n);
      }
    });
    gc::compact_loops(raw, num_draws, pairs, num_pairs);
  });
}

gc_status gc_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                      uint64_t seed, int32_t permute, uint32_t** pairs, size_t* num_pairs) {
  return guard([&] {
    if (pairs == nullptr || num_pairs == nullptr) gc::fail_kernel("null argument");
    if (scale == 0 || scale > 31) gc::fail_kernel("R-MAT scale must be in 1..31");
    if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) {
      gc::fail_kernel("R-MAT probabilities must be non-negative with a + b + c <= 1");
    }
    const uint64_t n = uint64_t(1) << scale;
    const uint64_t draws = static_cast<uint64_t>(edge_factor) << scale;
    const double two32 = 4294967296.0;
    const uint64_t ta = static_cast<uint64_t>(a * two32);
    const uint64_t tab = static_cast<uint64_t>((a + b) * two32);
    const uint64_t tabc = static_cast<uint64_t>((a + b + c) * two32);
    const uint64_t per_edge = (scale + 1) / 2;
    std::vector<uint32_t> perm;
    if (permute) {
      perm.resize(n);
      for (uint64_t i = 0; i < n; ++i) perm[i] = static_cast<uint32_t>(i);
      const uint64_t pseed = seed ^ 0x5851F42D4C957F2Dull;
      for (uint64_t i = n - 1; i >= 1; --i) {
        uint64_t r = gc::splitmix_at(pseed, n - 1 - i);
        uint64_t j = static_cast<uint64_t>((static_cast<unsigned __int128>(r) * (i + 1)) >> 64);
        std::swap(perm[i], perm[j]);
      }
    }
    std::vector<uint32_t> raw(2 * draws);
    gc::parallel_for(draws, [&](uint64_t bgn, uint64_t end) {
      for (uint64_t i = bgn; i < end; ++i) {
        uint32_t src = 0, dst = 0;
        uint64_t word = 0;
        for (uint32_t l = 0; l < scale; ++l) {
          if ((l & 1) == 0) word = gc::splitmix_at(seed, i * per_edge + l / 2);
          uint64_t r = (l & 1) == 0 ? (word >> 32) : (word & 0xFFFFFFFFull);
          uint32_t bit = scale - 1 - l;
          if (r < ta) {
          } else if (r < tab) {
            dst |= 1u << bit;
          } else if (r < tabc) {
            src |= 1u << bit;
          } else {
            src |= 1u << bit;
            dst |= 1u << bit;
          }
        }
        if (permute) {
          src = perm[src];
          dst = perm[dst];
        }
        raw[2 * i] = src;
        raw[2 * i + 1] = dst;
      }
    });
    gc::compact_loops(raw, draws, pairs, num_pairs);
  });
}

gc_status gc_gen_grid(uint32_t rows, uint32_t cols, uint32_t** pairs, size_t* num_pairs) {
  return guard([&] {
    if (pairs == nullptr || num_pairs == nullptr) gc::fail_kernel("null argument");
    uint64_t n = uint64_t(rows) * cols;
    if (n > 0xFFFFFFFFull) gc::fail_kernel("grid too large for u32 node ids");
    uint64_t count = 0;
    if (rows > 0 && cols > 0) count = uint64_t(rows) * (cols - 1) + uint64_t(rows - 1) * cols;
    uint32_t* out = gc::alloc_pairs(count);
    uint64_t w = 0;
    for (uint32_t r = 0; r < rows; ++r) {
      for (uint32_t c = 0; c < cols; ++c) {
        uint32_t id = r * cols + c;
        if (c + 1 < cols) {
          out[2 * w] = id;
          out[2 * w + 1] = id + 1;
          ++w;
        }
        if (r + 1 < rows) {
          out[2 * w] = id;
          out[2 * w + 1] = id + cols;
          ++w;
        }
      }
    }
    *pairs = out;
    *num_pairs = count;
  });
}

}  // extern "C"
