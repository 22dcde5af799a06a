// Host construction of PageRank's binned layout of the cold entries (the
// static arrays that pr_bin_scatter, pr_bin_accum and pr_spmv_bin of
// pagerank.hip read). Plain C++ over the degree-ordered symmetric CSR, no HIP:
// the library uploads the result, and a host-only unit test replays the three
// kernels' indexing on it.
//
// For a CSR of n rows sorted by degree (the `zero_rows` rows of degree 0
// first), the push set [n - push_count, n) and the hot set [n - hot_count, n):
//  - Pull rows [z, push_base), z = min(zero_rows, push_base). Each row's
//    entries at or above hot_base form its hot stream, stored as 16-bit
//    offsets from hot_base (hoff[r] .. hoff[r + 1]); a row with none holds one
//    offset equal to hot_count (the kernels read a zero there).
//  - Its other entries are cold. Bins cut the pull rows into consecutive
//    ranges of at most `bin_rows` rows and (unless one row has more)
//    `bin_values` cold entries; bin b's values sit at [bin_val[b],
//    bin_val[b + 1]), padded to a multiple of eight with row index equal to
//    the bin's row count (the accumulator the kernel ignores).
//  - Cold sources [z, hot_base): the cold entries seen from their sources,
//    which on a symmetric CSR are the entries below push_base of the rows
//    t < hot_base. Sources come in groups of 64 consecutive ids; group g takes
//    its largest cold count of slot rows, rounded up to a multiple of eight
//    (blocks of eight rows, each of one group: block_group), entry j of source
//    z + 64 g + l occupying slot row (first row of g) + j, lane l; block q
//    keeps the eight slots of lane l at slots[(64 q + l) * 8 .. + 8].
//    Consecutive blocks form sub-ranges of at most `stage` real entries
//    (stage >= 512, so any block fits). In a
//    sub-range the entries are staged bin by bin (in slot-row, then lane
//    order within a bin); padding slots hold `stage`. run r of sub-range s
//    (runs [run_first[s], run_first[s + 1])) copies the staged range
//    [run_info & 0xFFFF, + (run_info >> 16)) to values [run_dst[r], ...), and
//    uloc gives each value its row within the bin.
//  - Hub pieces: the entries of the push-set rows that point into the push
//    set, cut into pieces of at most 512 entries (hub row, first entry),
//    from the top row down, so that these pushes spread evenly over the waves.
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace gc {

struct PrBinLayout {
  uint32_t zero_rows = 0, push_base = 0, hot_base = 0;
  uint64_t cold_entries = 0;
  std::vector<uint32_t> hoff;       // [push_base + 1]: 16-bit stream start of each row (0 up to z)
  std::vector<uint16_t> hot16;      // the 16-bit stream
  std::vector<uint32_t> bin_row;    // [bins + 1]
  std::vector<uint32_t> bin_val;    // [bins + 1], multiples of eight
  std::vector<uint16_t> uloc;       // [bin_val.back()]
  std::vector<uint32_t> block_group;  // group of each block of eight slot rows
  std::vector<uint16_t> slots;      // [blocks * 64 * 8]
  std::vector<uint32_t> sub;        // [subs + 1] block ranges
  std::vector<uint32_t> run_first;  // [subs + 1]
  std::vector<uint32_t> run_info;   // staging begin | length << 16
  std::vector<uint32_t> run_dst;    // first value index of the run
  std::vector<uint32_t> hub_pieces; // (hub row, first entry) pairs
  uint32_t pieces() const { return static_cast<uint32_t>(hub_pieces.size() / 2); }
  uint32_t bins() const { return bin_row.empty() ? 0u : static_cast<uint32_t>(bin_row.size() - 1); }
  uint32_t subs() const { return sub.empty() ? 0u : static_cast<uint32_t>(sub.size() - 1); }
};

// ParFor: par(count, fn) runs fn(begin, end) over a split of [0, count).
template <class ParFor>
PrBinLayout build_pr_bin_layout(const uint32_t* off, const uint32_t* tgt, uint32_t n, uint32_t zero_rows,
                                uint32_t push_count, uint32_t hot_count, uint32_t stage, uint32_t bin_rows,
                                uint32_t bin_values, ParFor par) {
  if (push_count == 0 || push_count > n || hot_count < push_count || hot_count > n || hot_count >= 0xFFFFu ||
      stage < 512 || stage >= 0xFFFFu || bin_rows == 0 || bin_rows >= 0xFFFFu || bin_values == 0)
    throw std::runtime_error("PageRank: binned layout sizes out of range");
  PrBinLayout p;
  const uint32_t pb = n - push_count, hb = n - hot_count;
  const uint32_t z = std::min(zero_rows, pb);
  p.zero_rows = z;
  p.push_base = pb;
  p.hot_base = hb;
  const uint32_t pull = pb - z;

  // pieces of the push-set rows' entries inside the push set
  for (uint32_t h = n; h-- > pb;) {
    const uint32_t b = static_cast<uint32_t>(std::lower_bound(tgt + off[h], tgt + off[h + 1], pb) - tgt);
    for (uint32_t q = b; q < off[h + 1]; q += 512) {
      p.hub_pieces.push_back(h);
      p.hub_pieces.push_back(q);
    }
  }

  // the 16-bit stream of the pull rows' entries in [hb, n)
  std::vector<uint32_t> hsplit(pull);
  par(pull, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t r = z + static_cast<uint32_t>(i);
      hsplit[i] = static_cast<uint32_t>(std::lower_bound(tgt + off[r], tgt + off[r + 1], hb) - tgt);
    }
  });
  p.hoff.assign(static_cast<size_t>(pb) + 1, 0);
  uint64_t h_total = 0;
  for (uint32_t i = 0; i < pull; ++i) {
    const uint32_t r = z + i;
    p.hoff[r] = static_cast<uint32_t>(h_total);
    h_total += std::max(1u, off[r + 1] - hsplit[i]);
    if (h_total >= 0xFFFFFFFFull) throw std::runtime_error("PageRank: binned stream too long");
  }
  p.hoff[pb] = static_cast<uint32_t>(h_total);
  p.hot16.assign(h_total, 0);
  par(pull, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t r = z + static_cast<uint32_t>(i);
      uint16_t* out = p.hot16.data() + p.hoff[r];
      if (hsplit[i] == off[r + 1]) {
        out[0] = static_cast<uint16_t>(hot_count);
      } else {
        for (uint32_t q = hsplit[i]; q < off[r + 1]; ++q) *out++ = static_cast<uint16_t>(tgt[q] - hb);
      }
    }
  });

  // bins over the pull rows by their cold counts
  std::vector<uint32_t> rowbin(pull);
  {
    uint64_t vals = 0;
    uint32_t r = z;
    while (r < pb) {
      const uint32_t b = static_cast<uint32_t>(p.bin_row.size());
      p.bin_row.push_back(r);
      p.bin_val.push_back(static_cast<uint32_t>(vals));
      uint64_t cnt = 0;
      uint32_t rows = 0;
      while (r < pb && rows < bin_rows) {
        const uint32_t c = hsplit[r - z] - off[r];
        if (rows > 0 && cnt + c > bin_values) break;
        cnt += c;
        rowbin[r - z] = b;
        ++rows;
        ++r;
      }
      vals += (cnt + 7) & ~7ull;
      if (vals >= 0xFFFFFFFFull) throw std::runtime_error("PageRank: binned value buffer too long");
      p.cold_entries += cnt;
    }
    p.bin_row.push_back(pb);
    p.bin_val.push_back(static_cast<uint32_t>(vals));
  }
  const uint32_t bins = p.bins();

  // cold sources: counts, groups of 64, slot rows, blocks of eight rows
  const uint32_t srcs = hb > z ? hb - z : 0;
  const uint32_t groups = (srcs + 63) / 64;
  std::vector<uint32_t> cd(srcs);
  par(srcs, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t t = z + static_cast<uint32_t>(i);
      cd[i] = static_cast<uint32_t>(std::lower_bound(tgt + off[t], tgt + off[t + 1], pb) - tgt) - off[t];
    }
  });
  uint64_t cold_by_source = 0;
  for (uint32_t c : cd) cold_by_source += c;
  if (cold_by_source != p.cold_entries)
    throw std::runtime_error("PageRank: the cold entries differ by rows and by sources (CSR not symmetric)");
  constexpr uint32_t kPad = 0xFFFFFFFFu;
  std::vector<uint32_t> row_j;  // per slot row: the entry index j (kPad: padding row)
  std::vector<uint32_t> row_group;
  for (uint32_t gi = 0; gi < groups; ++gi) {
    uint32_t len = 0;
    for (uint32_t l = 0; l < 64 && gi * 64 + l < srcs; ++l) len = std::max(len, cd[gi * 64 + l]);
    const uint32_t len8 = (len + 7) & ~7u;
    for (uint32_t j = 0; j < len8; ++j) {
      row_group.push_back(gi);
      row_j.push_back(j < len ? j : kPad);
    }
  }
  const uint64_t blocks = row_group.size() / 8;
  if (blocks >= 0xFFFFFFFFull) throw std::runtime_error("PageRank: binned slot array too long");
  p.block_group.resize(blocks);
  for (uint64_t q = 0; q < blocks; ++q) p.block_group[q] = row_group[q * 8];
  auto real_in_row = [&](uint64_t k) {
    if (row_j[k] == kPad) return 0u;
    uint32_t c = 0;
    for (uint32_t l = 0; l < 64; ++l) {
      const uint64_t si = static_cast<uint64_t>(row_group[k]) * 64 + l;
      if (si < srcs && cd[si] > row_j[k]) ++c;
    }
    return c;
  };
  p.sub.push_back(0);
  {
    uint32_t acc = 0;
    for (uint64_t blk = 0; blk < blocks; ++blk) {
      uint32_t c = 0;
      for (uint32_t i = 0; i < 8; ++i) c += real_in_row(blk * 8 + i);
      if (acc > 0 && acc + c > stage) {
        p.sub.push_back(static_cast<uint32_t>(blk));
        acc = 0;
      }
      acc += c;
    }
    if (blocks > 0) p.sub.push_back(static_cast<uint32_t>(blocks));
  }
  const uint32_t subs = p.subs();

  // fn(slot index, row of the entry) for every real entry of sub-range s, in staging order
  auto for_entries = [&](uint32_t s, auto&& fn) {
    for (uint64_t blk = p.sub[s]; blk < p.sub[s + 1]; ++blk) {
      for (uint32_t i = 0; i < 8; ++i) {
        const uint64_t k = blk * 8 + i;
        if (row_j[k] == kPad) continue;
        for (uint32_t l = 0; l < 64; ++l) {
          const uint64_t si = static_cast<uint64_t>(row_group[k]) * 64 + l;
          if (si < srcs && cd[si] > row_j[k]) {
            const uint32_t t = z + static_cast<uint32_t>(si);
            const uint32_t u = tgt[off[t] + row_j[k]];
            if (u < z) throw std::runtime_error("PageRank: a row of degree 0 is a target (CSR not symmetric)");
            fn((blk * 64 + l) * 8 + i, u);
          }
        }
      }
    }
  };
  // entries per sub-range and bin, then each bin's write position per sub-range
  std::vector<uint32_t> counts(static_cast<size_t>(subs) * bins, 0);
  par(subs, [&](uint64_t b, uint64_t e) {
    for (uint64_t s = b; s < e; ++s) {
      uint32_t* c = counts.data() + s * bins;
      for_entries(static_cast<uint32_t>(s), [&](uint64_t, uint32_t u) { ++c[rowbin[u - z]]; });
    }
  });
  std::vector<uint32_t> pos(counts.size());
  {
    std::vector<uint64_t> at(p.bin_val.begin(), p.bin_val.end() - 1);
    for (uint32_t s = 0; s < subs; ++s) {
      for (uint32_t b = 0; b < bins; ++b) {
        pos[static_cast<size_t>(s) * bins + b] = static_cast<uint32_t>(at[b]);
        at[b] += counts[static_cast<size_t>(s) * bins + b];
      }
    }
  }
  p.run_first.assign(static_cast<size_t>(subs) + 1, 0);
  for (uint32_t s = 0; s < subs; ++s) {
    uint32_t runs = 0;
    for (uint32_t b = 0; b < bins; ++b) runs += counts[static_cast<size_t>(s) * bins + b] != 0;
    p.run_first[s + 1] = p.run_first[s] + runs;
  }
  p.run_info.assign(p.run_first[subs], 0);
  p.run_dst.assign(p.run_first[subs], 0);
  p.slots.assign(blocks * 64 * 8, static_cast<uint16_t>(stage));
  p.uloc.assign(p.bin_val.back(), 0);
  for (uint32_t b = 0; b < bins; ++b) {
    const uint16_t pad = static_cast<uint16_t>(p.bin_row[b + 1] - p.bin_row[b]);
    for (uint32_t v = p.bin_val[b]; v < p.bin_val[b + 1]; ++v) p.uloc[v] = pad;  // real ones overwritten below
  }
  par(subs, [&](uint64_t b, uint64_t e) {
    std::vector<uint32_t> stage_off(bins), cur(bins);
    for (uint64_t s = b; s < e; ++s) {
      const uint32_t* c = counts.data() + s * bins;
      const uint32_t* ps = pos.data() + s * bins;
      uint32_t at = 0, r = p.run_first[s];
      for (uint32_t bi = 0; bi < bins; ++bi) {
        stage_off[bi] = at;
        cur[bi] = 0;
        if (c[bi] != 0) {
          p.run_info[r] = at | (c[bi] << 16);
          p.run_dst[r] = ps[bi];
          ++r;
        }
        at += c[bi];
      }
      for_entries(static_cast<uint32_t>(s), [&](uint64_t slot, uint32_t u) {
        const uint32_t bi = rowbin[u - z];
        p.slots[slot] = static_cast<uint16_t>(stage_off[bi] + cur[bi]);
        p.uloc[ps[bi] + cur[bi]] = static_cast<uint16_t>(u - p.bin_row[bi]);
        ++cur[bi];
      });
    }
  });
  return p;
}

}  // namespace gc
