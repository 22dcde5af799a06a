// Host-only test of PageRank's binned layout (csrc/algos/pr_bin_layout.hpp)
// and of the sliced hot stream built from it (pr_slice_layout.hpp): on random
// symmetric graphs in degree order it replays the indexing of pr_bin_scatter,
// pr_bin_accum, pr_spmv_bin and pr_spmv_slice (pagerank.hip) with integer
// contributions and checks that every pull row's cold sum plus its 16-bit
// stream sum equals the sum over its neighbours, that every staged slot and
// every value is written exactly once, that the slices hand each lane its
// row's hot entries in order, every sliced entry read once and every item
// dealt to one wave, and that padding stays inert.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "pr_bin_layout.hpp"
#include "pr_slice_layout.hpp"

namespace {

int failures = 0;

void check(bool ok, const std::string& what) {
  if (!ok) {
    ++failures;
    if (failures < 20) std::printf("FAIL: %s\n", what.c_str());
  }
}

uint64_t rng_state = 0x9E3779B97F4A7C15ull;
uint64_t next_u64() {
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Csr {
  uint32_t n = 0;
  uint32_t zero_rows = 0;
  std::vector<uint32_t> off, tgt;
};

// A random symmetric graph (a few hubs, optional self-loops, isolated
// vertices), relabelled by (degree, id) ascending with sorted rows, as the
// library's degree-ordered layout.
Csr random_graph(uint32_t n, uint32_t edges, uint32_t hubs, bool loops) {
  std::vector<std::vector<uint32_t>> adj(n);
  for (uint32_t e = 0; e < edges; ++e) {
    uint32_t u = static_cast<uint32_t>(next_u64() % n);
    uint32_t v = static_cast<uint32_t>(next_u64() % n);
    if (hubs > 0 && next_u64() % 3 == 0) v = static_cast<uint32_t>(next_u64() % hubs);
    if (u == v && !loops) continue;
    adj[u].push_back(v);
    if (u != v) adj[v].push_back(u);
  }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return adj[a].size() < adj[b].size(); });
  std::vector<uint32_t> inv(n);
  for (uint32_t i = 0; i < n; ++i) inv[order[i]] = i;
  Csr g;
  g.n = n;
  g.off.push_back(0);
  for (uint32_t i = 0; i < n; ++i) {
    std::vector<uint32_t> row;
    for (uint32_t v : adj[order[i]]) row.push_back(inv[v]);
    std::sort(row.begin(), row.end());
    g.tgt.insert(g.tgt.end(), row.begin(), row.end());
    g.off.push_back(static_cast<uint32_t>(g.tgt.size()));
    if (row.empty()) ++g.zero_rows;
  }
  return g;
}

void serial(uint64_t count, const std::function<void(uint64_t, uint64_t)>& fn) {
  // uneven split, to exercise the chunked callers
  uint64_t b = 0;
  while (b < count) {
    const uint64_t e = std::min(count, b + 1 + next_u64() % 97);
    fn(b, e);
    b = e;
  }
}

// pr_spmv_slice's reads of the sliced layout, replayed for every item
void check_slices(const Csr& g, const gc::PrBinLayout& p, uint32_t hot, uint32_t cap, uint32_t waves,
                  const std::string& tag0) {
  const std::string tag = tag0 + " cap=" + std::to_string(cap) + " waves=" + std::to_string(waves);
  const gc::PrSliceLayout sl = gc::build_pr_slice_layout(p.hoff.data(), p.hot16.data(), p.zero_rows, p.push_base,
                                                         hot, p.hub_pieces, g.off.data(), g.tgt.data(), cap, waves,
                                                         serial);
  // the hub pieces' entries as offsets from push_base, 0xFFFF past each piece's end
  check(sl.hub16.size() == static_cast<size_t>(p.pieces()) * 512, tag + ": hub16 size");
  for (uint32_t i = 0; i < p.pieces() && sl.hub16.size() == static_cast<size_t>(p.pieces()) * 512; ++i) {
    const uint32_t h = p.hub_pieces[2 * i], b = p.hub_pieces[2 * i + 1];
    const uint32_t len = std::min(b + 512u, g.off[h + 1]) - b;
    for (uint32_t j = 0; j < 512; ++j) {
      const uint16_t x = sl.hub16[512 * static_cast<size_t>(i) + j];
      if (j < len) check(x == g.tgt[b + j] - p.push_base, tag + ": hub16 entry");
      else check(x == gc::kPrHubNone, tag + ": hub16 padding");
    }
  }
  const uint32_t z = p.zero_rows, pb = p.push_base, S = sl.slices;
  check(S == (pb - z + 63) / 64, tag + ": slice count");
  check(sl.slice_off.size() == S + 1u && sl.slice_off[S] + gc::kPrSlicePad == sl.data.size(), tag + ": slice offsets");
  check(sl.wave_first.size() == waves + 1u && sl.wave_first[0] == 0 && sl.wave_first[waves] == sl.item_count(),
        tag + ": wave ranges");
  for (uint32_t w = 0; w < waves; ++w) check(sl.wave_first[w] <= sl.wave_first[w + 1], tag + ": wave ranges ordered");
  std::vector<uint32_t> reads(sl.data.size(), 0);
  std::vector<uint32_t> hub_seen(p.pieces(), 0);
  // per slice: its pieces by partial slot (or the whole slice), each lane's entries in order
  std::vector<std::vector<std::pair<uint32_t, std::vector<std::vector<uint16_t>>>>> got(S);
  std::vector<uint32_t> whole(S, 0), slot_used(sl.partial_slots, 0);
  for (uint32_t k = 0; k < sl.item_count(); ++k) {
    const uint32_t* it = &sl.items[4 * static_cast<size_t>(k)];
    const uint32_t kind = it[0] >> 30, idx = it[0] & 0x3FFFFFFFu;
    if (kind == gc::kPrSliceHub) {
      check(idx < p.pieces() && it[1] == p.hub_pieces[2 * idx] && it[2] == p.hub_pieces[2 * idx + 1] &&
                it[3] == 512 * idx,
            tag + ": hub item");
      if (idx < p.pieces()) ++hub_seen[idx];
      continue;
    }
    check(kind <= 1 && idx < S, tag + ": slice item");
    if (kind > 1 || idx >= S) continue;
    const uint32_t ng = it[2] & 0xFFFFu, R = it[2] >> 16;
    check(R <= 6 && R % 2 == 0 && ng <= cap, tag + ": piece size");
    std::vector<std::vector<uint16_t>> lanes(64);
    for (uint32_t l = 0; l < 64; ++l) {
      for (uint32_t q = 0; q < ng; ++q)
        for (uint32_t i = 0; i < 8; ++i) {
          const uint64_t at = static_cast<uint64_t>(it[1]) + 512ull * q + 8 * l + i;
          check(at < sl.data.size(), tag + ": group read in range");
          if (at >= sl.data.size()) continue;
          ++reads[at];
          lanes[l].push_back(sl.data[at]);
        }
      // the remainder: one 16-byte load per lane, its first R entries used
      const uint64_t rem = static_cast<uint64_t>(it[1]) + 512ull * ng + static_cast<uint64_t>(R) * l;
      check(rem + 8 <= sl.data.size(), tag + ": remainder load in range");
      for (uint32_t i = 0; i < R && rem + i < sl.data.size(); ++i) {
        ++reads[rem + i];
        lanes[l].push_back(sl.data[rem + i]);
      }
      // the four group loads of a batch, clamped to the item's last group (to group 0 without groups)
      for (uint32_t q = 0; q < 4; ++q) {
        const uint64_t at = static_cast<uint64_t>(it[1]) + 512ull * std::min(q, ng > 0 ? ng - 1 : 0u) + 8 * l;
        check(at + 8 <= sl.data.size(), tag + ": clamped group load in range");
      }
    }
    if (kind == 0) {
      ++whole[idx];
      got[idx].push_back({0u, lanes});
    } else {
      check(it[3] < sl.partial_slots, tag + ": partial slot in range");
      if (it[3] < sl.partial_slots) ++slot_used[it[3]];
      got[idx].push_back({it[3], lanes});
    }
  }
  for (uint32_t i = 0; i < p.pieces(); ++i) check(hub_seen[i] == 1, tag + ": hub piece dealt once");
  for (size_t i = 0; i < reads.size(); ++i) check(reads[i] == (i < sl.slice_off[S] ? 1u : 0u), tag + ": sliced entry read once");
  // each wave's hub pieces come before its slice items
  for (uint32_t w = 0; w < waves; ++w) {
    bool slices_seen = false;
    for (uint32_t k = sl.wave_first[w]; k < sl.wave_first[w + 1]; ++k) {
      const bool hub = (sl.items[4 * static_cast<size_t>(k)] >> 30) == gc::kPrSliceHub;
      check(!(hub && slices_seen), tag + ": hub pieces first in a wave");
      slices_seen = slices_seen || !hub;
    }
  }
  for (uint32_t q = 0; q < sl.partial_slots; ++q) check(slot_used[q] == 1, tag + ": partial slot written once");
  std::vector<uint32_t> is_long(S, 0);
  for (uint32_t q = 0; q < sl.long_count(); ++q) {
    const uint32_t s = sl.long_slices[3 * q], p0 = sl.long_slices[3 * q + 1], np = sl.long_slices[3 * q + 2];
    check(s < S && whole[s] == 0 && got[s].size() == np, tag + ": wide slice pieces");
    if (s >= S) continue;
    is_long[s] = 1;
    std::sort(got[s].begin(), got[s].end(),
              [](const std::pair<uint32_t, std::vector<std::vector<uint16_t>>>& a,
                 const std::pair<uint32_t, std::vector<std::vector<uint16_t>>>& b) { return a.first < b.first; });
    for (uint32_t j = 0; j < got[s].size(); ++j) check(got[s][j].first == p0 + j, tag + ": wide slice slots in order");
  }
  for (uint32_t s = 0; s < S; ++s) {
    check(is_long[s] || (whole[s] == 1 && got[s].size() == 1), tag + ": every slice once");
    for (uint32_t l = 0; l < 64; ++l) {
      std::vector<uint16_t> seq;
      for (const auto& piece : got[s]) seq.insert(seq.end(), piece.second[l].begin(), piece.second[l].end());
      const uint64_t r = static_cast<uint64_t>(z) + 64ull * s + l;
      std::vector<uint16_t> want;
      if (r < pb) {
        for (uint32_t q = p.hoff[r]; q < p.hoff[r + 1]; ++q)
          if (p.hot16[q] != hot) want.push_back(p.hot16[q]);
      }
      std::vector<uint16_t> real;
      for (uint16_t t : seq) {
        if (t != hot) real.push_back(t);
      }
      check(real == want, tag + ": lane entries are the row's hot entries in order");
      // padding only after the row's entries
      bool pad = false, ok = true;
      for (uint16_t t : seq) {
        if (t == hot) pad = true;
        else if (pad) ok = false;
      }
      check(ok, tag + ": padding after the entries");
    }
  }
}

void run_case(const Csr& g, uint32_t push, uint32_t hot, uint32_t stage, uint32_t bin_rows, uint32_t bin_values) {
  const std::string tag = "n=" + std::to_string(g.n) + " push=" + std::to_string(push) + " hot=" +
                          std::to_string(hot) + " stage=" + std::to_string(stage) + " rows=" +
                          std::to_string(bin_rows) + " values=" + std::to_string(bin_values);
  gc::PrBinLayout p = gc::build_pr_bin_layout(g.off.data(), g.tgt.data(), g.n, g.zero_rows, push, hot, stage,
                                              bin_rows, bin_values, serial);
  const uint32_t n = g.n, z = p.zero_rows, pb = p.push_base, hb = p.hot_base;
  check(pb == n - push && hb == n - hot && z == std::min(g.zero_rows, pb), tag + ": bases");
  std::vector<uint64_t> val(n);
  for (uint32_t t = 0; t < n; ++t) val[t] = 1 + next_u64() % 1000003;

  // pr_bin_scatter
  std::vector<uint64_t> vals(p.bin_val.empty() ? 0 : p.bin_val.back(), 0);
  std::vector<uint32_t> written(vals.size(), 0);
  check(p.slots.size() == static_cast<size_t>(p.block_group.size()) * 64 * 8, tag + ": slot array size");
  uint64_t staged_total = 0;
  for (uint32_t s = 0; s < p.subs(); ++s) {
    std::vector<uint64_t> st(stage + 1, ~0ull);
    std::vector<uint32_t> hits(stage + 1, 0);
    for (uint32_t blk = p.sub[s]; blk < p.sub[s + 1]; ++blk) {
      for (uint32_t i = 0; i < 8; ++i) {
        const uint32_t gr = p.block_group[blk];
        for (uint32_t l = 0; l < 64; ++l) {
          const uint64_t t = static_cast<uint64_t>(z) + 64ull * gr + l;
          const uint64_t c = t < hb ? val[t] : 0;
          const uint16_t slot = p.slots[(static_cast<uint64_t>(blk) * 64 + l) * 8 + i];
          check(slot <= stage, tag + ": slot in range");
          if (slot > stage) continue;
          st[slot] = c;
          ++hits[slot];
        }
      }
    }
    uint32_t expect_from = 0;
    for (uint32_t r = p.run_first[s]; r < p.run_first[s + 1]; ++r) {
      const uint32_t from = p.run_info[r] & 0xFFFFu, len = p.run_info[r] >> 16;
      check(from == expect_from && len > 0, tag + ": runs tile the staging buffer");
      expect_from = from + len;
      for (uint32_t q = 0; q < len; ++q) {
        check(hits[from + q] == 1, tag + ": staged slot written once");
        const uint64_t d = static_cast<uint64_t>(p.run_dst[r]) + q;
        check(d < vals.size(), tag + ": run destination in range");
        if (d >= vals.size()) continue;
        vals[d] = st[from + q];
        ++written[d];
      }
    }
    check(expect_from <= stage, tag + ": sub-range within the stage");
    for (uint32_t q = expect_from; q < stage; ++q) check(hits[q] == 0, tag + ": no slot beyond the runs");
    staged_total += expect_from;
  }
  check(staged_total == p.cold_entries, tag + ": every cold entry staged");

  // pr_bin_accum
  std::vector<uint64_t> cold(pb - z, 0);
  for (uint32_t b = 0; b < p.bins(); ++b) {
    const uint32_t rows = p.bin_row[b + 1] - p.bin_row[b];
    check(rows >= 1 && rows <= bin_rows, tag + ": bin rows");
    check(p.bin_val[b] % 8 == 0 && p.bin_val[b + 1] % 8 == 0, tag + ": bin values aligned");
    std::vector<uint64_t> acc(rows + 1, 0);
    for (uint32_t v = p.bin_val[b]; v < p.bin_val[b + 1]; ++v) {
      check(p.uloc[v] <= rows, tag + ": row index in range");
      if (p.uloc[v] > rows) continue;
      if (p.uloc[v] == rows) {
        check(written[v] == 0, tag + ": padding value never written");
      } else {
        check(written[v] == 1, tag + ": value written once");
      }
      acc[p.uloc[v]] += vals[v];
    }
    for (uint32_t i = 0; i < rows; ++i) cold[p.bin_row[b] - z + i] = acc[i];
  }
  // hub pieces: every entry of a push-set row that points into the push set, once
  {
    std::vector<uint32_t> seen(g.tgt.size(), 0);
    for (uint32_t i = 0; i < p.pieces(); ++i) {
      const uint32_t h = p.hub_pieces[2 * i], b = p.hub_pieces[2 * i + 1];
      check(h >= pb && h < n && b >= g.off[h] && b < g.off[h + 1], tag + ": hub piece in its row");
      if (!(h >= pb && h < n && b >= g.off[h] && b < g.off[h + 1])) continue;
      for (uint32_t q = b; q < std::min(b + 512, g.off[h + 1]); ++q) ++seen[q];
    }
    for (uint32_t h = pb; h < n; ++h)
      for (uint32_t q = g.off[h]; q < g.off[h + 1]; ++q)
        check(seen[q] == (g.tgt[q] >= pb ? 1u : 0u), tag + ": hub piece coverage");
  }
  // pr_spmv_bin
  check(p.hoff.size() == static_cast<size_t>(pb) + 1 && p.hoff[pb] == p.hot16.size(), tag + ": stream bounds");
  for (uint32_t r = z; r < pb; ++r) {
    check(p.hoff[r + 1] > p.hoff[r], tag + ": every pull row has a stream entry");
    uint64_t hot_sum = 0;
    for (uint32_t q = p.hoff[r]; q < p.hoff[r + 1]; ++q) {
      const uint32_t t = p.hot16[q];
      check(t <= hot, tag + ": stream entry in range");
      if (t == hot) check(p.hoff[r + 1] - p.hoff[r] == 1, tag + ": the zero entry stands alone");
      else hot_sum += val[hb + t];
    }
    uint64_t expect = 0;
    for (uint32_t q = g.off[r]; q < g.off[r + 1]; ++q) expect += val[g.tgt[q]];
    check(cold[r - z] + hot_sum == expect, tag + ": row " + std::to_string(r) + " sum");
  }
  if (stage == 512 && bin_rows == 16384) {
    const uint32_t caps[] = {1, 2, 4};
    const uint32_t wave_counts[] = {1, 7, 4096};
    for (uint32_t cap : caps)
      for (uint32_t waves : wave_counts) check_slices(g, p, hot, cap, waves, tag);
  }
}

}  // namespace

int main() {
  struct Shape {
    uint32_t n, edges, hubs;
    bool loops;
  };
  const Shape shapes[] = {{1, 0, 0, false},     {5, 3, 0, true},      {64, 200, 4, false},
                          {300, 900, 10, true}, {2000, 12000, 30, false}, {5000, 60000, 64, true}};
  for (const Shape& sh : shapes) {
    const Csr g = random_graph(sh.n, sh.edges, sh.hubs, sh.loops);
    const uint32_t n = g.n;
    const uint32_t pushes[] = {1, 3, 17, 6144};
    const uint32_t hots[] = {0, 40, 700, 26624};  // 0: the push set itself
    for (uint32_t push : pushes) {
      if (push > n) continue;
      for (uint32_t hot : hots) {
        hot = std::max(push, std::min(hot, n));
        const uint32_t stages[] = {512, 1000, 36864};
        for (uint32_t stage : stages) {
          run_case(g, push, hot, stage, 16384, 65536);
          run_case(g, push, hot, stage, 1, 8);
          run_case(g, push, hot, stage, 7, 100);
        }
      }
    }
  }
  // argument checks
  const Csr g = random_graph(100, 300, 2, false);
  bool threw = false;
  try {
    gc::build_pr_bin_layout(g.off.data(), g.tgt.data(), g.n, g.zero_rows, 4, 2, 512, 16, 64, serial);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  check(threw, "hot set smaller than the push set is refused");
  threw = false;
  try {
    gc::build_pr_bin_layout(g.off.data(), g.tgt.data(), g.n, g.zero_rows, 4, 8, 100, 16, 64, serial);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  check(threw, "a stage below one block of slots is refused");
  // a CSR that is not symmetric is refused
  Csr d;
  d.n = 4;  // row 1 -> 2 and row 2 -> 0: equal cold counts by rows and by sources
  d.off = {0, 0, 1, 2, 2};
  d.tgt = {2, 0};
  d.zero_rows = 1;
  threw = false;
  try {
    gc::build_pr_bin_layout(d.off.data(), d.tgt.data(), d.n, d.zero_rows, 1, 1, 512, 16, 64, serial);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  check(threw, "a non-symmetric CSR is refused");
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("pr_bin layout tests passed\n");
  return 0;
}
