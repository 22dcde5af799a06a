// Host construction of the sliced layout of PageRank's 16-bit hot stream (the
// static arrays pr_spmv_slice of pagerank.hip reads). Plain C++, no HIP: the
// library uploads the result, and a host-only unit test replays the kernel's
// indexing on it.
//
// Input: the pull rows [z, push_base) with their hot streams (hoff, hot16 of
// pr_bin_layout.hpp; a row whose only entry is hot_count has none). Slice s
// holds rows z + 64 s + l, one per lane l; its width W is the largest entry
// count of its rows rounded up to even, split as G = W / 8 groups and a
// remainder of R = W % 8 entries per lane (0, 2, 4 or 6). From slice_off[s]:
//   group g:   [512 g + 8 l + i] = entry 8 g + i of lane l's row,
//   remainder: [512 G + R l + i] = entry 8 G + i of lane l's row,
// with hot_count (which reads a zero in the kernel) past a row's end or past
// the last pull row. A wave takes each group with one coalesced 16-byte load
// per lane and the remainder with one 16-byte load per lane (4-byte aligned,
// its first R entries used; the array carries 1024 entries of padding past
// the last slice for these and for the clamped group loads of narrow
// slices), and sums its own row with no exchange between lanes.
//
// Items (four words each): kind << 30 | index, then
//   slice, whole (kind 0): first entry of the slice, G | R << 16, unused;
//   slice piece (kind 1): first entry of its groups, groups | R << 16,
//     partial slot (64 doubles written there; the last piece of a slice takes
//     its remainder, R = 0 on the others);
//   hub piece (kind 2): hub row, first entry, first of its 512 entries in
//     hub16 (the piece's entries as offsets from push_base, 0xFFFF past the
//     piece's end).
// Slices wider than cap_groups groups are cut into pieces of cap_groups
// groups; long_slices lists (slice, first partial slot, pieces) for the
// finishing kernel, which adds a row's pieces in order. Items are dealt to
// `waves` waves by longest processing time first (largest item to the wave
// with the least work so far, by a cost of entries per lane plus a fixed cost
// per item), each wave's items in [wave_first[w], wave_first[w + 1]): its hub
// pieces first, then its slice items in the order dealt.
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <queue>
#include <stdexcept>
#include <utility>
#include <vector>

namespace gc {

struct PrSliceLayout {
  uint32_t slices = 0, partial_slots = 0;
  std::vector<uint32_t> slice_off;    // [slices + 1] entry offsets
  std::vector<uint16_t> data;         // [slice_off.back() + kPrSlicePad]
  std::vector<uint32_t> items;        // 4 words per item
  std::vector<uint32_t> wave_first;   // [waves + 1]
  std::vector<uint32_t> long_slices;  // (slice, first partial slot, pieces) triples
  std::vector<uint16_t> hub16;        // [hub pieces * 512]
  uint32_t item_count() const { return static_cast<uint32_t>(items.size() / 4); }
  uint32_t long_count() const { return static_cast<uint32_t>(long_slices.size() / 3); }
};

constexpr uint32_t kPrSliceWhole = 0, kPrSlicePiece = 1, kPrSliceHub = 2;
constexpr uint32_t kPrSlicePad = 1024;  // entries past the last slice
constexpr uint16_t kPrHubNone = 0xFFFF;  // a hub piece's entry past its end

// hub_pieces holds (hub row, first entry) pairs of the CSR (off, tgt): up to
// 512 entries from there, in the row, all of them at or above push_base.
template <class ParFor>
PrSliceLayout build_pr_slice_layout(const uint32_t* hoff, const uint16_t* hot16, uint32_t z, uint32_t push_base,
                                    uint32_t hot_count, const std::vector<uint32_t>& hub_pieces,
                                    const uint32_t* off, const uint32_t* tgt, uint32_t cap_groups, uint32_t waves,
                                    ParFor par) {
  if (cap_groups == 0 || cap_groups >= 0xFFFFu || waves == 0 || hot_count >= 0xFFFFu || push_base < z ||
      hub_pieces.size() % 2 != 0)
    throw std::runtime_error("PageRank: sliced layout sizes out of range");
  PrSliceLayout p;
  const size_t pieces = hub_pieces.size() / 2;
  std::vector<uint32_t> hub_len(pieces);
  p.hub16.assign(pieces * 512, kPrHubNone);
  for (size_t i = 0; i < pieces; ++i) {
    const uint32_t h = hub_pieces[2 * i], b = hub_pieces[2 * i + 1];
    hub_len[i] = std::min(b + 512u, off[h + 1]) - b;
    for (uint32_t j = 0; j < hub_len[i]; ++j) {
      const uint32_t t = tgt[b + j];
      if (t < push_base || t - push_base >= kPrHubNone)
        throw std::runtime_error("PageRank: a hub piece's entry lies outside the push set");
      p.hub16[512 * i + j] = static_cast<uint16_t>(t - push_base);
    }
  }
  const uint32_t pull = push_base - z;
  const uint32_t S = (pull + 63) / 64;
  p.slices = S;
  auto count_of = [&](uint32_t r) -> uint32_t {
    const uint32_t b = hoff[r], e = hoff[r + 1];
    return (e - b == 1 && hot16[b] == hot_count) ? 0u : e - b;
  };
  std::vector<uint32_t> width(S, 0);
  par(S, [&](uint64_t b, uint64_t e) {
    for (uint64_t s = b; s < e; ++s) {
      uint32_t w = 0;
      for (uint32_t l = 0; l < 64; ++l) {
        const uint64_t r = static_cast<uint64_t>(z) + 64 * s + l;
        if (r < push_base) w = std::max(w, count_of(static_cast<uint32_t>(r)));
      }
      width[s] = w + (w & 1u);
    }
  });
  auto groups_of = [&](uint32_t s) { return width[s] / 8; };
  auto rem_of = [&](uint32_t s) { return width[s] % 8; };
  p.slice_off.assign(static_cast<size_t>(S) + 1, 0);
  uint64_t total = 0;
  for (uint32_t s = 0; s < S; ++s) {
    p.slice_off[s] = static_cast<uint32_t>(total);
    total += 512ull * groups_of(s) + 64ull * rem_of(s);
    if (total >= 0xFFFFFFFFull) throw std::runtime_error("PageRank: sliced stream too long");
  }
  p.slice_off[S] = static_cast<uint32_t>(total);
  p.data.assign(total + kPrSlicePad, static_cast<uint16_t>(hot_count));
  par(S, [&](uint64_t b, uint64_t e) {
    for (uint64_t s = b; s < e; ++s) {
      const uint32_t G = groups_of(static_cast<uint32_t>(s)), R = rem_of(static_cast<uint32_t>(s));
      uint16_t* out = p.data.data() + p.slice_off[s];
      for (uint32_t l = 0; l < 64; ++l) {
        const uint64_t r = static_cast<uint64_t>(z) + 64 * s + l;
        if (r >= push_base) break;
        const uint32_t c = count_of(static_cast<uint32_t>(r));
        const uint16_t* in = hot16 + hoff[r];
        for (uint32_t j = 0; j < c; ++j) {
          const uint64_t at = j < 8 * G ? 512ull * (j / 8) + 8 * l + j % 8 : 512ull * G + R * l + (j - 8 * G);
          out[at] = in[j];
        }
      }
    }
  });

  // items, with their costs: entries per lane plus a fixed cost
  struct Item {
    uint32_t w[4];
    uint32_t cost;
  };
  std::vector<Item> all;
  all.reserve(static_cast<size_t>(S) + hub_len.size());
  constexpr uint32_t kWholeCost = 24, kPieceCost = 12, kHubCost = 12;
  for (uint32_t s = 0; s < S; ++s) {
    const uint32_t G = groups_of(s), R = rem_of(s);
    if (G <= cap_groups) {
      all.push_back({{kPrSliceWhole << 30 | s, p.slice_off[s], G | R << 16, 0}, kWholeCost + 8 * G + R});
      continue;
    }
    const uint32_t pieces = (G + cap_groups - 1) / cap_groups;
    p.long_slices.push_back(s);
    p.long_slices.push_back(p.partial_slots);
    p.long_slices.push_back(pieces);
    for (uint32_t q = 0; q < pieces; ++q) {
      const uint32_t g0 = q * cap_groups, ng = std::min(cap_groups, G - g0);
      const uint32_t r = q + 1 == pieces ? R : 0;
      all.push_back({{kPrSlicePiece << 30 | s, p.slice_off[s] + 512 * g0, ng | r << 16, p.partial_slots++},
                     kPieceCost + 8 * ng + r});
    }
  }
  for (size_t i = 0; i < hub_len.size(); ++i) {
    all.push_back({{kPrSliceHub << 30 | static_cast<uint32_t>(i), hub_pieces[2 * i], hub_pieces[2 * i + 1],
                    static_cast<uint32_t>(512 * i)},
                   kHubCost + (hub_len[i] + 63) / 64});
  }
  if (all.size() >= (1u << 30)) throw std::runtime_error("PageRank: too many sliced items");
  // longest processing time first; ties keep the items' order
  std::stable_sort(all.begin(), all.end(), [](const Item& a, const Item& b) { return a.cost > b.cost; });
  using Load = std::pair<uint64_t, uint32_t>;
  std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
  for (uint32_t w = 0; w < waves; ++w) heap.push({0, w});
  std::vector<uint32_t> owner(all.size());
  std::vector<uint32_t> per(waves, 0);
  for (size_t i = 0; i < all.size(); ++i) {
    Load top = heap.top();
    heap.pop();
    owner[i] = top.second;
    ++per[top.second];
    top.first += all[i].cost;
    heap.push(top);
  }
  p.wave_first.assign(static_cast<size_t>(waves) + 1, 0);
  for (uint32_t w = 0; w < waves; ++w) p.wave_first[w + 1] = p.wave_first[w] + per[w];
  p.items.assign(4 * all.size(), 0);
  std::vector<uint32_t> at(p.wave_first.begin(), p.wave_first.end() - 1);
  for (int hubs = 1; hubs >= 0; --hubs) {  // each wave's hub pieces first
    for (size_t i = 0; i < all.size(); ++i) {
      if (((all[i].w[0] >> 30) == kPrSliceHub) != (hubs == 1)) continue;
      const uint32_t k = at[owner[i]]++;
      for (int j = 0; j < 4; ++j) p.items[4 * static_cast<size_t>(k) + j] = all[i].w[j];
    }
  }
  return p;
}

}  // namespace gc
