// Launch geometry for wave64 kernels.
//
// Replaces launch::{thread_per_item, warp_per_item}
// (crates/grafeo-cuda-algos/src/launch.rs:7-41). The reference sizes a grid
// with one thread (or one 32-lane warp) per item and refuses inputs whose
// thread count overflows u32. Here every kernel indexes work with 64-bit
// indices in a grid-stride loop, so a grid is capped instead of refused: the
// cap keeps enough workgroups to fill all 256 CUs many times over, and any
// item count, including RMAT-24's 2^29 CSR entries, runs in one launch.
// Header-only and free of HIP types so a host-only test can check it.
#pragma once

#include <cstdint>

namespace gc {

constexpr uint32_t kWaveSize = 64;          // CDNA wavefront width
constexpr uint32_t kBlockThreads = 256;     // four waves per workgroup
constexpr uint32_t kMaxGridBlocks = 1u << 20;  // grid-stride cap

struct Grid1D {
  uint32_t blocks;
  uint32_t threads;
};

// One thread per item, grid-stride beyond kMaxGridBlocks workgroups.
// `items` == 0 gives a zero-block grid, which callers must not launch.
inline Grid1D thread_per_item(uint64_t items, uint32_t threads = kBlockThreads) {
  uint64_t blocks = (items + threads - 1) / threads;
  if (blocks > kMaxGridBlocks) blocks = kMaxGridBlocks;
  return Grid1D{static_cast<uint32_t>(blocks), threads};
}

// One 64-lane wave per item, `waves_per_block` waves per workgroup.
inline Grid1D wave_per_item(uint64_t items, uint32_t waves_per_block = kBlockThreads / kWaveSize) {
  uint64_t blocks = (items + waves_per_block - 1) / waves_per_block;
  if (blocks > kMaxGridBlocks) blocks = kMaxGridBlocks;
  return Grid1D{static_cast<uint32_t>(blocks), waves_per_block * kWaveSize};
}

// Total threads a grid covers (64-bit, no wrap-around).
inline uint64_t grid_threads(const Grid1D& g) {
  return static_cast<uint64_t>(g.blocks) * g.threads;
}

}  // namespace gc
