// Host-only unit tests of the module cache and the launch geometry, ported
// from crates/grafeo-cuda-algos/src/module_cache.rs:89-134 (PerContext with a
// closure loader) and src/launch.rs:43-74 (grid rounding, large inputs).
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "launch.hpp"
#include "module_cache.hpp"

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static void loads_once_per_key() {
  gc::PerDevice<uint32_t> cache;
  int loads = 0;
  auto load = [&](uint32_t v) {
    return [&, v] {
      ++loads;
      return v;
    };
  };
  CHECK(cache.get_or_try_insert(1, load(3)) == 3);
  CHECK(cache.get_or_try_insert(1, load(19)) == 3);
  CHECK(loads == 1);
  // A second device gets its own entry instead of the first one's.
  CHECK(cache.get_or_try_insert(2, load(88)) == 88);
  CHECK(cache.get_or_try_insert(1, load(0)) == 3);
  CHECK(cache.get_or_try_insert(2, load(0)) == 88);
  CHECK(loads == 2);
  CHECK(cache.size() == 2);
}

static void failed_load_is_not_cached() {
  gc::PerDevice<uint32_t> cache;
  bool threw = false;
  try {
    cache.get_or_try_insert(7, []() -> uint32_t { throw std::runtime_error("no driver"); });
  } catch (const std::runtime_error&) {
    threw = true;
  }
  CHECK(threw);
  CHECK(cache.size() == 0);
  CHECK(cache.get_or_try_insert(7, [] { return 3u; }) == 3);
}

static void thread_per_item_rounds_up_to_whole_blocks() {
  CHECK(gc::thread_per_item(1).blocks == 1);
  CHECK(gc::thread_per_item(256).blocks == 1);
  CHECK(gc::thread_per_item(257).blocks == 2);
  CHECK(gc::thread_per_item(3).threads == gc::kBlockThreads);
  CHECK(gc::thread_per_item(0).blocks == 0);
}

static void large_inputs_are_capped_not_refused() {
  // Regression of the reference: items * 32 wrapped in u32 and launched too few
  // warps. Here work indices are 64-bit and grids are grid-stride.
  uint64_t items = 0xFFFFFFFFull * 4;
  gc::Grid1D g = gc::thread_per_item(items);
  CHECK(g.blocks == gc::kMaxGridBlocks);
  gc::Grid1D w = gc::wave_per_item(items);
  CHECK(w.blocks == gc::kMaxGridBlocks);
  CHECK(w.threads == gc::kBlockThreads);
  CHECK(gc::wave_per_item(4).blocks == 1);
  CHECK(gc::wave_per_item(5).blocks == 2);
  CHECK(gc::grid_threads(gc::wave_per_item(5)) == 512);
}

int main() {
  loads_once_per_key();
  failed_load_is_not_cached();
  thread_per_item_rounds_up_to_whole_blocks();
  large_inputs_are_capped_not_refused();
  if (failures == 0) std::printf("all native unit tests passed\n");
  return failures == 0 ? 0 : 1;
}
