// Per-device cache of loaded code objects.
//
// Replaces module_cache::{PerContext, get_or_load}
// (crates/grafeo-cuda-algos/src/module_cache.rs:23-87): a code object is
// loaded once per device under a mutex and shared by every later call on that
// device; a second device gets its own entry (a module belongs to the device
// that was current when it was loaded); a failed load caches nothing, so the
// next call tries again. Header-only and free of HIP types so a host-only test
// can drive it with a fake loader.
#pragma once

#include <mutex>
#include <utility>
#include <vector>

namespace gc {

template <class T>
class PerDevice {
 public:
  // Returns the value cached for `key`, running `load()` on first use. If
  // `load` throws, nothing is cached and the exception propagates.
  template <class Load>
  T get_or_try_insert(int key, Load&& load) {
    std::lock_guard<std::mutex> lock(mu_);
    for (auto& entry : entries_) {
      if (entry.first == key) return entry.second;
    }
    T value = load();
    entries_.emplace_back(key, value);
    return value;
  }

  size_t size() {
    std::lock_guard<std::mutex> lock(mu_);
    return entries_.size();
  }

 private:
  std::mutex mu_;
  std::vector<std::pair<int, T>> entries_;
};

}  // namespace gc
