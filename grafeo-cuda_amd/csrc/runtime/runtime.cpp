// Errors, host threads, devices, code-object loading, kernel timing and
// device memory.
//
// Reference counterparts: crates/grafeo-cuda-core/src/error.rs (error
// categories and their Display text), crates/grafeo-cuda-core/src/device.rs
// (GpuDevice), crates/grafeo-cuda-algos/src/module_cache.rs (load once per
// device, failed loads not cached).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

#include "code_objects.hpp"
#include "common.hpp"
#include "module_cache.hpp"
#include "noalloc.hpp"

namespace gc {

// ---------------------------------------------------------------- errors --

namespace {
thread_local std::string t_last_error;

std::string display(gc_status status, const std::string& detail) {
  switch (status) {
    case GC_ERR_NO_DEVICE: return "no GPU device found";
    case GC_ERR_DRIVER: return "HIP runtime error: " + detail;
    case GC_ERR_OUT_OF_MEMORY: return detail;
    case GC_ERR_KERNEL: return "kernel error: " + detail;
    case GC_ERR_EXPORT: return "CSR export failed: " + detail;
    case GC_ERR_INVALID_CSR: return "invalid CSR: " + detail;
    default: return detail;
  }
}
}  // namespace

void fail(gc_status status, const std::string& detail) {
  throw Error{status, display(status, detail)};
}

void set_last_error(const std::string& message) { t_last_error = message; }

void hip_check(hipError_t err, const char* what) {
  if (err == hipSuccess) return;
  if (err == hipErrorOutOfMemory) {
    size_t free_b = 0, total_b = 0;
    (void)hipGetLastError();
    (void)hipMemGetInfo(&free_b, &total_b);
    fail(GC_ERR_OUT_OF_MEMORY, std::string("out of device memory in ") + what + ": " +
                                   std::to_string(free_b) + " bytes available");
  }
  if (err == hipErrorNoDevice) fail(GC_ERR_NO_DEVICE, "");
  fail(GC_ERR_DRIVER, std::string(what) + ": " + hipGetErrorString(err));
}

gc_status guard(const std::function<void()>& body) {
  try {
    body();
    return GC_OK;
  } catch (const Error& e) {
    set_last_error(e.message);
    return e.status;
  } catch (const std::bad_alloc&) {
    set_last_error("host out of memory");
    return GC_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    set_last_error(std::string("kernel error: ") + e.what());
    return GC_ERR_KERNEL;
  }
}

// ------------------------------------------------------- host parallelism --

unsigned host_threads() {
  static unsigned cached = [] {
    unsigned hw = std::thread::hardware_concurrency();
    if (hw == 0) hw = 1;
    unsigned t = std::min(hw, 16u);
    if (const char* env = std::getenv("GC_HOST_THREADS")) {
      long v = std::strtol(env, nullptr, 10);
      if (v > 0 && v <= 256) t = static_cast<unsigned>(v);
    }
    return t;
  }();
  return cached;
}

void parallel_for(uint64_t n, const std::function<void(uint64_t, uint64_t)>& fn,
                  uint64_t min_chunk) {
  if (n == 0) return;
  unsigned t = host_threads();
  uint64_t chunks = std::min<uint64_t>(t, (n + min_chunk - 1) / min_chunk);
  if (chunks <= 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(chunks);
  std::exception_ptr first_error;
  std::mutex err_mu;
  for (uint64_t c = 0; c < chunks; ++c) {
    uint64_t b = n * c / chunks, e = n * (c + 1) / chunks;
    pool.emplace_back([&, b, e] {
      try {
        fn(b, e);
      } catch (...) {
        std::lock_guard<std::mutex> lock(err_mu);
        if (!first_error) first_error = std::current_exception();
      }
    });
  }
  for (auto& th : pool) th.join();
  if (first_error) std::rethrow_exception(first_error);
}

// ---------------------------------------------------------------- device --

void Device::make_current() const { hip_check(hipSetDevice(ordinal), "hipSetDevice"); }

void Device::sync() const { hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

// One loaded code object (module) and its resolved kernels.
struct LoadedModule {
  hipModule_t module = nullptr;
  std::map<std::string, hipFunction_t> functions;
  std::mutex mu;
};

namespace {

const CodeObject* find_code_object(const char* family) {
  for (uint32_t i = 0; i < gc_embedded_code_object_count; ++i) {
    if (std::strcmp(gc_embedded_code_objects[i].name, family) == 0) {
      return &gc_embedded_code_objects[i];
    }
  }
  return nullptr;
}

// One cache per code-object family, keyed by device ordinal.
PerDevice<std::shared_ptr<LoadedModule>>& cache_for(const char* family) {
  static std::mutex mu;
  static std::map<std::string, std::unique_ptr<PerDevice<std::shared_ptr<LoadedModule>>>> caches;
  std::lock_guard<std::mutex> lock(mu);
  auto& slot = caches[family];
  if (!slot) slot.reset(new PerDevice<std::shared_ptr<LoadedModule>>());
  return *slot;
}

}  // namespace

Kernel get_kernel(Device& dev, const char* family, const char* name) {
  auto module = cache_for(family).get_or_try_insert(dev.ordinal, [&] {
    const CodeObject* co = find_code_object(family);
    if (co == nullptr) fail_kernel(std::string("no embedded code object named ") + family);
    dev.make_current();
    auto loaded = std::make_shared<LoadedModule>();
    hipError_t err = hipModuleLoadData(&loaded->module, co->begin);
    if (err != hipSuccess) {
      fail_kernel(std::string("failed to load gfx950 code object ") + family + ": " +
                  hipGetErrorString(err));
    }
    return loaded;
  });
  std::lock_guard<std::mutex> lock(module->mu);
  auto it = module->functions.find(name);
  if (it == module->functions.end()) {
    hipFunction_t fn = nullptr;
    hipError_t err = hipModuleGetFunction(&fn, module->module, name);
    if (err != hipSuccess) {
      fail_kernel(std::string("kernel ") + name + " not found in " + family + ": " +
                  hipGetErrorString(err));
    }
    it = module->functions.emplace(name, fn).first;
  }
  return Kernel{it->second, it->first.c_str()};
}

void noalloc_setup(Device& dev) {
  dev.make_current();
  void* ptr = nullptr;
  const hipError_t err = noalloc_scratch_address(&ptr);
  if (err != hipSuccess || ptr == nullptr) {
    (void)hipGetLastError();
    dev.noalloc_error = std::string("hipGetSymbolAddress of the static scratch failed: ") + hipGetErrorString(err);
    return;
  }
  try {
    // the kernels of gc_triangle_count_noalloc, loaded now so that it loads nothing
    (void)get_kernel(dev, "util", "fill_u32");
    (void)get_kernel(dev, "tc", "tc_bitmap16");
    (void)get_kernel(dev, "tc", "tc_bitmap16_w4");
    (void)get_kernel(dev, "tc", "tc_bitmap");
  } catch (const Error& e) {
    dev.noalloc_error = e.message;
    return;
  }
  dev.noalloc_scratch = ptr;
}

// ------------------------------------------------------------- profiling --

void profile_begin(Device& dev, const char* name, hipEvent_t* start) {
  if (!dev.profiling) return;
  if (!dev.profile_only.empty() &&
      std::find(dev.profile_only.begin(), dev.profile_only.end(), name) == dev.profile_only.end())
    return;
  hipEvent_t ev;
  if (!dev.event_pool.empty()) {
    ev = dev.event_pool.back();
    dev.event_pool.pop_back();
  } else {
    hip_check(hipEventCreate(&ev), "hipEventCreate");
  }
  hip_check(hipEventRecord(ev, dev.stream), "hipEventRecord");
  *start = ev;
}

void profile_end(Device& dev, const char* name, hipEvent_t start) {
  if (!dev.profiling || start == nullptr) return;
  hipEvent_t ev;
  if (!dev.event_pool.empty()) {
    ev = dev.event_pool.back();
    dev.event_pool.pop_back();
  } else {
    hip_check(hipEventCreate(&ev), "hipEventCreate");
  }
  hip_check(hipEventRecord(ev, dev.stream), "hipEventRecord");
  dev.pending.push_back(ProfileRecord{name, start, ev});
}

static void profile_collect(Device& dev) {
  if (dev.pending.empty()) return;
  dev.sync();
  for (auto& rec : dev.pending) {
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, rec.start, rec.stop), "hipEventElapsedTime");
    auto it = std::find_if(dev.totals.begin(), dev.totals.end(),
                           [&](const ProfileTotal& t) { return t.name == rec.name; });
    if (it == dev.totals.end()) {
      dev.totals.push_back(ProfileTotal{rec.name, 0.0, 0});
      it = dev.totals.end() - 1;
    }
    it->total_ms += ms;
    it->launches += 1;
    dev.event_pool.push_back(rec.start);
    dev.event_pool.push_back(rec.stop);
  }
  dev.pending.clear();
}

// --------------------------------------------------------- device memory --

void DeviceMemory::allocate(size_t bytes) {
  release();
  if (bytes == 0) return;
  void* p = nullptr;
  hipError_t err = hipMalloc(&p, bytes);
  if (err == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    fail(GC_ERR_OUT_OF_MEMORY, "graph too large for GPU memory: " + std::to_string(bytes) +
                                   " bytes required, " + std::to_string(free_b) +
                                   " bytes available");
  }
  hip_check(err, "hipMalloc");
  ptr_ = p;
  bytes_ = bytes;
}

void DeviceMemory::ensure(size_t bytes) {
  if (bytes <= bytes_ && ptr_ != nullptr) return;
  allocate(bytes);
}

void DeviceMemory::release() {
  if (ptr_ != nullptr) {
    (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
  }
}

void fill_words(Device& dev, void* p, uint64_t words, uint32_t value) {
  if (words == 0) return;
  Kernel k = get_kernel(dev, "util", "fill_u32");
  Grid1D grid = thread_per_item(words);
  launch(dev, k, grid.blocks, grid.threads, 0, static_cast<uint32_t*>(p), words, value);
}

uint64_t scan_scratch_words(uint64_t count) {
  constexpr uint64_t kPerBlock = 256 * 16;  // SCAN_PER_BLOCK of util.hip
  return (count + kPerBlock - 1) / kPerBlock + 1;
}

void exclusive_scan_u32(Device& dev, uint32_t* data, uint64_t count, uint32_t* block_buf) {
  if (count == 0) return;
  const uint32_t blocks = static_cast<uint32_t>(scan_scratch_words(count) - 1);
  launch(dev, get_kernel(dev, "util", "scan_block_sums"), blocks, 256u, 0,
         static_cast<const uint32_t*>(data), count, block_buf);
  launch(dev, get_kernel(dev, "util", "exclusive_scan_small"), 1u, 1024u, 0, block_buf, blocks,
         block_buf + blocks);
  launch(dev, get_kernel(dev, "util", "scan_apply"), blocks, 256u, 0, data, count,
         static_cast<const uint32_t*>(block_buf));
}

uint32_t wedge_table_bits(const Relabelled& rl) {
  uint32_t bits = 6;
  while (bits < 12 && (uint64_t(1) << bits) < 2ull * rl.max_out_degree) ++bits;
  return bits;
}

void copy_to_device(Device& dev, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, dev.stream), "hipMemcpyAsync H2D");
  dev.sync();
}

void copy_to_host(Device& dev, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, dev.stream), "hipMemcpyAsync D2H");
  dev.sync();
}

}  // namespace gc

// ============================================================ C boundary ==

using gc::guard;

extern "C" {

const char* gc_last_error(void) { return gc::t_last_error.c_str(); }

void gc_free(void* ptr) { std::free(ptr); }

gc_status gc_device_count(uint32_t* count) {
  return guard([&] {
    if (count == nullptr) gc::fail_kernel("count is null");
    int c = 0;
    hipError_t err = hipGetDeviceCount(&c);
    if (err != hipSuccess) {
      (void)hipGetLastError();
      c = 0;
    }
    *count = static_cast<uint32_t>(c);
  });
}

gc_status gc_device_new(uint32_t ordinal, gc_device** out) {
  return guard([&] {
    if (out == nullptr) gc::fail_kernel("out is null");
    *out = nullptr;
    int count = 0;
    hipError_t err = hipGetDeviceCount(&count);
    if (err != hipSuccess || count == 0) {
      (void)hipGetLastError();
      gc::fail(GC_ERR_NO_DEVICE, "");
    }
    if (ordinal >= static_cast<uint32_t>(count)) {
      gc::fail(GC_ERR_DRIVER, "invalid device ordinal " + std::to_string(ordinal) + " (" +
                                  std::to_string(count) + " devices)");
    }
    std::unique_ptr<gc::Device> dev(new gc::Device());
    dev->ordinal = static_cast<int>(ordinal);
    dev->make_current();
    hipDeviceProp_t prop;
    gc::hip_check(hipGetDeviceProperties(&prop, dev->ordinal), "hipGetDeviceProperties");
    dev->name = prop.name;
    if (dev->name.empty()) {
      char buf[256] = {0};
      if (hipDeviceGetName(buf, sizeof(buf) - 1, dev->ordinal) == hipSuccess) dev->name = buf;
    }
    if (dev->name.empty()) dev->name = std::string("AMD ") + prop.gcnArchName;
    dev->total_memory = prop.totalGlobalMem;
    dev->cu_count = prop.multiProcessorCount;
    dev->max_lds_per_block = prop.sharedMemPerBlock;
    gc::hip_check(hipStreamCreateWithFlags(&dev->stream, hipStreamNonBlocking), "hipStreamCreate");
    gc::hip_check(hipHostMalloc(&dev->pinned, gc::Device::kPinnedBytes, hipHostMallocDefault),
                  "hipHostMalloc");
    std::memset(dev->pinned, 0, gc::Device::kPinnedBytes);
    gc::hip_check(hipHostGetDevicePointer(&dev->pinned_device, dev->pinned, 0), "hipHostGetDevicePointer");
    gc::noalloc_setup(*dev);
    *out = reinterpret_cast<gc_device*>(dev.release());
  });
}

gc_device* gc_device_retain(gc_device* handle) {
  if (handle != nullptr) reinterpret_cast<gc::Device*>(handle)->refs.fetch_add(1);
  return handle;
}

void gc_device_release(gc_device* handle) {
  if (handle == nullptr) return;
  auto* dev = reinterpret_cast<gc::Device*>(handle);
  if (dev->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(dev->ordinal);
  if (dev->stream) {
    (void)hipStreamSynchronize(dev->stream);
    (void)hipStreamDestroy(dev->stream);
  }
  for (auto& rec : dev->pending) {
    (void)hipEventDestroy(rec.start);
    (void)hipEventDestroy(rec.stop);
  }
  for (auto ev : dev->event_pool) (void)hipEventDestroy(ev);
  if (dev->pinned) (void)hipHostFree(dev->pinned);
  delete dev;
}

gc_status gc_device_properties(const gc_device* handle, char* name, size_t name_cap,
                               uint64_t* total_memory, uint32_t* ordinal) {
  return guard([&] {
    if (handle == nullptr) gc::fail_kernel("device is null");
    auto* dev = reinterpret_cast<const gc::Device*>(handle);
    if (name != nullptr && name_cap > 0) {
      size_t len = std::min(name_cap - 1, dev->name.size());
      std::memcpy(name, dev->name.data(), len);
      name[len] = '\0';
    }
    if (total_memory) *total_memory = dev->total_memory;
    if (ordinal) *ordinal = static_cast<uint32_t>(dev->ordinal);
  });
}

gc_status gc_device_malloc(gc_device* handle, size_t bytes, void** ptr) {
  return guard([&] {
    if (handle == nullptr || ptr == nullptr) gc::fail_kernel("device or ptr is null");
    auto* dev = reinterpret_cast<gc::Device*>(handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    *ptr = nullptr;
    gc::hip_check(hipMalloc(ptr, bytes == 0 ? 1 : bytes), "hipMalloc");
  });
}

gc_status gc_device_free(gc_device* handle, void* ptr) {
  return guard([&] {
    if (handle == nullptr) gc::fail_kernel("device is null");
    auto* dev = reinterpret_cast<gc::Device*>(handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    if (ptr != nullptr) gc::hip_check(hipFree(ptr), "hipFree");
  });
}

gc_status gc_device_copy_to_host(gc_device* handle, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    if (handle == nullptr) gc::fail_kernel("device is null");
    auto* dev = reinterpret_cast<gc::Device*>(handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    gc::copy_to_host(*dev, dst, src, bytes);
  });
}

gc_status gc_profile_enable(gc_device* handle, int32_t on) {
  return guard([&] {
    if (handle == nullptr) gc::fail_kernel("device is null");
    auto* dev = reinterpret_cast<gc::Device*>(handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    gc::profile_collect(*dev);
    dev->totals.clear();
    dev->profiling = on != 0;
  });
}

gc_status gc_profile_only(gc_device* handle, const char* names) {
  return guard([&] {
    if (handle == nullptr) gc::fail_kernel("device is null");
    auto* dev = reinterpret_cast<gc::Device*>(handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->profile_only.clear();
    if (names == nullptr) return;
    std::string item;
    for (const char* p = names;; ++p) {
      if (*p == ',' || *p == '\0') {
        if (!item.empty()) dev->profile_only.push_back(item);
        item.clear();
        if (*p == '\0') break;
      } else if (*p != ' ') {
        item.push_back(*p);
      }
    }
  });
}

gc_status gc_profile_query(gc_device* handle, const char* name, double* total_ms,
                           uint64_t* launches) {
  return guard([&] {
    if (handle == nullptr || name == nullptr) gc::fail_kernel("device or name is null");
    auto* dev = reinterpret_cast<gc::Device*>(handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    gc::profile_collect(*dev);
    double ms = 0.0;
    uint64_t count = 0;
    for (auto& t : dev->totals) {
      if (t.name == name) {
        ms = t.total_ms;
        count = t.launches;
      }
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = count;
  });
}

uint32_t gc_code_object_count(void) { return gc_embedded_code_object_count; }

gc_status gc_code_object_info(uint32_t i, const char** name, const char** source_sha256,
                              const char** object_sha256, uint64_t* size_bytes) {
  return guard([&] {
    if (i >= gc_embedded_code_object_count) gc::fail_kernel("code object index out of range");
    const auto& co = gc_embedded_code_objects[i];
    if (name) *name = co.name;
    if (source_sha256) *source_sha256 = co.source_sha256;
    if (object_sha256) *object_sha256 = co.object_sha256;
    if (size_bytes) *size_bytes = static_cast<uint64_t>(co.end - co.begin);
  });
}

}  // extern "C"
