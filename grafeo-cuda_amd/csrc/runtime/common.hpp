// Internal runtime of libgrafeo_hip: errors, devices, device memory, kernel
// launches and the graph object shared by the algorithm drivers.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "grafeo_hip.h"
#include "launch.hpp"

namespace gc {

// ---------------------------------------------------------------- errors --

// Thrown inside the library and turned into a gc_status at the C boundary.
struct Error {
  gc_status status;
  std::string message;  // Display text, e.g. "invalid CSR: ..."
};

[[noreturn]] void fail(gc_status status, const std::string& detail);

// The io::Error text Rust's BufRead::lines gives for a line that is not UTF-8
// (parse_tsv_to_edges fails with it).
inline constexpr const char* kInvalidUtf8 = "stream did not contain valid UTF-8";
[[noreturn]] inline void fail_kernel(const std::string& detail) { fail(GC_ERR_KERNEL, detail); }
[[noreturn]] inline void fail_csr(const std::string& detail) { fail(GC_ERR_INVALID_CSR, detail); }

void hip_check(hipError_t err, const char* what);
void set_last_error(const std::string& message);

// Runs `body` and converts exceptions into a status (and gc_last_error text).
gc_status guard(const std::function<void()>& body);

// ---------------------------------------------------------------- timing --

class Stopwatch {
 public:
  Stopwatch() : start_(std::chrono::steady_clock::now()) {}
  double seconds() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - start_).count();
  }

 private:
  std::chrono::steady_clock::time_point start_;
};

// ------------------------------------------------------- host parallelism --

// Number of host threads used by builders and graph preparation
// (GC_HOST_THREADS, default: hardware concurrency capped at 16).
unsigned host_threads();

// Runs fn(begin, end) over [0, n) split into contiguous chunks, one per
// thread. Small ranges run on the calling thread.
void parallel_for(uint64_t n, const std::function<void(uint64_t, uint64_t)>& fn,
                  uint64_t min_chunk = 1u << 16);

// ------------------------------------------------------------- profiling --

struct ProfileRecord {
  std::string name;
  hipEvent_t start;
  hipEvent_t stop;
};

struct ProfileTotal {
  std::string name;
  double total_ms = 0.0;
  uint64_t launches = 0;
};

// ---------------------------------------------------------------- device --

struct LoadedModule;

struct Device {
  std::atomic<int> refs{1};
  int ordinal = 0;
  hipStream_t stream = nullptr;
  std::string name;
  uint64_t total_memory = 0;
  int cu_count = 0;
  size_t max_lds_per_block = 64 * 1024;
  uint32_t pr_hot_count_limit = 0xFFFFFFFFu;  // lowered if a launch with a larger LDS set fails
  std::recursive_mutex mu;  // serializes calls on this device
  // Small pinned staging area for scalar read-backs (counters, states).
  void* pinned = nullptr;
  void* pinned_device = nullptr;  // the same buffer's address for kernels
  static constexpr size_t kPinnedBytes = 4096;
  // Bytes [kPinnedKernelOffset, kPinnedBytes) are written by kernels (BFS
  // level counters, polled by the host); the bytes below by read-back copies.
  static constexpr size_t kPinnedKernelOffset = 2048;
  uint32_t publish_seq = 0;  // last sequence number given to a publishing kernel
  // Device address of the static scratch array of the allocation-free entry
  // point on this device (null if it could not be resolved: noalloc_error).
  void* noalloc_scratch = nullptr;
  std::string noalloc_error;
  // Kernel timing.
  bool profiling = false;
  std::vector<std::string> profile_only;  // kernels timed when profiling (empty: all)
  std::vector<ProfileRecord> pending;
  std::vector<ProfileTotal> totals;
  std::vector<hipEvent_t> event_pool;

  void make_current() const;
  void sync() const;
};

// A kernel handle resolved from an embedded code object.
struct Kernel {
  hipFunction_t fn = nullptr;
  const char* name = "";
};

// Looks up kernel `name` in code object `family` on `dev` (loaded once per
// device through the module cache).
Kernel get_kernel(Device& dev, const char* family, const char* name);

// Event-timing hooks around one launch (no-ops unless profiling is on).
void profile_begin(Device& dev, const char* name, hipEvent_t* start);
void profile_end(Device& dev, const char* name, hipEvent_t start);

template <class T>
inline void* karg(T& value) {
  return static_cast<void*>(&value);
}

// Launches `k` with the given arguments (passed by value, each must match the
// kernel parameter type exactly).
template <class... Args>
void launch(Device& dev, const Kernel& k, uint32_t grid, uint32_t block, uint32_t shmem,
            Args... args) {
  if (grid == 0) return;
  void* params[] = {karg(args)..., nullptr};
  hipEvent_t start = nullptr;
  profile_begin(dev, k.name, &start);
  hipError_t err = hipModuleLaunchKernel(k.fn, grid, 1, 1, block, 1, 1, shmem, dev.stream, params,
                                         nullptr);
  if (err != hipSuccess) {
    fail(GC_ERR_KERNEL, std::string(k.name) + " launch failed: " + hipGetErrorString(err));
  }
  profile_end(dev, k.name, start);
}

// As launch(), without the event-timing hooks (which may create events): for
// the allocation-free entry point.
template <class... Args>
void launch_untimed(Device& dev, const Kernel& k, uint32_t grid, uint32_t block, uint32_t shmem,
                    Args... args) {
  if (grid == 0) return;
  void* params[] = {karg(args)..., nullptr};
  hipError_t err = hipModuleLaunchKernel(k.fn, grid, 1, 1, block, 1, 1, shmem, dev.stream, params,
                                         nullptr);
  if (err != hipSuccess) {
    fail(GC_ERR_KERNEL, std::string(k.name) + " launch failed: " + hipGetErrorString(err));
  }
}

// Resolves the device's allocation-free scratch (noalloc.hip) and loads the
// code objects the allocation-free entry point launches, so that its calls
// load and allocate nothing. Called once per device when it is opened; a
// failure is kept in noalloc_error and reported by that entry point only.
void noalloc_setup(Device& dev);

// As launch(), but returns the launch status instead of throwing (for launch
// shapes that a device may refuse, such as large dynamic LDS).
template <class... Args>
hipError_t try_launch(Device& dev, const Kernel& k, uint32_t grid, uint32_t block, uint32_t shmem,
                      Args... args) {
  if (grid == 0) return hipSuccess;
  void* params[] = {karg(args)..., nullptr};
  hipEvent_t start = nullptr;
  profile_begin(dev, k.name, &start);
  hipError_t err = hipModuleLaunchKernel(k.fn, grid, 1, 1, block, 1, 1, shmem, dev.stream, params,
                                         nullptr);
  if (err == hipSuccess) {
    profile_end(dev, k.name, start);
  } else {
    (void)hipGetLastError();
    if (start != nullptr) dev.event_pool.push_back(start);
  }
  return err;
}

// --------------------------------------------------------- device memory --

// Owning device allocation (hipMalloc on the device current at allocation).
class DeviceMemory {
 public:
  DeviceMemory() = default;
  DeviceMemory(const DeviceMemory&) = delete;
  DeviceMemory& operator=(const DeviceMemory&) = delete;
  DeviceMemory(DeviceMemory&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) {
    o.ptr_ = nullptr;
    o.bytes_ = 0;
  }
  DeviceMemory& operator=(DeviceMemory&& o) noexcept {
    if (this != &o) {
      release();
      ptr_ = o.ptr_;
      bytes_ = o.bytes_;
      o.ptr_ = nullptr;
      o.bytes_ = 0;
    }
    return *this;
  }
  ~DeviceMemory() { release(); }

  void allocate(size_t bytes);          // replaces any previous allocation
  void ensure(size_t bytes);            // grows (never shrinks) the allocation
  void release();
  void* get() const { return ptr_; }
  size_t bytes() const { return bytes_; }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr_);
  }

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

void copy_to_device(Device& dev, void* dst, const void* src, size_t bytes);
void copy_to_host(Device& dev, void* dst, const void* src, size_t bytes);  // synchronizes

// ----------------------------------------------------------------- graph --

// Degree-ordered relabelled copy of a graph: new id i is old vertex perm[i];
// vertices are sorted by degree ascending, ties by old id, so hubs sit at the
// top of the id range and the rank/distance vectors of the hubs are one
// contiguous, cache-resident block. Each row is sorted by new id; for a
// symmetric graph the entries with target > row form the degree-oriented DAG
// (lower degree -> higher degree, ties by lower old id), which is exactly the
// orientation of from_edge_pairs_oriented.
struct Relabelled {
  DeviceMemory perm;     // u32[n]   new -> old
  DeviceMemory inv;      // u32[n]   old -> new
  DeviceMemory offsets;  // u32[n+1]
  DeviceMemory targets;  // u32[m]
  DeviceMemory fwd;      // u32[n]   first index of row i with target > i
  std::vector<uint32_t> host_inv;
  std::vector<uint32_t> host_offsets;  // kept for host-side planning
  uint64_t oriented_entries = 0;       // sum of (offsets[i+1] - fwd[i])
  uint64_t intersect_entries = 0;      // sum over DAG entries i->j of (d+(i) + d+(j))
  uint64_t wedges = 0;                 // sum over rows of d+(i) (d+(i) - 1) / 2
  uint32_t dangling = 0;               // rows of degree 0
  uint32_t max_degree = 0;
  uint32_t max_out_degree = 0;         // max over rows of (offsets[i+1] - fwd[i])
  // In-lists (entries below the row) longer than kHubChunk are cut into
  // slices of kHubChunk entries: (vertex, begin, end) triples, heaviest first.
  static constexpr uint32_t kHubChunk = 2048;
  DeviceMemory hub_units;
  uint32_t num_hub_units = 0;
};

// The learned scheduler: the FP32 MLP of mlp.hip (12 -> 32 -> 32 -> 3) with
// its input normalisation, resident on one device.
constexpr uint32_t kSchedFeatures = 12;
constexpr uint32_t kSchedChoices = 3;
constexpr uint32_t kSchedParams = 32 * 12 + 32 + 32 * 32 + 32 + 3 * 32 + 3;  // 1571

struct SchedNorm {  // MlpNorm of mlp.hip
  float mean[kSchedFeatures];
  float scale[kSchedFeatures];
};

struct Sched {
  std::atomic<int> refs{1};
  Device* dev = nullptr;  // holds one reference
  DeviceMemory params;    // f32[kSchedParams]
  DeviceMemory m1, m2;    // Adam moments
  DeviceMemory grad;      // f32[kSchedParams] + loss
  SchedNorm norm{};
  uint32_t adam_step = 0;
  uint64_t version = 0;  // incremented whenever params or norm change
};

void sched_release(Sched* s);
// y[count][3] = the scheduler's network on x[count][12] (GPU, synchronizes).
void sched_predict(Sched& s, const float* x, uint32_t count, float* y);

// Instantiated HIP graphs owned by a graph (destroyed with it).
struct GraphExecs {
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  GraphExecs() = default;
  GraphExecs(const GraphExecs&) = delete;
  GraphExecs& operator=(const GraphExecs&) = delete;
  ~GraphExecs() {
    for (int i = 0; i < 2; ++i) {
      if (exec[i] != nullptr) (void)hipGraphExecDestroy(exec[i]);
      if (graph[i] != nullptr) (void)hipGraphDestroy(graph[i]);
    }
  }
};

// The static layout of PageRank's binned cold entries (pagerank.hip,
// pr_bin_scatter / pr_bin_accum / pr_spmv_bin) for one push set, hot set and
// set of sizes; built once per graph and kept with it.
struct PrBinPlan {
  // key
  uint32_t push_count = 0, hot_count = 0, stage = 0, bin_rows = 0, bin_values = 0;
  // rows [zero_rows, push_base) pull; sources [zero_rows, hot_base) are cold
  uint32_t zero_rows = 0, push_base = 0, hot_base = 0;
  uint32_t hot_entries = 0;  // length of the 16-bit stream
  uint32_t chunks = 0;       // 512-entry chunks of the 16-bit stream
  uint32_t subs = 0;         // scatter sub-ranges (workgroups)
  uint32_t bins = 0;         // accumulation bins (workgroups)
  uint64_t values = 0;       // length of the value buffer (with padding)
  uint64_t cold_entries = 0; // real cold entries
  uint64_t slot_rows = 0;    // slot rows (padded to whole blocks of eight)
  double build_seconds = 0.0;
  DeviceMemory hot16, hoff, starts, chunk_row;          // 16-bit stream, row starts, chunk rows
  DeviceMemory slots, block_group, sub, run_first;      // scatter layout
  DeviceMemory run_info, run_dst;                       // runs: staging range, destination
  DeviceMemory bin_row, bin_val, uloc, vals, cold_sum;  // bins, row indices, values, row sums
  DeviceMemory hub_pieces;                              // (hub row, first entry) of the pieces
  uint32_t pieces = 0;
  uint64_t runs = 0, piece_entries = 0;
  uint32_t grid = 0;  // workgroups of the last pr_spmv_bin launches
  // the sliced hot stream (pr_slice_layout.hpp), when the plan was built for pr_spmv_slice
  bool sliced = false;
  uint32_t cap_groups = 0, waves = 0;  // key
  uint32_t slices = 0, slice_items = 0, long_count = 0, partial_slots = 0;
  uint64_t slice_entries = 0;
  DeviceMemory slice_data, items, wave_first, long_slices, slice_partial, hub16;
};

struct Graph {
  Device* dev = nullptr;  // holds one reference
  Sched* bfs_sched = nullptr;  // learned BFS schedule (one reference), or null: the fixed rule
  Sched* tc_sched = nullptr;   // learned triangle-count schedule (one reference), or null
  std::vector<uint32_t> last_schedule;  // choices of the last BFS levels / TC buckets
  GraphExecs bfs_chain;                 // captured BFS level chains, one per start queue
  // Degree buckets of the relabelled rows and their work (triangle counting).
  struct TcBucket {
    uint32_t row_lo, row_hi;
    uint64_t rows, entries, probes, max_out;
    uint64_t hot_probes;  // probes among the bitmap kernels' hot ids (the top 32,768)
  };
  std::vector<TcBucket> tc_buckets;
  bool tc_buckets_ready = false;
  uint32_t n = 0;
  uint32_t m = 0;
  DeviceMemory offsets;  // u32[n+1]
  DeviceMemory targets;  // u32[m]
  // Host copy kept until the derived layouts are built.
  std::vector<uint32_t> host_offsets;
  std::vector<uint32_t> host_targets;

  bool prepared = false;
  bool symmetric = false;
  uint32_t self_loops = 0;
  double prepare_seconds = 0.0;
  Relabelled rl;

  // Scratch reused across calls (grown on demand, freed with the graph).
  DeviceMemory fwd_original;  // u32[n] first index with target > row (original layout)
  bool fwd_original_ready = false;
  DeviceMemory ktruss_rev;    // u32[m] reverse entry of each forward relabelled entry
  bool ktruss_rev_ready = false;
  // The middle side of the k-truss supports (kt_mid_support): one record per entry of the hot
  // rows, the bits of the DAG entries it walks, its work units.
  DeviceMemory kt_mid_rec;      // uint2 per entry of rows [hot_base, n): {first entry, entries}
  DeviceMemory kt_midbits;      // one bit per CSR entry
  DeviceMemory kt_mid_units;    // uint4 {v, first entry, end entry, 0}
  uint32_t kt_mid_num_units = 0;
  uint32_t kt_mid_rec_base = 0;
  bool kt_mid_ready = false;
  uint32_t kt_mid_key[2] = {0, 0};  // per-edge cost and unit the build is for
  DeviceMemory bfs_top;       // u32[4][n] the four highest neighbours of each relabelled row
  bool bfs_top_ready = false;
  // The DAG lists' entries among the top 32768 ids as 16-bit offsets, eight to
  // a 16-byte oct, each row from an oct boundary (tc_bitmap16).
  DeviceMemory tc_split;      // u32[n] first such entry of each relabelled row
  DeviceMemory tc_hw;         // uint2[n] {first oct, octs} of each row
  DeviceMemory tc_words;      // 16-byte octs of eight offsets of all rows, part by part
  bool tc_hot16_ready = false;
  uint32_t tc_hot16_parts = 0;  // the middle-vertex parts the words are grouped by
  uint32_t tc_hot16_top = 0;    // octs per mask of the top ids in the words (0: no masks)
  // The middle-vertex side of the 16-bit count (tc_mid16): the DAG rows' hot parts in
  // natural order, one record per entry of the hot rows, the bits of the DAG entries
  // it counts, and its work units.
  DeviceMemory tc_mid_octs;     // 16-byte octs, each row from an oct boundary
  DeviceMemory tc_mid_rec;      // uint2 per entry of rows [hot_base, n): {first oct, octs}
  DeviceMemory tc_midbits;      // one bit per CSR entry (set on the DAG entries counted at v)
  DeviceMemory tc_mid_units;    // uint4 {v, first entry, end entry, 0}
  DeviceMemory tc_rowwork;      // u8 per row: 1 when tc_bitmap16 still has entries of the row to walk
  DeviceMemory tc_mid32_rec;    // uint2 per entry of rows [mid_base, hot_base): {first entry, entries}
  DeviceMemory tc_mid32_units;  // uint4 {v, first entry, end entry, 0}
  uint32_t tc_mid32_num_units = 0;
  uint32_t tc_mid32_rec_base = 0;
  uint32_t tc_mid32_base = 0;
  uint32_t tc_mid_num_units = 0;
  uint32_t tc_mid_rec_base = 0;
  bool tc_mid_ready = false;
  uint32_t tc_mid_key[5] = {0, 0, 0, 0, 0};  // parts, top octs, per-edge cost, unit, 32-bit tier on
  uint32_t pr_merge_push = 0;   // push set size the PageRank row-start bitmap was built for (0: none)
  std::unique_ptr<PrBinPlan> pr_bin;  // PageRank's binned layout (null until the first binned call)
  // The learned triangle-count schedule (method per degree bucket): a function
  // of the graph and the network, kept with the network and version it came from.
  std::vector<uint32_t> tc_choice;
  const Sched* tc_choice_sched = nullptr;
  uint64_t tc_choice_version = 0;
  std::vector<std::pair<std::string, std::unique_ptr<DeviceMemory>>> buffers;

  // Named scratch buffer of at least `bytes` bytes (contents unspecified).
  template <class T>
  T* buffer(const char* name, size_t bytes) {
    for (auto& b : buffers) {
      if (b.first == name) {
        b.second->ensure(bytes == 0 ? 1 : bytes);
        return b.second->as<T>();
      }
    }
    buffers.emplace_back(name, std::unique_ptr<DeviceMemory>(new DeviceMemory()));
    buffers.back().second->ensure(bytes == 0 ? 1 : bytes);
    return buffers.back().second->as<T>();
  }
};

Graph* graph_from_handle(gc_graph* h);

// Locks the graph's device, makes it current and returns the graph.
struct GraphCall {
  Graph* g;
  std::unique_lock<std::recursive_mutex> lock;
  explicit GraphCall(gc_graph* h) : g(graph_from_handle(h)), lock(g->dev->mu) { g->dev->make_current(); }
  Device& dev() { return *g->dev; }
};

// Fills `words` 32-bit words at `p` with `value` on the device's stream.
void fill_words(Device& dev, void* p, uint64_t words, uint32_t value);

// In-place exclusive prefix sum of `count` u32 values on the device;
// `block_buf` holds scan_scratch_words(count) words, the last of which
// receives the total.
uint64_t scan_scratch_words(uint64_t count);
void exclusive_scan_u32(Device& dev, uint32_t* data, uint64_t count, uint32_t* block_buf);

// log2 of the LDS hash-table size of the middle-vertex triangle kernels: the
// smallest power of two holding twice the largest DAG out-degree, between 64
// and 4096 entries (rows above that search global memory instead).
uint32_t wedge_table_bits(const Relabelled& rl);

// The row schedule of the bitmap triangle kernels over all rows of a prepared
// graph (tc.cpp): the first hot id (the top 32768, a multiple of 32), the rows
// handed out one at a time, the middle-vertex parts and the shortest row split
// over them (GC_TC_HEAVY_DEGREE, GC_TC_PARTS, GC_TC_SPLIT).
void tc_bitmap_schedule(const Graph& g, uint32_t* hot_base, uint32_t* heavy, uint32_t* parts, uint32_t* split_min);

// Builds the derived layouts if not done yet (on the device; GC_PREPARE=host
// runs the host restatement instead).
void prepare(Graph& g);

// Stable LSD radix sort of the low `bits` bits of n 64-bit keys, 8 bits per
// pass, ping-ponging between keys and tmp; returns the buffer holding the
// sorted keys (csr_gpu.cpp, kernels of sort.hip).
unsigned long long* radix_sort(Device& dev, unsigned long long* keys, unsigned long long* tmp, uint64_t n,
                               uint32_t bits);
void ensure_fwd_original(Graph& g);

// Host-side CSR validation with the reference's messages (throws InvalidCsr).
void validate_csr(const uint32_t* offsets, size_t num_offsets, const uint32_t* targets,
                  size_t num_targets);

}  // namespace gc
