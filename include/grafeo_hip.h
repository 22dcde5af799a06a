/*
 * grafeo_hip.h - C ABI of the MI355X (gfx950) graph-kernel library.
 *
 * Drop-in replacement for the GPU path of grafeo-cuda: the Rust crates
 * `grafeo-cuda-core` (device, CSR model, graph upload) and
 * `grafeo-cuda-algos` (triangle counting, k-truss, BFS, PageRank), and the
 * PyO3 module `grafeo_cuda` that wraps them. Every entry point below names the
 * reference item it replaces (paths relative to the reference repository).
 *
 * Conventions
 *  - Plain pointers and sizes only. Host arrays are read or written by the
 *    call and never retained; results are owned by the caller, except arrays
 *    the library allocates (k-truss edges, builder output), which are released
 *    with gc_free / gc_csr_free.
 *  - Every call returns a gc_status. On failure gc_last_error() returns the
 *    message of the failing call on this thread (the Display text of the
 *    reference's CudaError).
 *  - All device work of one gc_device runs on that device's own stream and
 *    calls on one device are serialized; each call returns after its results
 *    are written (host outputs) or after its device work is complete (device
 *    outputs).
 *  - "secs" outputs carry the wall-clock time of the call, as the reference's
 *    `time_seconds` fields do.
 */
#ifndef GRAFEO_HIP_H
#define GRAFEO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, one per variant of `CudaError`
 * (crates/grafeo-cuda-core/src/error.rs:5-39). */
typedef enum gc_status {
  GC_OK = 0,
  GC_ERR_NO_DEVICE = 1,     /* CudaError::NoDevice */
  GC_ERR_DRIVER = 2,        /* CudaError::Driver (HIP runtime error) */
  GC_ERR_OUT_OF_MEMORY = 3, /* CudaError::OutOfMemory */
  GC_ERR_KERNEL = 4,        /* CudaError::Kernel (launch and argument errors) */
  GC_ERR_EXPORT = 5,        /* CudaError::Export */
  GC_ERR_INVALID_CSR = 6    /* CudaError::InvalidCsr */
} gc_status;

/* Message of the last failed call on the calling thread ("" if none).
 * Replaces `impl Display for CudaError` (crates/grafeo-cuda-core/src/error.rs:5-39). */
const char* gc_last_error(void);

/* Releases memory the library allocated (k-truss edges, parsed edge pairs,
 * generator output). */
void gc_free(void* ptr);

/* ------------------------------------------------------------------------ */
/* Device                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct gc_device gc_device;

/* Number of visible GPUs (0 when the runtime finds none). */
gc_status gc_device_count(uint32_t* count);

/* Opens GPU `ordinal` with its own stream; refcount 1.
 * Replaces GpuDevice::new (crates/grafeo-cuda-core/src/device.rs:30-51). */
gc_status gc_device_new(uint32_t ordinal, gc_device** out);

/* Adds a reference (GpuDevice is Clone, crates/grafeo-cuda-core/src/device.rs:20-26). */
gc_device* gc_device_retain(gc_device* dev);

/* Drops a reference; the last one frees the stream and cached modules. */
void gc_device_release(gc_device* dev);

/* DeviceProperties {name, ordinal, total_memory}
 * (crates/grafeo-cuda-core/src/device.rs:8-16,55-58). `name` receives a
 * NUL-terminated string truncated to `name_cap` bytes. */
gc_status gc_device_properties(const gc_device* dev, char* name, size_t name_cap,
                               uint64_t* total_memory, uint32_t* ordinal);

/* ------------------------------------------------------------------------ */
/* CSR model                                                                 */
/* ------------------------------------------------------------------------ */

/* Library-owned CSR arrays: CsrArrays (crates/grafeo-cuda-core/src/csr_export.rs:15-30). */
typedef struct gc_csr {
  uint32_t* offsets;  /* num_nodes + 1 entries */
  uint32_t* targets;  /* num_edges entries */
  uint32_t num_nodes;
  uint32_t num_edges;
} gc_csr;

void gc_csr_free(gc_csr* csr);

/* validate_csr (crates/grafeo-cuda-core/src/csr_export.rs:71-120): offsets
 * non-empty, offsets[0] == 0, non-decreasing, offsets[n] == num_targets, both
 * lengths fit u32, targets < n, each list strictly increasing. Returns
 * GC_ERR_INVALID_CSR with the reference's message otherwise. */
gc_status gc_validate_csr(const uint32_t* offsets, size_t num_offsets,
                          const uint32_t* targets, size_t num_targets);

/* CsrArrays::validate (crates/grafeo-cuda-core/src/csr_export.rs:37-54). */
gc_status gc_csr_validate(const gc_csr* csr);

/* from_edge_pairs (crates/grafeo-cuda-core/src/csr_export.rs:202-239).
 * `pairs` holds num_pairs (src, dst) pairs interleaved. If `symmetric`, each
 * pair also adds (dst, src) unless it is a self-loop; the result is sorted
 * and deduplicated. A node id >= num_nodes is GC_ERR_INVALID_CSR. */
gc_status gc_csr_from_edge_pairs(const uint32_t* pairs, size_t num_pairs, uint32_t num_nodes,
                                 int symmetric, gc_csr* out);

/* from_edge_pairs_oriented (crates/grafeo-cuda-core/src/csr_export.rs:299-350):
 * drops loops, deduplicates undirected edges, orients each from the
 * lower-degree endpoint to the higher (ties: lower id first). */
gc_status gc_csr_from_edge_pairs_oriented(const uint32_t* pairs, size_t num_pairs,
                                          uint32_t num_nodes, gc_csr* out);

/* from_edge_pairs (crates/grafeo-cuda-core/src/csr_export.rs:202-239) built on
 * the GPU (stable radix sort of packed keys, dedup, offsets by binary
 * search): the same arrays as gc_csr_from_edge_pairs, returned to the host
 * (gc_csr_free). Out-of-range ids are GC_ERR_INVALID_CSR with the same
 * message. */
gc_status gc_csr_from_edge_pairs_device(gc_device* dev, const uint32_t* pairs, size_t num_pairs,
                                        uint32_t num_nodes, int32_t symmetric, gc_csr* out);

/* from_edge_pairs_oriented (crates/grafeo-cuda-core/src/csr_export.rs:299-350)
 * built on the GPU: canonical (min, max) keys radix-sorted, loops and copies
 * dropped, undirected degrees, each edge oriented from the lower degree to the
 * higher (ties: lower id first), sorted again. The same arrays as
 * gc_csr_from_edge_pairs_oriented, returned to the host (gc_csr_free);
 * out-of-range ids are GC_ERR_INVALID_CSR with the same message. */
gc_status gc_csr_from_edge_pairs_oriented_device(gc_device* dev, const uint32_t* pairs, size_t num_pairs,
                                                 uint32_t num_nodes, gc_csr* out);

/* parse_tsv_to_edges (crates/grafeo-cuda-core/src/csr_export.rs:246-287)
 * parsed on the GPU: the same pairs and num_nodes as gc_parse_tsv_to_edges
 * (ids in order of first appearance), `*pairs` library-allocated (gc_free).
 * The file is read on the host and parsed on the device (line breaks, one lane
 * per line, ids numbered by stable radix sorts). Files of 4 GiB or more are a
 * GC_ERR_KERNEL error; a missing file is GC_ERR_EXPORT. */
gc_status gc_parse_tsv_to_edges_device(gc_device* dev, const char* path, uint32_t** pairs, size_t* num_pairs,
                                       uint32_t* num_nodes);

/* parse_tsv_to_edges (crates/grafeo-cuda-core/src/csr_export.rs:246-287).
 * `*pairs` is library-allocated (gc_free); ids are remapped in order of first
 * appearance. A missing file is GC_ERR_EXPORT. */
gc_status gc_parse_tsv_to_edges(const char* path, uint32_t** pairs, size_t* num_pairs,
                                uint32_t* num_nodes);

/* ------------------------------------------------------------------------ */
/* Graph in device memory                                                    */
/* ------------------------------------------------------------------------ */

typedef struct gc_graph gc_graph;

/* GpuGraph::from_csr_slices (crates/grafeo-cuda-core/src/gpu_graph.rs:66-87):
 * validates, then copies both arrays to the device. The graph keeps a
 * reference to `dev`. */
gc_status gc_graph_from_csr(gc_device* dev, const uint32_t* offsets, size_t num_offsets,
                            const uint32_t* targets, size_t num_targets, gc_graph** out);

/* GpuGraph::from_csr_arrays (crates/grafeo-cuda-core/src/gpu_graph.rs:46-60). */
gc_status gc_graph_from_csr_arrays(gc_device* dev, const gc_csr* csr, gc_graph** out);

void gc_graph_free(gc_graph* g);

/* GpuGraph.num_nodes / num_edges (crates/grafeo-cuda-core/src/gpu_graph.rs:20-23). */
uint32_t gc_graph_num_nodes(const gc_graph* g);
uint32_t gc_graph_num_edges(const gc_graph* g);

/* Device addresses of the uploaded CSR (GpuGraph.offsets / .targets,
 * crates/grafeo-cuda-core/src/gpu_graph.rs:16-19). */
const uint32_t* gc_graph_device_offsets(const gc_graph* g);
const uint32_t* gc_graph_device_targets(const gc_graph* g);

/* Builds the graph's derived device layouts now instead of on first use: the
 * degree-ordered relabelled CSR used by the symmetric kernels, and the
 * symmetry check. No reference counterpart (the reference keeps only the
 * uploaded arrays); `secs` is the preparation time. */
gc_status gc_graph_prepare(gc_graph* g, double* secs);

/* 1 if every entry u->v of the CSR has the entry v->u, else 0 (prepares). */
gc_status gc_graph_is_symmetric(gc_graph* g, int32_t* symmetric);

/* Sizes of the degree-oriented DAG the symmetric kernels run on (prepares):
 * its entry count m_o (entries u->v whose (degree, id) rank is above u's),
 * the sum over those entries of (d+(u) + d+(v)) (the list entries a merge
 * intersection per DAG edge streams), and the number of DAG wedges,
 * sum over u of d+(u)(d+(u)-1)/2 (the list entries the middle-vertex triangle
 * kernel streams). Used to state the triangle kernels' algorithmic bytes. No
 * reference counterpart. */
gc_status gc_graph_dag_stats(gc_graph* g, uint64_t* oriented_entries, uint64_t* intersect_entries,
                             uint64_t* wedges);

/* ------------------------------------------------------------------------ */
/* Algorithms                                                                */
/* ------------------------------------------------------------------------ */

/* triangle_counting::triangle_count (crates/grafeo-cuda-algos/src/triangle_counting.rs:40-152):
 * number of (u, v, w), u < v < w, with v and w in N(u) and w in N(v)
 * (the triangle count of a symmetric CSR; self-loops never count). */
gc_status gc_triangle_count(gc_graph* g, uint64_t* total_triangles, double* secs);

/* Allocation-free triangle count. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). It returns what
 * gc_triangle_count returns, computed by the fixed rule's bitmap kernel. It
 * allocates no device memory, loads no code object, builds nothing and
 * creates no events. The kernel's work counters and u64 result live in a static
 * per-GPU array of GC_NOALLOC_SCRATCH_BYTES bytes. The array is resolved when
 * the device is opened (gc_device_new) and held until the process ends. Every
 * call on one GPU shares that array, which keeps the last call's contents:
 * each call clears it first, and the library serializes such calls per GPU
 * (also across gc_device handles of one ordinal). A caller that bypasses the
 * library to use the array concurrently would race.
 * Preconditions: a symmetric graph already prepared by gc_graph_prepare
 * (otherwise GC_ERR_KERNEL, before any device work). A graph with no edges
 * gives 0. If the library's 16-bit triangle lists exist (built by an earlier
 * gc_triangle_count), the faster tc_bitmap16 kernel runs; otherwise tc_bitmap
 * runs. GC_ERR_DRIVER if the scratch could not be resolved on this device.
 * The scratch this count needs does not grow with the graph; a need above the
 * array's size is GC_ERR_KERNEL. */
#define GC_NOALLOC_SCRATCH_BYTES 4096
gc_status gc_triangle_count_noalloc(gc_graph* g, uint64_t* total_triangles);

/* triangle_counting::triangle_count_oriented
 * (crates/grafeo-cuda-algos/src/triangle_counting.rs:162-210): the sum over
 * CSR entries u->v of |N(u) intersect N(v)| (the triangle count of a
 * degree-oriented CSR). */
gc_status gc_triangle_count_oriented(gc_graph* g, uint64_t* total_triangles, double* secs);

/* bfs::bfs (crates/grafeo-cuda-algos/src/bfs.rs:42-131): level-synchronous BFS
 * along the CSR lists. `distances` (host, num_nodes entries) receives -1 for
 * unreachable nodes. A source >= num_nodes is GC_ERR_KERNEL. */
gc_status gc_bfs(gc_graph* g, uint32_t source, int32_t* distances, uint32_t* reachable,
                 int32_t* max_depth, double* secs);

/* As gc_bfs, with `d_distances` a device array of num_nodes int32. */
gc_status gc_bfs_device(gc_graph* g, uint32_t source, int32_t* d_distances, uint32_t* reachable,
                        int32_t* max_depth, double* secs);

/* pagerank::pagerank (crates/grafeo-cuda-algos/src/pagerank.rs:46-199):
 * r0 = 1/n; r'[v] = (1-d)/n + d*(sum_{u in N(v), deg u > 0} r[u]/deg u + D/n)
 * with D the rank of nodes of degree 0; stops when the L1 change drops below
 * `tolerance` or after `max_iterations`. FP32 ranks, FP64 reductions.
 * `ranks` (host, num_nodes floats) receives the latest iterate. */
gc_status gc_pagerank(gc_graph* g, float damping, uint32_t max_iterations, float tolerance,
                      float* ranks, uint32_t* iterations, int32_t* converged, double* secs);

/* As gc_pagerank, with `d_ranks` a device array of num_nodes floats. */
gc_status gc_pagerank_device(gc_graph* g, float damping, uint32_t max_iterations, float tolerance,
                             float* d_ranks, uint32_t* iterations, int32_t* converged,
                             double* secs);

/* Sizes of the binned layout the last gc_pagerank / gc_pagerank_device call on a
 * symmetric graph used (the default SpMV; zeros if none was built). Diagnostic,
 * no reference counterpart: the benchmark states its kernels' bytes from them.
 * Fills min(count, 25) values: push set, hot set, staged entries per scatter
 * workgroup, rows and values per bin (the limits), cold entries, value buffer
 * length, slot rows, 16-bit stream entries, stream chunks, scatter sub-ranges,
 * bins, runs, hub pieces, hub-piece entries, pull rows, cold sources, the
 * layout's build time in nanoseconds, the pull kernel's workgroups; then 1 if
 * the hot stream is sliced (GC_PR_KERNEL=slice), slices, sliced entries (with
 * padding), work items, wide slices and their pieces. */
gc_status gc_pagerank_layout(gc_graph* g, uint64_t* stats, uint32_t count);

/* Multi-source BFS (an addition with no reference function; for each source
 * it returns what gc_bfs returns, bfs::bfs, crates/grafeo-cuda-algos/src/bfs.rs:42-131).
 * `distances` (host) receives num_sources rows of num_nodes int32 distances,
 * row s for sources[s] (-1 unreachable, 0 at the source); `reachable` and
 * `max_depth` (host arrays of num_sources, each nullable) get every source's
 * reachable count and last non-empty level. Up to 128 sources share one
 * traversal (a 128-bit source mask per vertex); more run 128 at a time.
 * Sources may repeat. A source >= num_nodes is GC_ERR_KERNEL with gc_bfs's
 * message, checked before any work. */
gc_status gc_msbfs(gc_graph* g, const uint32_t* sources, uint32_t num_sources, int32_t* distances,
                   uint32_t* reachable, int32_t* max_depth, double* secs);

/* As gc_msbfs, with `d_distances` a device array of num_sources * num_nodes
 * int32 (row-major, one row per source). */
gc_status gc_msbfs_device(gc_graph* g, const uint32_t* sources, uint32_t num_sources,
                          int32_t* d_distances, uint32_t* reachable, int32_t* max_depth,
                          double* secs);

/* k_truss::k_truss (crates/grafeo-cuda-algos/src/k_truss.rs:42-161): peels
 * edges with fewer than k-2 supporting triangles until none is removed.
 * `*edges` (library-allocated, gc_free) receives `*num_edges` pairs (u, v),
 * u < v, in CSR order; `iterations` counts the final round that removes
 * nothing. k < 3 is GC_ERR_KERNEL. */
gc_status gc_k_truss(gc_graph* g, uint32_t k, uint32_t** edges, size_t* num_edges,
                     uint32_t* iterations, double* secs);

/* Trussness of every edge (an addition with no reference function; it is
 * consistent with k_truss::k_truss, crates/grafeo-cuda-algos/src/k_truss.rs:42-161):
 * truss[i] is the largest k whose k-truss (as gc_k_truss computes it) holds
 * edge i, or 2 for an edge in no triangle, so gc_k_truss(g, k) returns exactly
 * the edges with truss >= k. `*edges` (pairs (u, v), u < v, in CSR order, as
 * gc_k_truss emits them) and `*truss` are library-allocated (gc_free each);
 * `*max_truss` is the largest value (0 for a graph without edges). Needs a
 * symmetric CSR (GC_ERR_KERNEL otherwise). */
gc_status gc_truss_numbers(gc_graph* g, uint32_t** edges, uint32_t** truss, size_t* num_edges,
                           uint32_t* max_truss, double* secs);

/* ------------------------------------------------------------------------ */
/* Batched triangle counting (an addition with no reference function: per   */
/* graph it returns what gc_triangle_count returns,                           */
/* triangle_counting::triangle_count, triangle_counting.rs:40-152)            */
/* ------------------------------------------------------------------------ */

typedef struct gc_batch gc_batch;

/* Builds a batch of `num_graphs` symmetric CSRs (graph i: offsets[i] of
 * num_nodes[i] + 1 entries, targets[i] of num_targets[i] entries). Every CSR
 * is validated first (GC_ERR_INVALID_CSR with the graph index appended); a
 * non-symmetric CSR is GC_ERR_KERNEL. The batch holds its own device copy of
 * the degree-oriented DAGs and a reference to the device. */
gc_status gc_batch_from_csrs(gc_device* dev, uint32_t num_graphs, const uint32_t* const* offsets,
                             const uint32_t* num_nodes, const uint32_t* const* targets,
                             const uint64_t* num_targets, gc_batch** out);
void gc_batch_free(gc_batch* b);
uint32_t gc_batch_num_graphs(const gc_batch* b);

/* Triangles of every graph (`counts`, host, num_graphs values) from one replay
 * of a captured HIP graph holding one persistent counting launch. */
gc_status gc_batch_triangle_count(gc_batch* b, uint64_t* counts, double* secs);

/* Launch shape: config[0] = out-degree from which a row uses the LDS bitmap
 * (0: never), config[1] = DAG edges per merge task, config[2] = rows per
 * bitmap task, config[3] = workgroups per CU. Results never depend on it.
 * gc_batch_tune times candidate shapes on the batch (coordinate descent from
 * the default), keeps the fastest and returns it with its time in ms. */
gc_status gc_batch_set_config(gc_batch* b, const uint32_t* config);
gc_status gc_batch_get_config(gc_batch* b, uint32_t* config);
gc_status gc_batch_tune(gc_batch* b, uint32_t* config, double* best_ms);

/* ------------------------------------------------------------------------ */
/* Learned scheduler (an addition with no reference function; it changes no */
/* result, only which kernels compute it)                                   */
/* ------------------------------------------------------------------------ */

/* A 12 -> 32 -> 32 -> 3 FP32 MLP (ReLU hidden layers) resident on a device.
 * Its 3 outputs predict the log-time (natural log of microseconds) of three
 * choices for one decision; the smallest allowed prediction wins. Parameter
 * layout (GC_SCHED_PARAMS floats): W1[32][12], b1[32], W2[32][32], b2[32],
 * W3[3][32], b3[3]; inputs are normalised as (x - mean[i]) * scale[i]. */
#define GC_SCHED_FEATURES 12
#define GC_SCHED_CHOICES 3
#define GC_SCHED_PARAMS 1571

typedef struct gc_sched gc_sched;

/* He-uniform weights from SplitMix64(seed), zero biases, mean 0, scale 1. */
gc_status gc_sched_new(gc_device* dev, uint64_t seed, gc_sched** out);
void gc_sched_free(gc_sched* s);
/* Copies out / in the parameters and the input normalisation (each nullable;
 * setting parameters also resets the optimizer state). */
gc_status gc_sched_get(gc_sched* s, float* params, float* mean, float* scale);
gc_status gc_sched_set(gc_sched* s, const float* params, const float* mean, const float* scale);
/* y[count][3] = network(x[count][12]) on the GPU. */
gc_status gc_sched_predict(gc_sched* s, const float* x, uint32_t count, float* y);
/* Loss L = sum mask (y - t)^2 / sum mask over a batch (x[count][12],
 * t[count][3], mask[count][3] in {0, 1}) and its gradient dL/dparams
 * (GC_SCHED_PARAMS floats), by the GPU backward kernels. */
gc_status gc_sched_loss_grad(gc_sched* s, const float* x, const float* t, const float* mask,
                             uint32_t count, float* loss, float* grad);
/* `steps` full-batch Adam steps (beta 0.9 / 0.999, eps 1e-8) on the GPU;
 * *final_loss is the loss after the last step. */
gc_status gc_sched_train(gc_sched* s, const float* x, const float* t, const float* mask,
                         uint32_t count, uint32_t steps, float lr, float* final_loss);

/* Attaches schedulers to a graph (each holds a reference; null detaches):
 * with `bfs`, BFS picks each level among top-down, bottom-up and the
 * device-chained levels (captured HIP graphs of many levels with one
 * read-back per graph); with `tc`, the
 * symmetric triangle count picks, per degree bucket, among the bitmap, hash
 * and binary-search intersections. Without them the fixed rules run (BFS:
 * Beamer, Asanovic and Patterson's alpha = 14, beta = 24 switch; triangle
 * count: the bitmap intersection over all rows). Results are identical. */
gc_status gc_graph_set_scheduler(gc_graph* g, gc_sched* bfs, gc_sched* tc);
/* The choices taken by the last BFS (one per level run) or triangle count
 * (one per degree bucket) on this graph. */
gc_status gc_graph_last_schedule(gc_graph* g, uint32_t* choices, uint32_t capacity, uint32_t* count);

/* Training data. BFS from `source`: for each level (at most max_levels), the
 * 12 features of the decision (features[level][12]) and the time of each
 * choice run from the same state (times_us[level][3], -1 where the choice is
 * not allowed); the traversal continues along the fastest choice. */
gc_status gc_bfs_level_timings(gc_graph* g, uint32_t source, uint32_t max_levels, float* features,
                               float* times_us, uint32_t* levels);
/* Symmetric triangle count: per degree bucket (rows of degree [2^b, 2^(b+1))),
 * its 12 features and the time of each intersection method on it alone;
 * every method's count is checked against the others. */
gc_status gc_tc_bucket_timings(gc_graph* g, uint32_t max_buckets, float* features, float* times_us,
                               uint32_t* buckets);

/* GpuGraph from edge pairs without a host CSR: the CSR is built on the GPU
 * (as gc_csr_from_edge_pairs_device) and becomes the graph's arrays; it is
 * valid by construction, so no validate_csr pass runs (an addition: the
 * reference builds the CSR on the host, harness.rs:84-115). */
gc_status gc_graph_from_edge_pairs(gc_device* dev, const uint32_t* pairs, size_t num_pairs,
                                   uint32_t num_nodes, int32_t symmetric, gc_graph** out);

/* GpuGraph from a TSV edge list with no host CSR (an addition: the reference
 * parses and builds on the host, harness.rs:84-115): parsed as
 * gc_parse_tsv_to_edges_device, then built as gc_graph_from_edge_pairs
 * (`symmetric` as there); the pairs never leave the device. `num_pairs`
 * (optional) receives the edge lines parsed. */
gc_status gc_graph_from_tsv(gc_device* dev, const char* path, int32_t symmetric, gc_graph** out,
                            size_t* num_pairs);

/* The degree-ordered layout the kernels run on (built on first use; no
 * reference counterpart), as device pointers owned by the graph: offsets
 * (num_nodes + 1), targets (num_edges, each row sorted), perm (new id -> the
 * caller's id) and fwd (per row, the first entry above the row). */
gc_status gc_graph_relabelled(gc_graph* g, const uint32_t** offsets, const uint32_t** targets,
                              const uint32_t** perm, const uint32_t** fwd);

/* Device memory for the *_device entry points' outputs (hipMalloc on the
 * device of `dev`). No reference counterpart. */
gc_status gc_device_malloc(gc_device* dev, size_t bytes, void** ptr);
gc_status gc_device_free(gc_device* dev, void* ptr);

/* Copies `bytes` from device address `src` to host memory `dst`. */
gc_status gc_device_copy_to_host(gc_device* dev, void* dst, const void* src, size_t bytes);

/* ------------------------------------------------------------------------ */
/* Kernel timing (hipEvent pairs around each launch)                         */
/* ------------------------------------------------------------------------ */

/* Turns per-kernel event timing on (1) or off (0) for launches on `dev`, and
 * clears the accumulated figures. No reference counterpart: the reference
 * times whole calls only (crates/grafeo-cuda-algos/src/harness.rs:13-36). */
gc_status gc_profile_enable(gc_device* dev, int32_t on);

/* Restricts event timing to the comma-separated kernel `names` (NULL or ""
 * times every kernel again). Each timed launch adds two event records to the
 * stream, so a benchmark times only the kernels it reports. No reference
 * counterpart (as gc_profile_enable). */
gc_status gc_profile_only(gc_device* dev, const char* names);

/* Total milliseconds and launch count recorded for kernel `name` since the
 * last gc_profile_enable. Waits for outstanding launches. */
gc_status gc_profile_query(gc_device* dev, const char* name, double* total_ms, uint64_t* launches);

/* ------------------------------------------------------------------------ */
/* Synthetic graph generators (benchmark inputs)                             */
/* ------------------------------------------------------------------------ */

/* Erdos-Renyi-style multigraph draws: `num_draws` pairs (u, v) uniform over
 * [0, n)^2 from SplitMix64(seed); self-loops are dropped, so *num_pairs <=
 * num_draws. No reference counterpart (the reference loads TSV files). */
gc_status gc_gen_erdos_renyi(uint32_t n, uint64_t num_draws, uint64_t seed, uint32_t** pairs,
                             size_t* num_pairs);

/* R-MAT draws: edge_factor * 2^scale pairs, quadrant probabilities (a, b, c,
 * 1-a-b-c) per bit level; if `permute`, vertex ids pass through a seeded
 * random permutation; self-loops are dropped. */
gc_status gc_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                      uint64_t seed, int32_t permute, uint32_t** pairs, size_t* num_pairs);

/* rows x cols 4-neighbour grid, node id r*cols + c, each undirected edge once. */
gc_status gc_gen_grid(uint32_t rows, uint32_t cols, uint32_t** pairs, size_t* num_pairs);

/* ------------------------------------------------------------------------ */
/* Build information                                                         */
/* ------------------------------------------------------------------------ */

/* Number of embedded gfx950 code objects, and for entry `i` its kernel-family
 * name, the SHA-256 of the HIP source it was built from and of the code
 * object itself. Replaces the embedded-PTX checks
 * (crates/grafeo-cuda-algos/src/lib.rs:26-63). */
uint32_t gc_code_object_count(void);
gc_status gc_code_object_info(uint32_t i, const char** name, const char** source_sha256,
                              const char** object_sha256, uint64_t* size_bytes);

#ifdef __cplusplus
}
#endif

#endif /* GRAFEO_HIP_H */
