/*
 * bulkx.h -- C-ABI of the MI355X (gfx950) algorithm layer.
 *
 * Every entry point is stream-ordered: it enqueues work on `stream` (a
 * hipStream_t passed as void*, NULL = the null stream) and returns without
 * waiting, exactly like a bulk::async launch of the reference.  Pointers named
 * in/out/keys/values are DEVICE pointers owned by the caller; `init` arguments
 * are HOST pointers to one element of the stated dtype.
 *
 * Scratch memory: every algorithm takes (workspace, workspace_bytes).  Pass a
 * device buffer of at least bulkx_workspace_size(...) bytes to make the call
 * allocation-free, or workspace = NULL to let the call allocate and free its
 * temporaries stream-ordered (hipMallocAsync / hipFreeAsync), which is what
 * the reference's thrust::detail::temporary_array does per call.  A workspace
 * must not be used by two calls that may execute at the same time.
 *
 * Return value: 0 (BULKX_SUCCESS) or a BULKX_ERR_* code; bulkx_status_string()
 * names it.  Argument checks happen before anything is enqueued.
 */
#ifndef BULKX_H
#define BULKX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- element types (the reference instantiates int, unsigned, long, float, double) ---- */
enum bulkx_dtype {
  BULKX_INT32 = 0,
  BULKX_UINT32 = 1,
  BULKX_INT64 = 2,
  BULKX_UINT64 = 3,
  BULKX_FLOAT32 = 4,
  BULKX_FLOAT64 = 5
};

/* ---- binary operations: thrust::plus / thrust::minimum / thrust::maximum ---- */
enum bulkx_op { BULKX_PLUS = 0, BULKX_MIN = 1, BULKX_MAX = 2 };

/* ---- comparators: thrust::less / thrust::greater ---- */
enum bulkx_compare { BULKX_LESS = 0, BULKX_GREATER = 1 };

/* ---- key predicates for reduce_by_key: thrust::equal_to ---- */
enum bulkx_pred { BULKX_EQUAL_TO = 0 };

/* ---- algorithms (for bulkx_workspace_size) ---- */
enum bulkx_algo {
  BULKX_ALGO_INCLUSIVE_SCAN = 0,
  BULKX_ALGO_EXCLUSIVE_SCAN = 1,
  BULKX_ALGO_REDUCE = 2,
  BULKX_ALGO_MERGE = 3,
  BULKX_ALGO_MERGE_BY_KEY = 4,
  BULKX_ALGO_SORT_BY_KEY = 5,
  BULKX_ALGO_REDUCE_BY_KEY = 6,
  BULKX_ALGO_REDUCE_INTERVALS = 7,
  /* tuning-table keys of the batch executor's families (no workspace):
   * scan / reduce jobs (dtype INT32 or FLOAT32), reduce_by_key jobs (dtype =
   * the value dtype: INT32, FLOAT32 or INT64) and merge jobs (dtype INT32) */
  BULKX_ALGO_BATCH_SCAN = 8,
  BULKX_ALGO_BATCH_REDUCE_BY_KEY = 9,
  BULKX_ALGO_BATCH_MERGE = 10
};

/* ---- status codes ---- */
enum bulkx_status {
  BULKX_SUCCESS = 0,
  BULKX_ERR_INVALID_ARGUMENT = 1,
  BULKX_ERR_UNSUPPORTED_DTYPE = 2,
  BULKX_ERR_UNSUPPORTED_OP = 3,
  BULKX_ERR_WORKSPACE_TOO_SMALL = 4,
  BULKX_ERR_NO_DEVICE = 5,
  BULKX_ERR_HIP = 6,
  BULKX_ERR_OUT_OF_MEMORY = 7,
  BULKX_ERR_BUDGET = 8,
  BULKX_ERR_NOT_READY = 9,
  BULKX_ERR_INTERNAL = 10,
  BULKX_ERR_LOOKBACK = 11 /* a look-back wait gave up: the results of that launch are wrong */
};

/* Name of a status code (static string; never NULL). */
const char* bulkx_status_string(int status);

/* Last HIP error text seen by this thread (after BULKX_ERR_HIP). */
const char* bulkx_last_error(void);

/* Library version, "major.minor.patch". */
const char* bulkx_version(void);

/* Size in bytes of one element of `dtype`, 0 if unknown. */
size_t bulkx_dtype_size(int dtype);

/*
 * Deferred device-side failure of the plain entry points.  A tile of a
 * single-pass scan or reduce_by_key waits for its predecessors' published
 * prefixes (decoupled look-back); a wait bounded by the spin limit that runs
 * out sets the current device's sticky look-back error word, and that
 * launch's results are wrong.  bulkx_stream_status waits for `stream`, then
 * reads and clears the word: BULKX_ERR_LOOKBACK if any launch on this device
 * gave up since the last query, else BULKX_SUCCESS.  The counterpart of the
 * reference's sticky launch errors, reported by cudaGetLastError after a
 * synchronize (bulk/detail/throw_on_error.hpp:33-49, terminate.hpp:34).
 */
int bulkx_stream_status(void* stream);

/*
 * Polls a look-back wait makes before it gives up (default 2^22, about 0.2 s).
 * 0 makes every look-back give up at once without reading its predecessors,
 * so a test can drive the error path from the host while every kernel still
 * exits normally.  Applies to launches enqueued afterwards on the current
 * device and to batches whose job list is set afterwards.  No counterpart
 * in the reference (its waits are unbounded).
 */
int bulkx_set_lookback_spin_limit(uint32_t polls);

/*
 * Device properties the launcher sizes against (replaces
 * concurrent_group<>::hardware_concurrency(), bulk/execution_policy.hpp:368-376,
 * and the runtime_introspection cache, bulk/detail/cuda_launcher/runtime_introspection.inl:34-157).
 * Returns BULKX_ERR_NO_DEVICE when no GPU is visible.
 */
int bulkx_device_info(int* num_cus, int* lds_bytes_per_cu, int* wavefront_size,
                      int* max_threads_per_group, long long* hbm_bytes);

/*
 * choose_sizes replacement (bulk/choose_sizes.hpp:29-75, bulk/detail/choose_sizes.inl:32-117):
 * the (num_groups, group_size, grain_size) the library launches for `algo`
 * over n elements of `dtype`, taken from the tuning table when one is loaded.
 */
int bulkx_choose_sizes(int algo, int dtype, int64_t n,
                       int64_t* num_groups, int* group_size, int* grain_size);

/*
 * Load a tuning table written by the autotuner (JSON text).  path = NULL
 * reloads from $BULKX_TUNING or the packaged default.  Returns the number of
 * entries loaded through *entries.
 */
int bulkx_load_tuning(const char* path, int* entries);

/*
 * Autotuner interface (no counterpart in the reference, whose group and grain
 * sizes are constants in each driver): every algorithm is compiled with
 * bulkx_tuning_variants() tile shapes; the table maps (algo, dtype, size
 * class) to one of them.  Size classes: 0: n < 2^16, 1: < 2^20, 2: < 2^24,
 * 3: < 2^26, 4: >= 2^26.  The dtype key is the key/element dtype, and the value dtype
 * for reduce_by_key.  variant = -1 removes the entry (built-in default).
 */
int bulkx_tuning_variants(int algo, int* count);
int bulkx_set_tuning(int algo, int dtype, int size_class, int variant);
/* The variant the table (or the built-in default, 0) selects for (algo,
 * dtype, n); for stable_sort_by_key, variant 4 is the radix form. */
int bulkx_tuned_variant(int algo, int dtype, int64_t n, int* variant);

/* Bytes of scratch `algo` needs for n1 (+ n2) elements; *bytes may be 0. */
int bulkx_workspace_size(int algo, int dtype, int value_dtype, int64_t n1, int64_t n2,
                         size_t* bytes);

/*
 * inclusive_scan -- replaces ::inclusive_scan(first, last, result, init, op),
 * scan.cu:87-148 (whose per-group body is bulk::inclusive_scan, scan.hpp:455-486).
 * result[i] = init op x[0] op ... op x[i].  init = NULL means no init
 * (bulk::inclusive_scan without init, scan.hpp:489-526).  in == out is allowed.
 */
int bulkx_inclusive_scan(int dtype, int op, const void* in, void* out, int64_t n,
                         const void* init, void* workspace, size_t workspace_bytes,
                         void* stream);

/*
 * inclusive_scan_noalloc -- bulkx_inclusive_scan with no workspace argument
 * and no allocation of any kind: the look-back scratch is a static 4 MiB
 * (+256 B) device array that the library holds on each device until the
 * process ends.  No counterpart in the reference; for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Limits: n <= 2^32 for int32/uint32, 2^31 for
 * float32, 2^30 for int64/uint64, 2^29 for float64; a larger n returns
 * BULKX_ERR_WORKSPACE_TOO_SMALL before anything is enqueued.  Every call on
 * one GPU shares that GPU's array: two calls whose executions overlap on
 * one GPU (different streams, no ordering between them) race and give wrong
 * results; calls ordered on one stream are safe.
 */
int bulkx_inclusive_scan_noalloc(int dtype, int op, const void* in, void* out, int64_t n,
                                 const void* init, void* stream);

/*
 * exclusive_scan -- device-wide form of bulk::exclusive_scan(group, ...),
 * scan.hpp:553-591: result[0] = init, result[i] = init op x[0] op ... op x[i-1].
 * init must not be NULL.  in == out is allowed.
 */
int bulkx_exclusive_scan(int dtype, int op, const void* in, void* out, int64_t n,
                         const void* init, void* workspace, size_t workspace_bytes,
                         void* stream);

/*
 * reduce -- replaces T my_reduce(first, last, init, op), reduce.cu:58-97.
 * Writes init op x[0] op ... op x[n-1] to the DEVICE pointer out (one element);
 * n == 0 writes init.  The reference returns the value to the host with a
 * blocking copy (reduce.cu:96); here the caller decides when to read it.
 */
int bulkx_reduce(int dtype, int op, const void* in, int64_t n, const void* init,
                 void* out, void* workspace, size_t workspace_bytes, void* stream);

/*
 * merge -- replaces my_merge(first1, last1, first2, last2, result, comp),
 * merge.cu:127-162: stable merge of two sorted key ranges; on ties the element
 * of the first range comes first (merge.hpp:111, merge_path merge.hpp:35-60).
 * kout must hold n1 + n2 keys and must not alias the inputs.
 */
int bulkx_merge(int key_dtype, int comp, const void* keys1, int64_t n1,
                const void* keys2, int64_t n2, void* keys_out,
                void* workspace, size_t workspace_bytes, void* stream);

/*
 * merge_by_key -- device-wide form of bulk::merge_by_key (merge.hpp:507-607,
 * agent body merge.hpp:173-305): merge keys as bulkx_merge does and carry
 * each key's value along.
 */
int bulkx_merge_by_key(int key_dtype, int value_dtype, int comp,
                       const void* keys1, const void* values1, int64_t n1,
                       const void* keys2, const void* values2, int64_t n2,
                       void* keys_out, void* values_out,
                       void* workspace, size_t workspace_bytes, void* stream);

/*
 * stable_sort_by_key -- replaces stable_merge_sort_by_key(keys_first, keys_last,
 * values_first, comp), merge_sort_by_key.cu:150-209: in-place stable sort of
 * keys, values permuted alongside.  values may be NULL (keys only,
 * bulk::stable_sort, sort.hpp:139-149).
 */
int bulkx_stable_sort_by_key(int key_dtype, int value_dtype, int comp,
                             void* keys, void* values, int64_t n,
                             void* workspace, size_t workspace_bytes, void* stream);

/*
 * reduce_by_key -- replaces my_reduce_by_key(keys_first, keys_last, values_first,
 * keys_result, values_result, pred, op), reduce_by_key.cu:176-272.  One output
 * per maximal run of keys with pred(k[i-1], k[i]); the output key is the run's
 * last key (reduce_by_key.hpp:87-92), the value is the op-reduction of the run.
 * The number of runs is written to the DEVICE pointer num_runs (int64).
 * Outputs must hold up to n elements and must not alias the inputs.
 */
int bulkx_reduce_by_key(int key_dtype, int value_dtype, int pred, int op,
                        const void* keys, const void* values, int64_t n,
                        void* keys_out, void* values_out, int64_t* num_runs,
                        void* workspace, size_t workspace_bytes, void* stream);

/*
 * reduce_intervals -- replaces reduce_intervals(first, decomp, result, op),
 * reduce_intervals.hpp:31-40, over aligned_decomposition(n, num_partitions,
 * tile_size) (decomposition.hpp:142-218): result[g] is the op-reduction of
 * partition g.  tile_size = interval size and num_partitions = ceil(n/size)
 * gives the blocked form (reduce_intervals.hpp:43-47).  Empty partitions
 * (possible when num_partitions exceeds the tile count) produce the identity.
 */
int bulkx_reduce_intervals(int dtype, int op, const void* in, int64_t n,
                           int64_t num_partitions, int64_t tile_size, void* out,
                           void* stream);

/*
 * tail_flags / head_flags -- materialised forms of the reference's flag
 * iterators (tail_flags.hpp:26-109, head_flags.hpp:26-217):
 * tail[i] = (i == n-1) || !pred(k[i], k[i+1]);  head[i] = (i == 0) || !pred(k[i], k[i-1]).
 * Flags are written as uint8 0/1.
 */
int bulkx_tail_flags(int key_dtype, int pred, const void* keys, int64_t n, uint8_t* flags,
                     void* stream);
int bulkx_head_flags(int key_dtype, int pred, const void* keys, int64_t n, uint8_t* flags,
                     void* stream);

/* ------------------------------------------------------------------------
 * Batching: many uneven jobs run as segmented work -- one launch per op
 * family over the tiles of all its jobs, plus the sort radix passes --
 * recorded once into a hipGraph and replayed.  No reference counterpart: the reference issues one launch chain
 * per call.
 * ------------------------------------------------------------------------ */
enum bulkx_job_kind {
  BULKX_JOB_INCLUSIVE_SCAN = 0,
  BULKX_JOB_REDUCE_BY_KEY = 1,
  BULKX_JOB_MERGE_BY_KEY = 2,
  BULKX_JOB_SORT_BY_KEY = 3,
  BULKX_JOB_REDUCE = 4
};

/*
 * One job.  Device pointers; unused ones are NULL.
 *  scan:          in0 -> out0, n0 elements, dtype; op plus.
 *  reduce:        in0 (n0) -> out0 (one element); op plus.
 *  reduce_by_key: keys in0, values in1, n0 -> keys out0, values out1,
 *                 run count (int64) -> out2; key dtype int32, value dtype
 *                 value_dtype; equal_to / plus.
 *  merge_by_key:  keys in0 (n0) + in2 (n1), values in1 + in3 -> out0 / out1.
 *  sort_by_key:   keys out0, values out1 sorted in place (n0); in* unused.
 */
typedef struct bulkx_job {
  int32_t kind;
  int32_t dtype;
  int32_t value_dtype;
  int32_t reserved;
  int64_t n0;
  int64_t n1;
  const void* in0;
  const void* in1;
  const void* in2;
  const void* in3;
  void* out0;
  void* out1;
  void* out2;
} bulkx_job;

typedef struct bulkx_batch bulkx_batch;

/* Create a batch executor that can hold up to max_jobs jobs and up to
 * max_elements elements in total, bound to `stream`.  Each op family's tile
 * shape is read from the tuning table now (BULKX_ALGO_BATCH_*, size class 0;
 * shape 0 without an entry) and fixed for the batch's life. */
int bulkx_batch_create(int max_jobs, int64_t max_elements, void* stream, bulkx_batch** batch);
int bulkx_batch_destroy(bulkx_batch* batch);
/* Replace the batch's job list (host array, copied). */
int bulkx_batch_set_jobs(bulkx_batch* batch, const bulkx_job* jobs, int num_jobs);
/* Enqueue the batch on its stream: use_graph 1 replays the recorded graph
 * (four concurrent branches), 0 launches every kernel eagerly on the batch's
 * stream with per-stage timing events, 2 launches the graph's branches
 * eagerly on the batch's side streams (no graph, no timing). */
int bulkx_batch_launch(bulkx_batch* batch, int use_graph);
/* Run each job as its own plain call on the batch's stream (the per-call baseline). */
int bulkx_batch_launch_per_call(bulkx_batch* batch);
/* Waits for the batch's last launch and returns BULKX_ERR_LOOKBACK if one of
 * its look-back waits gave up (see bulkx_stream_status), else BULKX_SUCCESS. */
int bulkx_batch_status(bulkx_batch* batch);
/* Allocates now the scratch that job lists of up to `elements` elements can
 * need (sort ping-pong space, reduce_by_key run slots -- about 12 bytes per
 * element -- and the per-call path's workspace), so that later
 * bulkx_batch_set_jobs calls within that size allocate and free nothing. */
int bulkx_batch_reserve(bulkx_batch* batch, int64_t elements);
/* Bytes of the growable scratch the batch holds now: the sort / reduce_by_key
 * space (*tmp_bytes) and the per-call path's workspace (*workspace_bytes);
 * either pointer may be NULL.  A set_jobs call within the reserved size
 * leaves both unchanged.  No counterpart in the reference; for callers that
 * budget HBM. */
int bulkx_batch_scratch_bytes(bulkx_batch* batch, size_t* tmp_bytes, size_t* workspace_bytes);
/* Record the batch's graph ahead of time (it is recorded on first use
 * otherwise), so no later graph launch pays a capture.  No counterpart in the reference; for
 * latency-sensitive callers (the server calls it for each of its slots). */
int bulkx_batch_prepare_graphs(bulkx_batch* batch);
/* Device time in ms of the phases of the last eager launch (use_graph = 0):
 * the first-phase kernels (one per op family) and the sort radix passes
 * together.  Waits for that launch to finish.  BULKX_ERR_NOT_READY if
 * there is none. */
int bulkx_batch_last_timing(bulkx_batch* batch, float* phase0_ms, float* passes_ms);
/* Per-stage device time in ms of the last eager launch, in launch order:
 * [0] int32 scan/reduce, [1] float32 scan/reduce, [2..4] reduce_by_key with
 * int32 / float32 / int64 values, [5] merge_by_key, [6] sort jobs of one
 * block-sorted tile, [7] the radix passes of the larger sort jobs; then the
 * radix passes split by kernel: [8] downsweeps, [9] upsweeps, [10] counts
 * scans.  Writes num_stages entries (zeros beyond 11). */
int bulkx_batch_stage_timing(bulkx_batch* batch, float* stage_ms, int num_stages);
/* Sort geometry of the batch: a sort job of at most *run_elements keys is
 * sorted by one block-sort tile in LDS; a larger one by four radix passes
 * over tiles of *radix_tile keys (each pass reads and writes the job once).
 * No counterpart in the reference (its merge_sort_by_key.cu:163-171 tile is
 * fixed and its merge passes grow with log2 n). */
int bulkx_batch_sort_geometry(int64_t* run_elements, int64_t* radix_tile);
/* The tile shape each op family of this batch runs, in the order of
 * bulkx_batch_stage_timing's first seven stages (int32 scan/reduce, float32
 * scan/reduce, reduce_by_key with int32 / float32 / int64 values, merge,
 * one-tile sorts): shape[i] (0 or 1) and tile_elements[i] (elements per
 * workgroup).  Either pointer may be NULL.  No counterpart in the
 * reference, whose tile sizes are constants in each driver. */
int bulkx_batch_family_shapes(bulkx_batch* batch, int* shape, int64_t* tile_elements);

/* ------------------------------------------------------------------------
 * Serving: a request loop that admits jobs under an HBM budget, keeps their
 * inputs and outputs device-resident, batches the requests that arrive while
 * earlier ones run, and records per-request latency.
 * ------------------------------------------------------------------------ */
typedef struct bulkx_server bulkx_server;

/* Server flags. */
enum bulkx_serve_flags {
  BULKX_SERVE_PER_CALL = 1, /* launch every request alone through the plain entry points */
  BULKX_SERVE_NO_GRAPH = 2  /* launch batches without hipGraph replay */
};

/* hbm_budget_bytes bounds the device memory of admitted-but-unfinished jobs;
 * max_inflight bounds the batches on the device at once (one stream each);
 * a batch holds at most max_batch_jobs jobs and max_batch_elements elements. */
int bulkx_server_create(int64_t hbm_budget_bytes, int max_inflight, int max_batch_jobs,
                        int64_t max_batch_elements, int flags, bulkx_server** server);
int bulkx_server_destroy(bulkx_server* server);
/* Submit a job whose data already resides in device memory.  `bytes` is its
 * device footprint charged to the budget.  Returns an id through *request_id;
 * BULKX_ERR_BUDGET if the job alone exceeds the budget.  The server's streams
 * are non-blocking: the job's inputs must be complete on the device when it
 * is submitted (the caller synchronizes the stream that produced them). */
int bulkx_server_submit(bulkx_server* server, const bulkx_job* job, int64_t bytes,
                        int64_t* request_id);
/* Like bulkx_server_submit, but the request arrives at `arrival_us` on the
 * server clock (microseconds since creation; requests must be submitted in
 * arrival order).  The dispatcher does not see it before that time, so an
 * open-loop arrival process can be replayed exactly. */
int bulkx_server_submit_at(bulkx_server* server, const bulkx_job* job, int64_t bytes,
                           double arrival_us, int64_t* request_id);
/* Submit a job whose buffers are in host memory (pinned with hipHostMalloc
 * for the copies to run asynchronously; pageable memory works but its
 * copies hold the server's worker).  When the request is admitted the server
 * allocates its device buffers, copies the inputs in ahead of its batch and
 * the outputs back behind it (reduce_by_key: all n0 slots of keys and
 * values, and the count), then frees the buffers, all on the slot's stream;
 * the request is complete when its results are in host memory.  The host
 * buffers must stay valid until then.  *device_bytes (if not NULL) receives
 * the footprint charged to the budget.  arrival_us < 0 means now.  No
 * counterpart in the reference (its drivers take device ranges only): for
 * callers whose data starts and ends on the host. */
int bulkx_server_submit_host(bulkx_server* server, const bulkx_job* job, double arrival_us,
                             int64_t* request_id, int64_t* device_bytes);
/* Microseconds elapsed on the server clock. */
double bulkx_server_now_us(bulkx_server* server);
/* Block until every submitted request has completed. */
int bulkx_server_drain(bulkx_server* server);
/* Per-request timing in microseconds since server creation: arrival (submit),
 * launch and completion.  0 when unknown yet. */
int bulkx_server_request_times(bulkx_server* server, int64_t request_id,
                               double* t_arrival_us, double* t_launch_us, double* t_done_us);
/* Counters: batches launched, requests completed, peak admitted bytes. */
int bulkx_server_stats(bulkx_server* server, int64_t* batches, int64_t* completed,
                       int64_t* peak_bytes);

#ifdef __cplusplus
}
#endif

#endif /* BULKX_H */
