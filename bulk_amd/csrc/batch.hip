// Batch executor: many uneven scan / reduce / reduce_by_key / merge_by_key /
// sort_by_key jobs run as segmented work in one persistent launch (plus one
// launch per sort merge round), recorded once into a hipGraph.
//
// Work decomposition.  Every job is cut into fixed-size tiles; the tiles of
// all jobs form one global queue, ordered by job.  A persistent grid of
// resident workgroups claims tiles from an atomic counter, maps the tile to
// its job through a device-resident tile->job table and runs the same tile
// body the device-wide kernel runs (scan_tile, rbk_tile, merge_two_tile,
// block_sort_tile).  Scan, reduce and reduce_by_key jobs keep a decoupled
// look-back chain of their own (per-job status words); because tiles are
// claimed in queue order, a tile's predecessors always belong to running or
// finished workgroups.  Sort jobs need global order between merge rounds,
// so each round is one more persistent launch over the sort tiles.
//
// Everything a launch needs (tile counts, tables, counters) lives in device
// memory, so the captured graph stays valid when the job list changes: a new
// list is uploaded with stream-ordered copies and the graph is replayed.
#include <algorithm>
#include <vector>

#include "host_common.hpp"
#include "merge_impl.hpp"
#include "reduce_by_key_impl.hpp"
#include "scan_impl.hpp"
#include "sort_impl.hpp"

namespace bulkx {

constexpr int kBatchBlock = 512;
constexpr int kBatchScanItems = 32;   // 16384 elements per scan / reduce tile
constexpr int kBatchRbkItems = 8;     // 4096 elements per reduce_by_key tile
constexpr int kBatchMergeItems = 8;   // 4096 outputs per merge / sort tile
constexpr int64_t kBatchScanTile = int64_t(kBatchBlock) * kBatchScanItems;
constexpr int64_t kBatchRbkTile = int64_t(kBatchBlock) * kBatchRbkItems;
constexpr int64_t kBatchMergeTile = int64_t(kBatchBlock) * kBatchMergeItems;
constexpr int kBatchMaxPasses = 24;
constexpr int kStatusWordsPerTile = 6;  // widest look-back record used here (24 bytes)

enum BatchOp : int32_t {
  OP_NONE = 0,
  OP_SCAN_I32, OP_SCAN_F32,
  OP_REDUCE_I32, OP_REDUCE_F32,
  OP_RBK_I32_I32, OP_RBK_I32_F32, OP_RBK_I32_I64,
  OP_MERGE_I32, OP_MERGE_F32,
  OP_SORT_I32, OP_SORT_U32, OP_SORT_F32
};

struct BatchJobDev {
  int32_t op;
  int32_t passes;  // sort: merge rounds
  int64_t n0, n1;
  const void* in0;
  const void* in1;
  const void* in2;
  const void* in3;
  void* out0;
  void* out1;
  void* out2;
  uint64_t* status;  // look-back words (null for single-tile jobs)
  void* tmp_k;       // sort ping-pong buffers
  void* tmp_v;
  int64_t tiles0;       // phase-0 tiles
  int64_t tile_begin0;  // first global phase-0 tile
  int64_t sort_tile_begin;
};

struct BatchDesc {
  int64_t total_tiles0;
  int64_t total_sort_tiles;
  int64_t num_status_words;
  uint32_t counter0;
  uint32_t err;
  uint32_t pass_counter[kBatchMaxPasses];
};

union BatchSmem {
  ScanSmem<int32_t, OpPlus, kBatchBlock> scan_i;
  ScanSmem<float, OpPlus, kBatchBlock> scan_f;
  RbkSmem<int32_t, int32_t, OpPlus, kBatchBlock, kBatchRbkItems> rbk_i;
  RbkSmem<int32_t, float, OpPlus, kBatchBlock, kBatchRbkItems> rbk_f;
  RbkSmem<int32_t, int64_t, OpPlus, kBatchBlock, kBatchRbkItems> rbk_l;
  MergeSmem<int32_t, uint32_t, true, kBatchBlock, kBatchMergeItems> merge_i;
  MergeSmem<float, uint32_t, true, kBatchBlock, kBatchMergeItems> merge_f;
  MergeSmem<uint32_t, uint32_t, true, kBatchBlock, kBatchMergeItems> merge_u;
  SortSmem<int32_t, uint32_t, true, kBatchBlock, kBatchMergeItems> sort_i;
  SortSmem<uint32_t, uint32_t, true, kBatchBlock, kBatchMergeItems> sort_u;
  SortSmem<float, uint32_t, true, kBatchBlock, kBatchMergeItems> sort_f;
};

__global__ void batch_clear_kernel(BatchDesc* d, uint64_t* status) {
  const int64_t nw = d->num_status_words;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nw;
       i += int64_t(gridDim.x) * blockDim.x)
    status[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    d->counter0 = 0;
    d->err = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < kBatchMaxPasses) d->pass_counter[threadIdx.x] = 0;
}

template <class K>
__device__ __forceinline__ void sort_block_job(BatchSmem& sm, const BatchJobDev& job, int64_t lt);

template <class K>
__device__ __forceinline__ SortSmem<K, uint32_t, true, kBatchBlock, kBatchMergeItems>& sort_smem(
    BatchSmem& sm);
template <>
__device__ __forceinline__ SortSmem<int32_t, uint32_t, true, kBatchBlock, kBatchMergeItems>&
sort_smem<int32_t>(BatchSmem& sm) { return sm.sort_i; }
template <>
__device__ __forceinline__ SortSmem<uint32_t, uint32_t, true, kBatchBlock, kBatchMergeItems>&
sort_smem<uint32_t>(BatchSmem& sm) { return sm.sort_u; }
template <>
__device__ __forceinline__ SortSmem<float, uint32_t, true, kBatchBlock, kBatchMergeItems>&
sort_smem<float>(BatchSmem& sm) { return sm.sort_f; }

template <class K>
__device__ __forceinline__ void sort_block_job(BatchSmem& sm, const BatchJobDev& job, int64_t lt) {
  K* keys = static_cast<K*>(job.out0);
  uint32_t* vals = static_cast<uint32_t*>(job.out1);
  // an odd number of rounds starts in the temporary so the last lands in place
  K* dk = (job.passes & 1) ? static_cast<K*>(job.tmp_k) : keys;
  uint32_t* dv = (job.passes & 1) ? static_cast<uint32_t*>(job.tmp_v) : vals;
  block_sort_tile<CmpLess, K, uint32_t, true, kBatchBlock, kBatchMergeItems>(
      sort_smem<K>(sm), keys, vals, job.n0, dk, dv, lt);
}

__global__ __launch_bounds__(kBatchBlock) void batch_phase0_kernel(BatchDesc* d,
                                                                   const BatchJobDev* jobs,
                                                                   const uint32_t* tile_job) {
  __shared__ BatchSmem sm;
  __shared__ uint32_t s_t;
  const int64_t total = d->total_tiles0;
  while (true) {
    __syncthreads();  // the previous tile's LDS traffic is over
    if (threadIdx.x == 0) s_t = atomicAdd(&d->counter0, 1u);
    __syncthreads();
    const int64_t t = s_t;
    if (t >= total) break;
    const BatchJobDev& job = jobs[tile_job[t]];
    const int64_t lt = t - job.tile_begin0;
    const bool single = job.tiles0 == 1;
    switch (job.op) {
      case OP_SCAN_I32:
        scan_tile<int32_t, OpPlus, kBatchBlock, kBatchScanItems, false>(
            sm.scan_i, static_cast<const int32_t*>(job.in0), static_cast<int32_t*>(job.out0), job.n0,
            0, false, lt, job.status, &d->err, single, nullptr);
        break;
      case OP_SCAN_F32:
        scan_tile<float, OpPlus, kBatchBlock, kBatchScanItems, false>(
            sm.scan_f, static_cast<const float*>(job.in0), static_cast<float*>(job.out0), job.n0,
            0.f, false, lt, job.status, &d->err, single, nullptr);
        break;
      case OP_REDUCE_I32:
        scan_tile<int32_t, OpPlus, kBatchBlock, kBatchScanItems, false, false>(
            sm.scan_i, static_cast<const int32_t*>(job.in0), nullptr, job.n0, 0, false, lt,
            job.status, &d->err, single, static_cast<int32_t*>(job.out0));
        break;
      case OP_REDUCE_F32:
        scan_tile<float, OpPlus, kBatchBlock, kBatchScanItems, false, false>(
            sm.scan_f, static_cast<const float*>(job.in0), nullptr, job.n0, 0.f, false, lt,
            job.status, &d->err, single, static_cast<float*>(job.out0));
        break;
      case OP_RBK_I32_I32:
        rbk_tile<int32_t, int32_t, OpPlus, kBatchBlock, kBatchRbkItems>(
            sm.rbk_i, static_cast<const int32_t*>(job.in0), static_cast<const int32_t*>(job.in1),
            job.n0, static_cast<int32_t*>(job.out0), static_cast<int32_t*>(job.out1),
            static_cast<int64_t*>(job.out2), lt, job.status, &d->err, single);
        break;
      case OP_RBK_I32_F32:
        rbk_tile<int32_t, float, OpPlus, kBatchBlock, kBatchRbkItems>(
            sm.rbk_f, static_cast<const int32_t*>(job.in0), static_cast<const float*>(job.in1),
            job.n0, static_cast<int32_t*>(job.out0), static_cast<float*>(job.out1),
            static_cast<int64_t*>(job.out2), lt, job.status, &d->err, single);
        break;
      case OP_RBK_I32_I64:
        rbk_tile<int32_t, int64_t, OpPlus, kBatchBlock, kBatchRbkItems>(
            sm.rbk_l, static_cast<const int32_t*>(job.in0), static_cast<const int64_t*>(job.in1),
            job.n0, static_cast<int32_t*>(job.out0), static_cast<int64_t*>(job.out1),
            static_cast<int64_t*>(job.out2), lt, job.status, &d->err, single);
        break;
      case OP_MERGE_I32:
        merge_two_tile<CmpLess, int32_t, uint32_t, true, kBatchBlock, kBatchMergeItems>(
            sm.merge_i, static_cast<const int32_t*>(job.in0), static_cast<const uint32_t*>(job.in1),
            job.n0, static_cast<const int32_t*>(job.in2), static_cast<const uint32_t*>(job.in3),
            job.n1, static_cast<int32_t*>(job.out0), static_cast<uint32_t*>(job.out1), lt);
        break;
      case OP_MERGE_F32:
        merge_two_tile<CmpLess, float, uint32_t, true, kBatchBlock, kBatchMergeItems>(
            sm.merge_f, static_cast<const float*>(job.in0), static_cast<const uint32_t*>(job.in1),
            job.n0, static_cast<const float*>(job.in2), static_cast<const uint32_t*>(job.in3),
            job.n1, static_cast<float*>(job.out0), static_cast<uint32_t*>(job.out1), lt);
        break;
      case OP_SORT_I32: sort_block_job<int32_t>(sm, job, lt); break;
      case OP_SORT_U32: sort_block_job<uint32_t>(sm, job, lt); break;
      case OP_SORT_F32: sort_block_job<float>(sm, job, lt); break;
      default: break;
    }
  }
}

template <class K>
__device__ __forceinline__ MergeSmem<K, uint32_t, true, kBatchBlock, kBatchMergeItems>& pass_smem(
    BatchSmem& sm);
template <>
__device__ __forceinline__ MergeSmem<int32_t, uint32_t, true, kBatchBlock, kBatchMergeItems>&
pass_smem<int32_t>(BatchSmem& sm) { return sm.merge_i; }
template <>
__device__ __forceinline__ MergeSmem<float, uint32_t, true, kBatchBlock, kBatchMergeItems>&
pass_smem<float>(BatchSmem& sm) { return sm.merge_f; }
template <>
__device__ __forceinline__ MergeSmem<uint32_t, uint32_t, true, kBatchBlock, kBatchMergeItems>&
pass_smem<uint32_t>(BatchSmem& sm) { return sm.merge_u; }

template <class K>
__device__ __forceinline__ void sort_pass_job(BatchSmem& sm, const BatchJobDev& job, int pass,
                                              int64_t lt) {
  K* orig_k = static_cast<K*>(job.out0);
  uint32_t* orig_v = static_cast<uint32_t*>(job.out1);
  K* tmp_k = static_cast<K*>(job.tmp_k);
  uint32_t* tmp_v = static_cast<uint32_t*>(job.tmp_v);
  // the block sort wrote to "first"; rounds alternate from there
  const bool first_is_tmp = (job.passes & 1) != 0;
  const bool src_is_tmp = first_is_tmp ^ ((pass & 1) != 0);
  const K* sk = src_is_tmp ? tmp_k : orig_k;
  const uint32_t* sv = src_is_tmp ? tmp_v : orig_v;
  K* dk = src_is_tmp ? orig_k : tmp_k;
  uint32_t* dv = src_is_tmp ? orig_v : tmp_v;
  merge_pass_tile<CmpLess, K, uint32_t, true, kBatchBlock, kBatchMergeItems>(
      pass_smem<K>(sm), sk, sv, job.n0, kBatchMergeTile << pass, dk, dv, lt);
}

__global__ __launch_bounds__(kBatchBlock) void batch_pass_kernel(BatchDesc* d,
                                                                 const BatchJobDev* jobs,
                                                                 const uint32_t* sort_tile_job,
                                                                 int pass) {
  __shared__ BatchSmem sm;
  __shared__ uint32_t s_t;
  const int64_t total = d->total_sort_tiles;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) s_t = atomicAdd(&d->pass_counter[pass], 1u);
    __syncthreads();
    const int64_t t = s_t;
    if (t >= total) break;
    const BatchJobDev& job = jobs[sort_tile_job[t]];
    if (pass >= job.passes) continue;
    const int64_t lt = t - job.sort_tile_begin;
    switch (job.op) {
      case OP_SORT_I32: sort_pass_job<int32_t>(sm, job, pass, lt); break;
      case OP_SORT_U32: sort_pass_job<uint32_t>(sm, job, pass, lt); break;
      case OP_SORT_F32: sort_pass_job<float>(sm, job, pass, lt); break;
      default: break;
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static int passes_for(int64_t tiles) {
  int p = 0;
  while ((int64_t(1) << p) < tiles) ++p;
  return p;
}

static int batch_op(const bulkx_job& j) {
  switch (j.kind) {
    case BULKX_JOB_INCLUSIVE_SCAN:
      return j.dtype == BULKX_INT32 ? OP_SCAN_I32 : j.dtype == BULKX_FLOAT32 ? OP_SCAN_F32 : OP_NONE;
    case BULKX_JOB_REDUCE:
      return j.dtype == BULKX_INT32 ? OP_REDUCE_I32 : j.dtype == BULKX_FLOAT32 ? OP_REDUCE_F32 : OP_NONE;
    case BULKX_JOB_REDUCE_BY_KEY:
      if (j.dtype != BULKX_INT32) return OP_NONE;
      return j.value_dtype == BULKX_INT32   ? OP_RBK_I32_I32
             : j.value_dtype == BULKX_FLOAT32 ? OP_RBK_I32_F32
             : j.value_dtype == BULKX_INT64   ? OP_RBK_I32_I64
                                              : OP_NONE;
    case BULKX_JOB_MERGE_BY_KEY:
      if (dtype_size(j.value_dtype) != 4) return OP_NONE;
      return j.dtype == BULKX_INT32 ? OP_MERGE_I32 : j.dtype == BULKX_FLOAT32 ? OP_MERGE_F32 : OP_NONE;
    case BULKX_JOB_SORT_BY_KEY:
      if (dtype_size(j.value_dtype) != 4) return OP_NONE;
      return j.dtype == BULKX_INT32     ? OP_SORT_I32
             : j.dtype == BULKX_UINT32  ? OP_SORT_U32
             : j.dtype == BULKX_FLOAT32 ? OP_SORT_F32
                                        : OP_NONE;
    default: return OP_NONE;
  }
}

// Workspace one plain entry point needs for job j.
static size_t per_call_workspace(const bulkx_job& j) {
  switch (j.kind) {
    case BULKX_JOB_INCLUSIVE_SCAN: return scan_workspace_bytes(j.dtype, j.n0);
    case BULKX_JOB_REDUCE: return reduce_workspace_bytes(j.dtype, j.n0);
    case BULKX_JOB_REDUCE_BY_KEY: return reduce_by_key_workspace_bytes(j.dtype, j.value_dtype, j.n0);
    case BULKX_JOB_SORT_BY_KEY: return sort_workspace_bytes(j.dtype, j.value_dtype, j.n0);
    default: return 0;
  }
}

}  // namespace bulkx

using namespace bulkx;

struct bulkx_batch {
  hipStream_t stream = nullptr;
  hipStream_t capture_stream = nullptr;  // graphs are recorded here (capture is not
                                         // permitted on the null stream)
  int max_jobs = 0;
  int64_t max_elements = 0;
  int64_t max_tiles = 0;
  int max_passes = 0;
  int grid = 0;
  // device
  BatchDesc* d_desc = nullptr;
  BatchJobDev* d_jobs = nullptr;
  uint32_t* d_tile_job = nullptr;
  uint32_t* d_sort_tile_job = nullptr;
  uint64_t* d_status = nullptr;
  char* d_tmp = nullptr;
  size_t status_words_cap = 0;
  size_t tmp_bytes_cap = 0;
  char* d_ws = nullptr;  // workspace of the per-call path
  size_t ws_bytes = 0;
  // pinned host staging
  BatchDesc* h_desc = nullptr;
  BatchJobDev* h_jobs = nullptr;
  uint32_t* h_tile_job = nullptr;
  uint32_t* h_sort_tile_job = nullptr;
  // current job list
  std::vector<bulkx_job> jobs;
  int num_jobs = 0;
  int64_t total_tiles0 = 0, total_sort_tiles = 0;
  bool pending_upload = false;
  // phase timing of eager launches
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  bool timed = false;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int batch_enqueue_kernels(bulkx_batch* b, hipStream_t s, bool timed) {
  batch_clear_kernel<<<256, 256, 0, s>>>(b->d_desc, b->d_status);
  BULKX_LAUNCHED();
  if (timed) BULKX_HIP(hipEventRecord(b->ev[0], s));
  batch_phase0_kernel<<<b->grid, kBatchBlock, 0, s>>>(b->d_desc, b->d_jobs, b->d_tile_job);
  BULKX_LAUNCHED();
  if (timed) BULKX_HIP(hipEventRecord(b->ev[1], s));
  for (int p = 0; p < b->max_passes; ++p) {
    batch_pass_kernel<<<b->grid, kBatchBlock, 0, s>>>(b->d_desc, b->d_jobs, b->d_sort_tile_job, p);
    BULKX_LAUNCHED();
  }
  if (timed) BULKX_HIP(hipEventRecord(b->ev[2], s));
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_create(int max_jobs, int64_t max_elements, void* stream,
                                  bulkx_batch** out) {
  if (!out || max_jobs <= 0 || max_elements <= 0) return BULKX_ERR_INVALID_ARGUMENT;
  DeviceProps props;
  BULKX_TRY(current_device_props(&props));
  bulkx_batch* b = new bulkx_batch();
  b->stream = static_cast<hipStream_t>(stream);
  b->max_jobs = max_jobs;
  b->max_elements = max_elements;
  // every job can add one partial tile
  b->max_tiles = max_elements / kBatchRbkTile + 2 * int64_t(max_jobs) + 1;
  b->max_passes = passes_for((max_elements + kBatchMergeTile - 1) / kBatchMergeTile);
  if (b->max_passes > kBatchMaxPasses) {
    delete b;
    return BULKX_ERR_INVALID_ARGUMENT;
  }
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_phase0_kernel,
                                                              kBatchBlock, 0);
  if (e != hipSuccess || per_cu < 1) per_cu = 1;
  b->grid = per_cu * props.num_cus;
  b->status_words_cap = size_t(b->max_tiles) * kStatusWordsPerTile;
  b->tmp_bytes_cap = 0;  // sort ping-pong space grows on demand in set_jobs
  b->ws_bytes = 256;  // the per-call workspace grows on demand in set_jobs
  int st = BULKX_SUCCESS;
  auto dmalloc = [&](void** p, size_t bytes) {
    if (st != BULKX_SUCCESS) return;
    hipError_t r = hipMalloc(p, bytes);
    if (r != hipSuccess) {
      set_last_error("hipMalloc", r);
      st = r == hipErrorOutOfMemory ? BULKX_ERR_OUT_OF_MEMORY : BULKX_ERR_HIP;
    }
  };
  auto hmalloc = [&](void** p, size_t bytes) {
    if (st != BULKX_SUCCESS) return;
    hipError_t r = hipHostMalloc(p, bytes, hipHostMallocDefault);
    if (r != hipSuccess) {
      set_last_error("hipHostMalloc", r);
      st = BULKX_ERR_HIP;
    }
  };
  dmalloc(reinterpret_cast<void**>(&b->d_desc), sizeof(BatchDesc));
  dmalloc(reinterpret_cast<void**>(&b->d_jobs), sizeof(BatchJobDev) * max_jobs);
  dmalloc(reinterpret_cast<void**>(&b->d_tile_job), sizeof(uint32_t) * b->max_tiles);
  dmalloc(reinterpret_cast<void**>(&b->d_sort_tile_job), sizeof(uint32_t) * b->max_tiles);
  dmalloc(reinterpret_cast<void**>(&b->d_status), sizeof(uint64_t) * b->status_words_cap);
  dmalloc(reinterpret_cast<void**>(&b->d_ws), b->ws_bytes);
  hmalloc(reinterpret_cast<void**>(&b->h_desc), sizeof(BatchDesc));
  hmalloc(reinterpret_cast<void**>(&b->h_jobs), sizeof(BatchJobDev) * max_jobs);
  hmalloc(reinterpret_cast<void**>(&b->h_tile_job), sizeof(uint32_t) * b->max_tiles);
  hmalloc(reinterpret_cast<void**>(&b->h_sort_tile_job), sizeof(uint32_t) * b->max_tiles);
  if (st != BULKX_SUCCESS) {
    bulkx_batch_destroy(b);
    return st;
  }
  memset(b->h_desc, 0, sizeof(BatchDesc));
  hipError_t r = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
  for (int i = 0; i < 3 && r == hipSuccess; ++i) r = hipEventCreate(&b->ev[i]);
  if (r != hipSuccess) {
    set_last_error("hipStreamCreateWithFlags", r);
    bulkx_batch_destroy(b);
    return BULKX_ERR_HIP;
  }
  r = hipMemcpy(b->d_desc, b->h_desc, sizeof(BatchDesc), hipMemcpyHostToDevice);
  if (r != hipSuccess) {
    set_last_error("hipMemcpy", r);
    bulkx_batch_destroy(b);
    return BULKX_ERR_HIP;
  }
  *out = b;
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_destroy(bulkx_batch* b) {
  if (!b) return BULKX_SUCCESS;
  (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void* p : {static_cast<void*>(b->d_desc), static_cast<void*>(b->d_jobs),
                  static_cast<void*>(b->d_tile_job), static_cast<void*>(b->d_sort_tile_job),
                  static_cast<void*>(b->d_status), static_cast<void*>(b->d_tmp),
                  static_cast<void*>(b->d_ws)})
    if (p) (void)hipFree(p);
  for (void* p : {static_cast<void*>(b->h_desc), static_cast<void*>(b->h_jobs),
                  static_cast<void*>(b->h_tile_job), static_cast<void*>(b->h_sort_tile_job)})
    if (p) (void)hipHostFree(p);
  delete b;
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_set_jobs(bulkx_batch* b, const bulkx_job* jobs, int num_jobs) {
  if (!b || num_jobs < 0 || num_jobs > b->max_jobs || (num_jobs > 0 && !jobs))
    return BULKX_ERR_INVALID_ARGUMENT;
  // Validate and lay out on the host first; nothing is touched on failure.
  int64_t elements = 0, tiles0 = 0, sort_tiles = 0;
  size_t status_words = 0, tmp_bytes = 0, ws_need = 256;
  for (int i = 0; i < num_jobs; ++i) {
    const bulkx_job& j = jobs[i];
    const int op = batch_op(j);
    if (op == OP_NONE) return BULKX_ERR_UNSUPPORTED_DTYPE;
    if (j.n0 < 0 || j.n1 < 0) return BULKX_ERR_INVALID_ARGUMENT;
    const int64_t n = j.n0 + (op == OP_MERGE_I32 || op == OP_MERGE_F32 ? j.n1 : 0);
    elements += n;
    ws_need = std::max(ws_need, per_call_workspace(j));
    if (n == 0 && j.kind != BULKX_JOB_REDUCE && j.kind != BULKX_JOB_REDUCE_BY_KEY) continue;
    const bool needs_out2 = j.kind == BULKX_JOB_REDUCE_BY_KEY;
    if (!j.out0 || (needs_out2 && (!j.out1 || !j.out2)) ||
        (n > 0 && !j.in0 && j.kind != BULKX_JOB_SORT_BY_KEY))
      return BULKX_ERR_INVALID_ARGUMENT;
    int64_t t;
    switch (op) {
      case OP_SCAN_I32: case OP_SCAN_F32: case OP_REDUCE_I32: case OP_REDUCE_F32:
        t = (n + kBatchScanTile - 1) / kBatchScanTile;
        if (t > 1) status_words += size_t(t) * kStatusWordsPerTile;
        break;
      case OP_RBK_I32_I32: case OP_RBK_I32_F32: case OP_RBK_I32_I64:
        t = (n + kBatchRbkTile - 1) / kBatchRbkTile;
        if (t > 1) status_words += size_t(t) * kStatusWordsPerTile;
        break;
      default:
        t = (n + kBatchMergeTile - 1) / kBatchMergeTile;
        if (op >= OP_SORT_I32) {
          sort_tiles += t;
          if (t > 1) tmp_bytes += align_up(size_t(n) * 4, 256) * 2;
        }
        break;
    }
    if (t < 1) t = 1;  // empty reduce / reduce_by_key: one tile writes the result
    tiles0 += t;
  }
  if (elements > b->max_elements || tiles0 > b->max_tiles || sort_tiles > b->max_tiles ||
      status_words > b->status_words_cap)
    return BULKX_ERR_BUDGET;

  // The staging buffers may still feed the previous upload.
  BULKX_HIP(hipStreamSynchronize(b->stream));
  if (tmp_bytes > b->tmp_bytes_cap) {  // grow the sort ping-pong space (never inside a launch)
    if (b->d_tmp) BULKX_HIP(hipFree(b->d_tmp));
    b->d_tmp = nullptr;
    b->tmp_bytes_cap = 0;
    const size_t want = tmp_bytes + tmp_bytes / 4;
    BULKX_HIP(hipMalloc(reinterpret_cast<void**>(&b->d_tmp), want));
    b->tmp_bytes_cap = want;
  }
  if (ws_need > b->ws_bytes) {  // grow the workspace of the per-call path
    if (b->d_ws) BULKX_HIP(hipFree(b->d_ws));
    b->d_ws = nullptr;
    b->ws_bytes = 0;
    BULKX_HIP(hipMalloc(reinterpret_cast<void**>(&b->d_ws), ws_need));
    b->ws_bytes = ws_need;
  }
  b->jobs.assign(jobs, jobs + num_jobs);
  b->num_jobs = num_jobs;
  int64_t t0 = 0, ts = 0;
  size_t sw = 0, tb = 0;
  for (int i = 0; i < num_jobs; ++i) {
    const bulkx_job& j = jobs[i];
    const int op = batch_op(j);
    BatchJobDev& dj = b->h_jobs[i];
    memset(&dj, 0, sizeof(dj));
    dj.op = op;
    dj.n0 = j.n0;
    dj.n1 = j.n1;
    dj.in0 = j.in0;
    dj.in1 = j.in1;
    dj.in2 = j.in2;
    dj.in3 = j.in3;
    dj.out0 = j.out0;
    dj.out1 = j.out1;
    dj.out2 = j.out2;
    const int64_t n = j.n0 + (op == OP_MERGE_I32 || op == OP_MERGE_F32 ? j.n1 : 0);
    int64_t t;
    if (op <= OP_REDUCE_F32) t = (n + kBatchScanTile - 1) / kBatchScanTile;
    else if (op <= OP_RBK_I32_I64) t = (n + kBatchRbkTile - 1) / kBatchRbkTile;
    else t = (n + kBatchMergeTile - 1) / kBatchMergeTile;
    const bool chained = op <= OP_RBK_I32_I64;
    if (n == 0 && !(j.kind == BULKX_JOB_REDUCE || j.kind == BULKX_JOB_REDUCE_BY_KEY)) {
      dj.op = OP_NONE;
      dj.tiles0 = 0;
      continue;
    }
    if (t < 1) t = 1;
    dj.tiles0 = t;
    dj.tile_begin0 = t0;
    if (chained && t > 1) {
      dj.status = b->d_status + sw;
      sw += size_t(t) * kStatusWordsPerTile;
    }
    if (op >= OP_SORT_I32) {
      dj.passes = passes_for(t);
      dj.sort_tile_begin = ts;
      if (t > 1) {
        dj.tmp_k = b->d_tmp + tb;
        tb += align_up(size_t(n) * 4, 256);
        dj.tmp_v = b->d_tmp + tb;
        tb += align_up(size_t(n) * 4, 256);
      }
      for (int64_t k = 0; k < t; ++k) b->h_sort_tile_job[ts + k] = uint32_t(i);
      ts += t;
    }
    for (int64_t k = 0; k < t; ++k) b->h_tile_job[t0 + k] = uint32_t(i);
    t0 += t;
  }
  memset(b->h_desc, 0, sizeof(BatchDesc));
  b->h_desc->total_tiles0 = t0;
  b->h_desc->total_sort_tiles = ts;
  b->h_desc->num_status_words = int64_t(sw);
  b->total_tiles0 = t0;
  b->total_sort_tiles = ts;
  BULKX_HIP(hipMemcpyAsync(b->d_jobs, b->h_jobs, sizeof(BatchJobDev) * (num_jobs > 0 ? num_jobs : 1),
                           hipMemcpyHostToDevice, b->stream));
  if (t0 > 0)
    BULKX_HIP(hipMemcpyAsync(b->d_tile_job, b->h_tile_job, sizeof(uint32_t) * t0,
                             hipMemcpyHostToDevice, b->stream));
  if (ts > 0)
    BULKX_HIP(hipMemcpyAsync(b->d_sort_tile_job, b->h_sort_tile_job, sizeof(uint32_t) * ts,
                             hipMemcpyHostToDevice, b->stream));
  BULKX_HIP(hipMemcpyAsync(b->d_desc, b->h_desc, sizeof(BatchDesc), hipMemcpyHostToDevice,
                           b->stream));
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_launch(bulkx_batch* b, int use_graph) {
  if (!b) return BULKX_ERR_INVALID_ARGUMENT;
  if (!use_graph) {
    b->timed = true;
    return batch_enqueue_kernels(b, b->stream, true);
  }
  if (!b->exec) {
    BULKX_HIP(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
    int st = batch_enqueue_kernels(b, b->capture_stream, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->capture_stream, &g);
    if (st != BULKX_SUCCESS) return st;
    if (e != hipSuccess) {
      set_last_error("hipStreamEndCapture", e);
      return BULKX_ERR_HIP;
    }
    b->graph = g;
    BULKX_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  }
  BULKX_HIP(hipGraphLaunch(b->exec, b->stream));
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_launch_per_call(bulkx_batch* b) {
  if (!b) return BULKX_ERR_INVALID_ARGUMENT;
  void* s = b->stream;
  for (const bulkx_job& j : b->jobs) {
    int st = BULKX_SUCCESS;
    switch (j.kind) {
      case BULKX_JOB_INCLUSIVE_SCAN:
        st = bulkx_inclusive_scan(j.dtype, BULKX_PLUS, j.in0, j.out0, j.n0, nullptr, b->d_ws,
                                  b->ws_bytes, s);
        break;
      case BULKX_JOB_REDUCE:
        st = bulkx_reduce(j.dtype, BULKX_PLUS, j.in0, j.n0, nullptr, j.out0, b->d_ws, b->ws_bytes, s);
        break;
      case BULKX_JOB_REDUCE_BY_KEY:
        st = bulkx_reduce_by_key(j.dtype, j.value_dtype, BULKX_EQUAL_TO, BULKX_PLUS, j.in0, j.in1,
                                 j.n0, j.out0, j.out1, static_cast<int64_t*>(j.out2), b->d_ws,
                                 b->ws_bytes, s);
        break;
      case BULKX_JOB_MERGE_BY_KEY:
        st = bulkx_merge_by_key(j.dtype, j.value_dtype, BULKX_LESS, j.in0, j.in1, j.n0, j.in2,
                                j.in3, j.n1, j.out0, j.out1, nullptr, 0, s);
        break;
      case BULKX_JOB_SORT_BY_KEY:
        st = bulkx_stable_sort_by_key(j.dtype, j.value_dtype, BULKX_LESS, j.out0, j.out1, j.n0,
                                      b->d_ws, b->ws_bytes, s);
        break;
      default: st = BULKX_ERR_INVALID_ARGUMENT;
    }
    if (st != BULKX_SUCCESS) return st;
  }
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_last_timing(bulkx_batch* b, float* phase0_ms, float* passes_ms) {
  if (!b) return BULKX_ERR_INVALID_ARGUMENT;
  if (!b->timed) return BULKX_ERR_NOT_READY;
  BULKX_HIP(hipEventSynchronize(b->ev[2]));
  float a = 0.f, c = 0.f;
  BULKX_HIP(hipEventElapsedTime(&a, b->ev[0], b->ev[1]));
  BULKX_HIP(hipEventElapsedTime(&c, b->ev[1], b->ev[2]));
  if (phase0_ms) *phase0_ms = a;
  if (passes_ms) *passes_ms = c;
  return BULKX_SUCCESS;
}
