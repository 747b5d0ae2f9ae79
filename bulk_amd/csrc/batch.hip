// Batch executor: many uneven scan / reduce / reduce_by_key / merge_by_key /
// sort_by_key jobs run as segmented work, a fixed sequence of launches
// recorded once into a hipGraph.
//
// Work decomposition.  Every job is cut into fixed-size tiles, and the tiles
// of the jobs of one op family (int32 scan/reduce, float32 scan/reduce, the
// three reduce_by_key value types, merge, sort) form that family's tile
// table.  Each family is one launch with one workgroup per tile (the grids
// follow the job list); a workgroup maps its tile to the job through the
// device-resident tile->job table and runs the tile body the device-wide
// kernels run (scan_tile, rbk_tile, merge_two_tile, block_sort_tile).  Scan
// and reduce jobs keep a decoupled look-back chain of their own (per-job
// status words, tiles claimed from a per-job counter, so a tile's
// predecessors always run or are done); reduce_by_key jobs run the
// compact form per job (each tile's runs to scratch slots with its
// aggregate, one scan workgroup per job, then the runs moved to their
// places); merge jobs read their tile boundaries from a partition
// launch.  Sort jobs of one tile are block-sorted; larger ones run the four
// radix passes (upsweep, per-job counts scan, downsweep) over all their
// tiles together.  The graph forks into four concurrent branches (sort,
// merge, reduce_by_key, scan) after one clear kernel.
//
// Each family is compiled in two tile shapes (batch_shapes.hpp); a batch
// takes each family's shape from the tuning table when it is created, and
// the autotuner times both on the mixed-job stream.
//
// Everything a launch needs (tile counts, tables, counters) lives in device
// memory, so the captured graph stays valid when the job list changes: a new
// list is uploaded with stream-ordered copies, the size-dependent grids are
// updated in place, and the graph is replayed.
#include <algorithm>
#include <vector>

#include "batch_shapes.hpp"
#include "host_common.hpp"
#include "merge_impl.hpp"
#include "reduce_by_key_impl.hpp"
#include "scan_impl.hpp"
#include "sort_impl.hpp"
#include "tuning.hpp"

namespace bulkx {

// Tile shapes: fam_block / fam_items (batch_shapes.hpp), two per family.
// Sort jobs of at most one block-sort tile (512 x 16 = 8192 keys, in LDS)
// are sorted by the block sort alone; larger ones by the radix passes
// (radix_impl.hpp: four digit passes, each an upsweep of per-tile digit
// counts, a scan of the job's counts and a downsweep that scatters keys and
// values), which read and write a job 4 times however large it is.
constexpr int64_t kBatchSortRun = int64_t(kBatchSortBlock) * kBatchSortItems;
using BatchRadix = RadixConfig<uint32_t>;
constexpr int64_t kBatchRadixTile = BatchRadix::TILE;
constexpr int64_t kBatchMinTile = batch_min_tile();
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
  int32_t passes;  // sort: radix passes (0: one block-sorted tile)
  int64_t n0, n1;
  const void* in0;
  const void* in1;
  const void* in2;
  const void* in3;
  void* out0;
  void* out1;
  void* out2;
  uint64_t* status;  // look-back words (null for single-tile jobs)
  int64_t* splits;   // merge: the merge path of every tile boundary (tiles0 + 1)
  void* tmp_k;       // sort: ping-pong buffers; reduce_by_key: run slots (TILE + 1 per tile)
  void* tmp_v;
  uint32_t* counts;  // sort: digit counts of the radix tiles (256 per tile, tile-major);
                     // reduce_by_key: per-tile run count | first-element-is-a-head << 31
  uint32_t* totals;  // sort: keys per digit over the job (256), from the row scan
  int64_t tiles0;       // phase-0 tiles
  int64_t tile_begin0;  // first global phase-0 tile
  int64_t sort_tile_begin;  // first global radix tile
  int64_t sort_tiles;       // radix tiles
};

// Phase-0 work is split by op family (BatchFamily, batch_shapes.hpp); each
// family is one kernel compiled for its own register and LDS footprint (a
// single kernel over all ops would run at the occupancy of the heaviest
// body), in each of its tile shapes.
constexpr int kBatchStages = FAM_COUNT + 1;  // the family kernels, then the sort radix passes
// Launch-grid kinds: one per phase-0 family, the radix passes' upsweep, scan
// and downsweep, the aggregate scans of the three reduce_by_key families,
// and the merge partition.
constexpr int kRbkFamilies = 3;
constexpr int kGridRadixUp = FAM_COUNT;
constexpr int kGridRadixScan = FAM_COUNT + 1;
constexpr int kGridRadixDown = FAM_COUNT + 2;
constexpr int grid_kind_rbk_scan(int fam) { return FAM_COUNT + 3 + fam - FAM_RBK_I32; }
constexpr int kGridMergePartition = FAM_COUNT + 3 + kRbkFamilies;
constexpr int kGridKinds = kGridMergePartition + 1;
constexpr int kBatchBranches = 4;

struct BatchDesc {
  int64_t fam_begin[FAM_COUNT + 1];  // family f owns tiles [fam_begin[f], fam_begin[f+1])
  // multi-tile reduce_by_key jobs of family FAM_RBK_I32 + r:
  // rbk_jobs[rbk_begin[r], rbk_begin[r + 1])
  int64_t rbk_begin[kRbkFamilies + 1];
  int64_t total_sort_tiles;  // radix tiles of all sort jobs
  int64_t num_radix_jobs;    // sort jobs that run the radix passes
  int64_t num_status_words;
  int64_t num_jobs;
  LookbackGuard guard;  // the look-back waits of every family; err cleared per launch
};

template <int FAM, int SH>
struct FamilySmem;
template <int SH>
struct FamilySmem<FAM_SCAN_I32, SH> {
  ScanSmem<int32_t, OpPlus, fam_block(FAM_SCAN_I32, SH)> s;
};
template <int SH>
struct FamilySmem<FAM_SCAN_F32, SH> {
  ScanSmem<float, OpPlus, fam_block(FAM_SCAN_F32, SH)> s;
};
template <int SH>
struct FamilySmem<FAM_RBK_I32, SH> {
  RbkSmem<int32_t, int32_t, OpPlus, fam_block(FAM_RBK_I32, SH), fam_items(FAM_RBK_I32, SH)> s;
};
template <int SH>
struct FamilySmem<FAM_RBK_F32, SH> {
  RbkSmem<int32_t, float, OpPlus, fam_block(FAM_RBK_F32, SH), fam_items(FAM_RBK_F32, SH)> s;
};
template <int SH>
struct FamilySmem<FAM_RBK_I64, SH> {
  RbkSmem<int32_t, int64_t, OpPlus, fam_block(FAM_RBK_I64, SH), fam_items(FAM_RBK_I64, SH)> s;
};
template <int SH>
struct FamilySmem<FAM_MERGE, SH> {
  static constexpr int B = fam_block(FAM_MERGE, SH), I = fam_items(FAM_MERGE, SH);
  union {
    MergeSmem<int32_t, uint32_t, true, B, I> i;
    MergeSmem<float, uint32_t, true, B, I> f;
    MergeSmem<uint32_t, uint32_t, true, B, I> u;
  };
};
template <int SH>
struct FamilySmem<FAM_SORT, SH> {
  union {
    SortSmem<int32_t, uint32_t, true, kBatchSortBlock, kBatchSortItems> i;
    SortSmem<uint32_t, uint32_t, true, kBatchSortBlock, kBatchSortItems> u;
    SortSmem<float, uint32_t, true, kBatchSortBlock, kBatchSortItems> f;
  };
};
// The radix downsweep's LDS, one layout for every 4-byte key type.
using BatchRadixSmem = RadixSmem<uint32_t, uint32_t, true, BatchRadix::BLOCK, BatchRadix::ITEMS>;
template <class K>
__device__ __forceinline__ RadixSmem<K, uint32_t, true, BatchRadix::BLOCK, BatchRadix::ITEMS>&
radix_smem(BatchRadixSmem& sm) {
  static_assert(sizeof(K) == 4, "4-byte keys share the layout");
  return reinterpret_cast<RadixSmem<K, uint32_t, true, BatchRadix::BLOCK, BatchRadix::ITEMS>&>(sm);
}

__global__ void batch_clear_kernel(BatchDesc* d, uint64_t* status, uint32_t* job_ctr) {
  const int64_t nw = d->num_status_words;
  const int64_t gs = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nw; i += gs) status[i] = 0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < d->num_jobs; i += gs)
    job_ctr[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) d->guard.err = 0;
}

template <class K, class SM>
__device__ __forceinline__ void sort_block_job(SM& sm, const BatchJobDev& job, int64_t lt) {
  // a job of one tile, sorted in place (larger jobs run the radix passes)
  K* keys = static_cast<K*>(job.out0);
  uint32_t* vals = static_cast<uint32_t*>(job.out1);
  block_sort_tile<CmpLess, K, uint32_t, true, kBatchSortBlock, kBatchSortItems>(
      sm, keys, vals, job.n0, keys, vals, lt);
}

// Value type of a reduce_by_key family, and its tile aggregates (kept in the
// job's status words).
template <int FAM>
using RbkValue = typename std::conditional<
    FAM == FAM_RBK_I32, int32_t, typename std::conditional<FAM == FAM_RBK_F32, float, int64_t>::type>::type;
template <int FAM>
using RbkAgg = Seg<typename Acc<RbkValue<FAM>, OpPlus>::type>;

// One phase-0 tile of a job of family FAM in shape SH.
template <int FAM, int SH>
__device__ __forceinline__ void run_tile(FamilySmem<FAM, SH>& sm, BatchDesc* d,
                                         const BatchJobDev& job, int64_t lt) {
  constexpr int B = fam_block(FAM, SH), I = fam_items(FAM, SH);
  const bool single = job.tiles0 == 1;
  if constexpr (FAM == FAM_SCAN_I32) {
    if (job.op == OP_SCAN_I32)
      scan_tile<int32_t, OpPlus, B, I, false>(
          sm.s, static_cast<const int32_t*>(job.in0), static_cast<int32_t*>(job.out0), job.n0, 0,
          false, lt, job.status, &d->guard, single, nullptr);
    else
      scan_tile<int32_t, OpPlus, B, I, false, false>(
          sm.s, static_cast<const int32_t*>(job.in0), nullptr, job.n0, 0, false, lt, job.status,
          &d->guard, single, static_cast<int32_t*>(job.out0));
  } else if constexpr (FAM == FAM_SCAN_F32) {
    if (job.op == OP_SCAN_F32)
      scan_tile<float, OpPlus, B, I, false>(
          sm.s, static_cast<const float*>(job.in0), static_cast<float*>(job.out0), job.n0, 0.f,
          false, lt, job.status, &d->guard, single, nullptr);
    else
      scan_tile<float, OpPlus, B, I, false, false>(
          sm.s, static_cast<const float*>(job.in0), nullptr, job.n0, 0.f, false, lt, job.status,
          &d->guard, single, static_cast<float*>(job.out0));
  } else if constexpr (FAM == FAM_MERGE) {
    if (job.op == OP_MERGE_I32)
      merge_two_tile<CmpLess, int32_t, uint32_t, true, B, I>(
          sm.i, static_cast<const int32_t*>(job.in0), static_cast<const uint32_t*>(job.in1),
          job.n0, static_cast<const int32_t*>(job.in2), static_cast<const uint32_t*>(job.in3),
          job.n1, static_cast<int32_t*>(job.out0), static_cast<uint32_t*>(job.out1), lt,
          job.splits);
    else
      merge_two_tile<CmpLess, float, uint32_t, true, B, I>(
          sm.f, static_cast<const float*>(job.in0), static_cast<const uint32_t*>(job.in1),
          job.n0, static_cast<const float*>(job.in2), static_cast<const uint32_t*>(job.in3),
          job.n1, static_cast<float*>(job.out0), static_cast<uint32_t*>(job.out1), lt,
          job.splits);
  } else {
    switch (job.op) {
      case OP_SORT_I32: sort_block_job<int32_t>(sm.i, job, lt); break;
      case OP_SORT_U32: sort_block_job<uint32_t>(sm.u, job, lt); break;
      default: sort_block_job<float>(sm.f, job, lt); break;
    }
  }
}

// Phase-0 kernel of one family: one workgroup per tile, launched with one
// workgroup per tile of the family (the host resizes the grid, also inside
// the captured graph, whenever the job list changes).  Workgroup b serves a
// tile of job tile_job[begin + b].  Families with a look-back chain (scan,
// reduce, reduce_by_key) take the tile's index within the job from the
// job's own atomic counter, so a tile's predecessors always belong to
// workgroups that already run or are done; the counters of different jobs
// are different addresses, so claims do not serialise across jobs (one
// device-wide counter would cost ~11 ns per tile on MI355X).  Merge and sort
// tiles are independent and use the workgroup index directly.  (A
// persistent loop here would keep the tile body's per-lane address math live
// across iterations and cost ~40 registers per lane.)
template <int FAM, int SH>
__global__ __launch_bounds__(fam_block(FAM, SH)) void batch_phase0_kernel(
    BatchDesc* d, const BatchJobDev* __restrict__ jobs, const uint32_t* __restrict__ tile_job,
    uint32_t* __restrict__ job_ctr) {
  static_assert(!fam_is_rbk(FAM), "reduce_by_key families run batch_rbk_compact_kernel");
  __shared__ FamilySmem<FAM, SH> sm;
  const int64_t begin = d->fam_begin[FAM];
  const int64_t count = d->fam_begin[FAM + 1] - begin;
  const int64_t g = int64_t(blockIdx.x);
  if (g >= count) return;
  const uint32_t j = tile_job[begin + g];
  const BatchJobDev& job = jobs[j];
  int64_t lt;
  if constexpr (FAM == FAM_MERGE || FAM == FAM_SORT) {
    lt = begin + g - job.tile_begin0;
  } else {
    __shared__ uint32_t s_t;
    if (threadIdx.x == 0) s_t = atomicAdd(&job_ctr[j], 1u);
    __syncthreads();
    lt = wave_uniform(s_t);
  }
  run_tile<FAM, SH>(sm, d, job, lt);
}

// The merge path of every merge tile's first diagonal, one lane per tile
// (as merge_split_kernel); the last tile of a job also writes the end
// boundary (all of A).  Searched once here instead of by each tile.
template <int SH>
__global__ __launch_bounds__(256) void batch_merge_partition_kernel(
    BatchDesc* d, const BatchJobDev* __restrict__ jobs, const uint32_t* __restrict__ tile_job) {
  const int64_t begin = d->fam_begin[FAM_MERGE];
  const int64_t g = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (g >= d->fam_begin[FAM_MERGE + 1] - begin) return;
  const BatchJobDev& job = jobs[tile_job[begin + g]];
  const int64_t lt = begin + g - job.tile_begin0;
  const int64_t n = job.n0 + job.n1;
  constexpr int64_t T = fam_tile(FAM_MERGE, SH);
  const int64_t diag = lt * T < n ? lt * T : n;
  int64_t r;
  if (job.op == OP_MERGE_I32)
    r = lane_merge_path<CmpLess>(static_cast<const int32_t*>(job.in0), job.n0,
                                 static_cast<const int32_t*>(job.in2), job.n1, diag);
  else
    r = lane_merge_path<CmpLess>(static_cast<const float*>(job.in0), job.n0,
                                 static_cast<const float*>(job.in2), job.n1, diag);
  job.splits[lt] = r;
  if (lt == job.tiles0 - 1) job.splits[lt + 1] = job.n0;
}

// Compact reduce_by_key in the batch (the device-wide compact form,
// rbk_compact_kernel): pass 1 reads every tile of the family once -- a job
// of one tile writes its runs and run count directly, a tile of a
// multi-tile job writes the runs that end in it to its scratch slots
// (tmp_k / tmp_v, TILE + 1 per tile), its aggregate to the job's status
// words and its run count and head flag to the job's meta (counts); pass 2
// scans each multi-tile job's aggregates in place (one workgroup per job);
// pass 3 moves every tile's runs to their places, the value carried into
// the tile folded into its first run.  No tile ever waits on another.
// Bytes per element with 4-byte keys and values and runs of 8: 8 read + 3
// for the runs (written, read, written), against 13 for a three-pass form
// that reads the keys twice (aggregates, then the tiles).  Mixed stream of
// 4096 jobs, int32 and float32 families: 1.41 ms against 1.54 ms for the
// three-pass form (pass 1 578 us, pass 3 144 us for the int32 family).
template <int FAM, int SH>
__global__ __launch_bounds__(fam_block(FAM, SH)) void batch_rbk_compact_kernel(
    BatchDesc* d, const BatchJobDev* __restrict__ jobs, const uint32_t* __restrict__ tile_job) {
  __shared__ FamilySmem<FAM, SH> sm;
  const int64_t begin = d->fam_begin[FAM];
  const int64_t g = int64_t(blockIdx.x);
  if (g >= d->fam_begin[FAM + 1] - begin) return;
  const BatchJobDev& job = jobs[tile_job[begin + g]];
  using V = RbkValue<FAM>;
  constexpr int B = fam_block(FAM, SH), I = fam_items(FAM, SH);
  const int32_t* keys = static_cast<const int32_t*>(job.in0);
  const V* vals = static_cast<const V*>(job.in1);
  if (job.tiles0 <= 1) {
    rbk_tile<int32_t, V, OpPlus, B, I>(sm.s, keys, vals, job.n0, static_cast<int32_t*>(job.out0),
                                       static_cast<V*>(job.out1), static_cast<int64_t*>(job.out2),
                                       0, nullptr, &d->guard, true);
    return;
  }
  const RbkCompactOut<RbkAgg<FAM>> co{reinterpret_cast<RbkAgg<FAM>*>(job.status), job.counts};
  rbk_tile<int32_t, V, OpPlus, B, I>(sm.s, keys, vals, job.n0, static_cast<int32_t*>(job.tmp_k),
                                     static_cast<V*>(job.tmp_v), nullptr,
                                     begin + g - job.tile_begin0, nullptr, &d->guard, false,
                                     nullptr, nullptr, nullptr, 0, &co);
}

template <int FAM, int SH>
__global__ __launch_bounds__(256) void batch_rbk_move_kernel(
    BatchDesc* d, const BatchJobDev* __restrict__ jobs, const uint32_t* __restrict__ tile_job) {
  const int64_t begin = d->fam_begin[FAM];
  const int64_t g = int64_t(blockIdx.x);
  if (g >= d->fam_begin[FAM + 1] - begin) return;
  const BatchJobDev& job = jobs[tile_job[begin + g]];
  if (job.tiles0 <= 1) return;
  using V = RbkValue<FAM>;
  rbk_move_tile<int32_t, V, OpPlus, int(fam_tile(FAM, SH))>(
      static_cast<const int32_t*>(job.tmp_k), static_cast<const V*>(job.tmp_v),
      reinterpret_cast<const RbkAgg<FAM>*>(job.status), nullptr, job.counts, job.tiles0, job.n0,
      static_cast<int32_t*>(job.out0), static_cast<V*>(job.out1), static_cast<int64_t*>(job.out2),
      begin + g - job.tile_begin0);
}

template <int FAM>
__global__ __launch_bounds__(256) void batch_rbk_scan_kernel(BatchDesc* d,
                                                             const BatchJobDev* __restrict__ jobs,
                                                             const uint32_t* __restrict__ rbk_jobs) {
  const int r = FAM - FAM_RBK_I32;
  const int64_t b = d->rbk_begin[r] + int64_t(blockIdx.x);
  if (b >= d->rbk_begin[r + 1]) return;
  const BatchJobDev& job = jobs[rbk_jobs[b]];
  (void)rbk_scan_range<RbkValue<FAM>, OpPlus>(reinterpret_cast<RbkAgg<FAM>*>(job.status), 0,
                                              job.tiles0);
}

// Source and destination of a sort job's radix pass `pass`: even passes
// read the caller's arrays and write the temporaries, odd passes the
// reverse, so the fourth pass ends in place.
template <class K>
struct RadixBuffers {
  const K* sk;
  const uint32_t* sv;
  K* dk;
  uint32_t* dv;
  __device__ RadixBuffers(const BatchJobDev& job, int pass) {
    const bool src_is_tmp = (pass & 1) != 0;
    K* ok = static_cast<K*>(job.out0);
    uint32_t* ov = static_cast<uint32_t*>(job.out1);
    K* tk = static_cast<K*>(job.tmp_k);
    uint32_t* tv = static_cast<uint32_t*>(job.tmp_v);
    sk = src_is_tmp ? tk : ok;
    sv = src_is_tmp ? tv : ov;
    dk = src_is_tmp ? ok : tk;
    dv = src_is_tmp ? ov : tv;
  }
};

// Digit counts of every radix tile of pass `pass` (one workgroup per tile),
// each tile's 256 counts one coalesced row (tile-major).
__global__ __launch_bounds__(BatchRadix::BLOCK) void batch_radix_upsweep_kernel(
    BatchDesc* d, const BatchJobDev* __restrict__ jobs, const uint32_t* __restrict__ sort_tile_job,
    int pass) {
  __shared__ uint32_t s_hist[kRadixCountWords];
  const int64_t t = blockIdx.x;
  if (t >= d->total_sort_tiles) return;
  const BatchJobDev& job = jobs[sort_tile_job[t]];
  const int64_t lt = t - job.sort_tile_begin;
  const int shift = pass * kRadixBits;
  switch (job.op) {
    case OP_SORT_I32:
      radix_count_tile<int32_t, false, BatchRadix::BLOCK, BatchRadix::ITEMS, true>(
          s_hist, RadixBuffers<int32_t>(job, pass).sk, job.n0, shift, job.counts, job.sort_tiles, lt);
      break;
    case OP_SORT_U32:
      radix_count_tile<uint32_t, false, BatchRadix::BLOCK, BatchRadix::ITEMS, true>(
          s_hist, RadixBuffers<uint32_t>(job, pass).sk, job.n0, shift, job.counts, job.sort_tiles, lt);
      break;
    default:
      radix_count_tile<float, false, BatchRadix::BLOCK, BatchRadix::ITEMS, true>(
          s_hist, RadixBuffers<float>(job, pass).sk, job.n0, shift, job.counts, job.sort_tiles, lt);
      break;
  }
}

// Row scan of each radix job's counts: sixteen workgroups per job, one per
// group of 16 digits (radix_scan_rows), so a large job's scan is spread over
// the device; each count becomes the keys of its digit in the job's earlier
// tiles, and the job's per-digit totals are written beside them.
constexpr int kRadixRowGroups = kRadixDigits / 16;
constexpr int kRadixRowSlices = 16;
constexpr int kRadixRowBlock = 16 * kRadixRowSlices;
__global__ __launch_bounds__(kRadixRowBlock) void batch_radix_scan_kernel(
    BatchDesc* d, const BatchJobDev* __restrict__ jobs, const uint32_t* __restrict__ radix_jobs) {
  __shared__ uint32_t s_part[kRadixRowSlices][16];
  const int64_t b = blockIdx.x / kRadixRowGroups;
  if (b >= d->num_radix_jobs) return;
  const BatchJobDev& job = jobs[radix_jobs[b]];
  radix_scan_rows<kRadixRowSlices>(job.counts, job.sort_tiles, job.totals,
                                   int(blockIdx.x % kRadixRowGroups), s_part);
}

template <class K>
__device__ __forceinline__ void radix_down_job(BatchRadixSmem& sm, const BatchJobDev& job,
                                               int pass, int64_t lt) {
  const RadixBuffers<K> rb(job, pass);
  radix_scatter_tile<K, uint32_t, true, false, BatchRadix::BLOCK, BatchRadix::ITEMS>(
      radix_smem<K>(sm), rb.sk, rb.sv, job.n0, pass * kRadixBits, job.counts, job.sort_tiles,
      rb.dk, rb.dv, lt, job.totals);
}

// Scatter of pass `pass` for every radix tile; each XCD takes a contiguous
// range of the tiles (xcd_tile), so a job's neighbouring tiles share L2.
__global__ __launch_bounds__(BatchRadix::BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void batch_radix_downsweep_kernel(
    BatchDesc* d, const BatchJobDev* __restrict__ jobs, const uint32_t* __restrict__ sort_tile_job,
    int pass) {
  __shared__ BatchRadixSmem sm;
  const int64_t total = d->total_sort_tiles;
  if (int64_t(blockIdx.x) >= total) return;
  const int64_t t = xcd_tile(blockIdx.x, total);
  const BatchJobDev& job = jobs[sort_tile_job[t]];
  const int64_t lt = t - job.sort_tile_begin;
  switch (job.op) {
    case OP_SORT_I32: radix_down_job<int32_t>(sm, job, pass, lt); break;
    case OP_SORT_U32: radix_down_job<uint32_t>(sm, job, pass, lt); break;
    default: radix_down_job<float>(sm, job, pass, lt); break;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
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

static int host_family(int op) {
  switch (op) {
    case OP_SCAN_I32: case OP_REDUCE_I32: return FAM_SCAN_I32;
    case OP_SCAN_F32: case OP_REDUCE_F32: return FAM_SCAN_F32;
    case OP_RBK_I32_I32: return FAM_RBK_I32;
    case OP_RBK_I32_F32: return FAM_RBK_F32;
    case OP_RBK_I32_I64: return FAM_RBK_I64;
    case OP_MERGE_I32: case OP_MERGE_F32: return FAM_MERGE;
    default: return FAM_SORT;
  }
}

}  // namespace bulkx

// Whether the batch can run job j (its kind, dtypes and pointers); the same
// check bulkx_batch_set_jobs applies to every job of a list.
int bulkx::batch_check_job(const bulkx_job& j) {
  const int op = batch_op(j);
  if (op == OP_NONE) return BULKX_ERR_UNSUPPORTED_DTYPE;
  if (j.n0 < 0 || j.n1 < 0) return BULKX_ERR_INVALID_ARGUMENT;
  const int64_t n = j.n0 + (op == OP_MERGE_I32 || op == OP_MERGE_F32 ? j.n1 : 0);
  if (n == 0 && j.kind != BULKX_JOB_REDUCE && j.kind != BULKX_JOB_REDUCE_BY_KEY)
    return BULKX_SUCCESS;
  const bool rbk = j.kind == BULKX_JOB_REDUCE_BY_KEY;
  if (!j.out0 || (rbk && (!j.out1 || !j.out2)) ||
      (n > 0 && !j.in0 && j.kind != BULKX_JOB_SORT_BY_KEY))
    return BULKX_ERR_INVALID_ARGUMENT;
  if (rbk && n > 0 && !j.in1) return BULKX_ERR_INVALID_ARGUMENT;
  if (j.kind == BULKX_JOB_MERGE_BY_KEY &&
      ((j.n0 > 0 && !j.in1) || (j.n1 > 0 && (!j.in2 || !j.in3)) || !j.out1))
    return BULKX_ERR_INVALID_ARGUMENT;
  if (j.kind == BULKX_JOB_SORT_BY_KEY && !j.out1) return BULKX_ERR_INVALID_ARGUMENT;
  return BULKX_SUCCESS;
}

namespace bulkx {

// Workspace one plain entry point needs for job j.
static size_t per_call_workspace(const bulkx_job& j) {
  switch (j.kind) {
    case BULKX_JOB_INCLUSIVE_SCAN: return scan_workspace_bytes(j.dtype, j.n0);
    case BULKX_JOB_REDUCE: return reduce_workspace_bytes(j.dtype, j.n0);
    case BULKX_JOB_REDUCE_BY_KEY: return reduce_by_key_workspace_bytes(j.dtype, j.value_dtype, j.n0);
    case BULKX_JOB_SORT_BY_KEY: return sort_workspace_bytes(j.dtype, j.value_dtype, j.n0);
    case BULKX_JOB_MERGE_BY_KEY: return merge_workspace_bytes(j.dtype, j.value_dtype, j.n0, j.n1);
    default: return 0;
  }
}

}  // namespace bulkx

using namespace bulkx;

// A recorded launch sequence and its kernel nodes whose grid follows the job
// list.
struct BatchGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<std::pair<hipGraphNode_t, int>> sized_nodes;  // (node, grid kind)
  unsigned graph_grid[kGridKinds] = {};
};

struct bulkx_batch {
  hipStream_t stream = nullptr;
  hipStream_t capture_stream = nullptr;  // graphs are recorded here (capture is not
                                         // permitted on the null stream)
  // the graph's concurrent branches (sort chain, merge, reduce_by_key, scan)
  hipStream_t side[kBatchBranches] = {};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev[kBatchBranches] = {};
  int max_jobs = 0;
  int64_t max_elements = 0;
  int64_t max_tiles = 0;
  int64_t fam_tiles[FAM_COUNT] = {};  // phase-0 grid of each family = its tile count
  int shape[FAM_COUNT] = {};  // each family's tile shape (tuning table, at create)
  // device
  BatchDesc* d_desc = nullptr;
  BatchJobDev* d_jobs = nullptr;
  uint32_t* d_tile_job = nullptr;
  uint32_t* d_sort_tile_job = nullptr;
  uint32_t* d_rbk_jobs = nullptr;  // multi-tile reduce_by_key jobs, by family
  uint32_t* d_radix_jobs = nullptr;  // sort jobs that run the radix passes
  int64_t* d_merge_splits = nullptr;  // tile boundaries of the merge jobs
  uint32_t* d_job_ctr = nullptr;  // per-job tile counters of the look-back families
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
  uint32_t* h_rbk_jobs = nullptr;
  uint32_t* h_radix_jobs = nullptr;
  uint32_t* h_err = nullptr;  // the last launch's error word, copied back behind it
  bool last_per_call = false;  // the last launch ran the plain entry points
  int64_t rbk_jobs[kRbkFamilies] = {};  // multi-tile jobs of each reduce_by_key family
  // current job list
  std::vector<bulkx_job> jobs;
  int num_jobs = 0;
  int64_t total_tiles0 = 0, total_sort_tiles = 0, num_radix_jobs = 0;
  // stage timing of eager launches: ev[f]..ev[f+1] brackets family f,
  // ev[FAM_COUNT]..ev[FAM_COUNT+1] the sort radix passes; in pass p,
  // rev[3p]..rev[3p+1] brackets the upsweep, rev[3p+1]..rev[3p+2] the
  // counts scan and rev[3p+2]..rev[3p+3] the downsweep
  hipEvent_t ev[kBatchStages + 1] = {};
  hipEvent_t rev[3 * kRadixPasses + 1] = {};
  bool timed = false;
  // the recorded launch sequence (one graph for every job list)
  BatchGraph graph;
};

// Grid of each job-list-sized launch: the phase-0 families (one workgroup per
// tile), the radix passes (one workgroup per radix tile; the counts scan one
// per radix job), the reduce_by_key aggregate scans (one per multi-tile job)
// and the merge partition (one lane per merge tile).  Never 0: an idle
// launch runs one workgroup that finds no work.
static unsigned batch_grid(const bulkx_batch* b, int kind) {
  int64_t g;
  if (kind < FAM_COUNT) {
    g = b->fam_tiles[kind];
  } else if (kind == kGridMergePartition) {
    g = (b->fam_tiles[FAM_MERGE] + 255) / 256;
  } else if (kind == kGridRadixUp || kind == kGridRadixDown) {
    g = b->total_sort_tiles;
  } else if (kind == kGridRadixScan) {
    g = b->num_radix_jobs * kRadixRowGroups;
  } else {
    g = b->rbk_jobs[kind - grid_kind_rbk_scan(FAM_RBK_I32)];
  }
  return unsigned(g > 0 ? g : 1);
}

using Phase0Kernel = void (*)(BatchDesc*, const BatchJobDev*, const uint32_t*, uint32_t*);
// (reduce_by_key families run their own three kernels: no phase-0 entry)
#define BULKX_PHASE0(SH)                                                              \
  {batch_phase0_kernel<FAM_SCAN_I32, SH>, batch_phase0_kernel<FAM_SCAN_F32, SH>,      \
   nullptr, nullptr, nullptr, batch_phase0_kernel<FAM_MERGE, SH>,                     \
   batch_phase0_kernel<FAM_SORT, 0>}  /* one shape */
static const Phase0Kernel kPhase0[kBatchShapes][FAM_COUNT] = {BULKX_PHASE0(0), BULKX_PHASE0(1)};
#undef BULKX_PHASE0

using RbkUpKernel = void (*)(BatchDesc*, const BatchJobDev*, const uint32_t*);
static const RbkUpKernel kBatchRbkCompact[kBatchShapes][kRbkFamilies] = {
    {batch_rbk_compact_kernel<FAM_RBK_I32, 0>, batch_rbk_compact_kernel<FAM_RBK_F32, 0>,
     batch_rbk_compact_kernel<FAM_RBK_I64, 0>},
    {batch_rbk_compact_kernel<FAM_RBK_I32, 1>, batch_rbk_compact_kernel<FAM_RBK_F32, 1>,
     batch_rbk_compact_kernel<FAM_RBK_I64, 1>}};
static const RbkUpKernel kBatchRbkMove[kBatchShapes][kRbkFamilies] = {
    {batch_rbk_move_kernel<FAM_RBK_I32, 0>, batch_rbk_move_kernel<FAM_RBK_F32, 0>,
     batch_rbk_move_kernel<FAM_RBK_I64, 0>},
    {batch_rbk_move_kernel<FAM_RBK_I32, 1>, batch_rbk_move_kernel<FAM_RBK_F32, 1>,
     batch_rbk_move_kernel<FAM_RBK_I64, 1>}};
using PartitionKernel = void (*)(BatchDesc*, const BatchJobDev*, const uint32_t*);
static const PartitionKernel kMergePartition[kBatchShapes] = {batch_merge_partition_kernel<0>,
                                                              batch_merge_partition_kernel<1>};
static const RbkUpKernel kRbkScan[kRbkFamilies] = {batch_rbk_scan_kernel<FAM_RBK_I32>,
                                                   batch_rbk_scan_kernel<FAM_RBK_F32>,
                                                   batch_rbk_scan_kernel<FAM_RBK_I64>};


// The passes of family f on stream q (reduce_by_key: compact, scan, move).
static int batch_enqueue_family(bulkx_batch* b, int f, hipStream_t q) {
  const int sh = b->shape[f];
  if (f == FAM_MERGE) {
    kMergePartition[sh]<<<batch_grid(b, kGridMergePartition), 256, 0, q>>>(b->d_desc, b->d_jobs,
                                                                          b->d_tile_job);
    BULKX_LAUNCHED();
  }
  if (fam_is_rbk(f)) {
    const int r = f - FAM_RBK_I32;
    kBatchRbkCompact[sh][r]<<<batch_grid(b, f), fam_block(f, sh), 0, q>>>(b->d_desc, b->d_jobs,
                                                                     b->d_tile_job);
    BULKX_LAUNCHED();
    kRbkScan[r]<<<batch_grid(b, grid_kind_rbk_scan(f)), 256, 0, q>>>(b->d_desc, b->d_jobs,
                                                                     b->d_rbk_jobs);
    BULKX_LAUNCHED();
    kBatchRbkMove[sh][r]<<<batch_grid(b, f), 256, 0, q>>>(b->d_desc, b->d_jobs, b->d_tile_job);
    BULKX_LAUNCHED();
    return BULKX_SUCCESS;
  }
  kPhase0[sh][f]<<<batch_grid(b, f), fam_block(f, sh), 0, q>>>(b->d_desc, b->d_jobs,
                                                               b->d_tile_job, b->d_job_ctr);
  BULKX_LAUNCHED();
  return BULKX_SUCCESS;
}

// The radix passes of every sort job that has more than one tile.  With
// `rev` the pass kernels are bracketed by events (eager launches only).
static int batch_enqueue_radix(bulkx_batch* b, hipStream_t s, const hipEvent_t* rev) {
  for (int p = 0; p < kRadixPasses; ++p) {
    if (rev) BULKX_HIP(hipEventRecord(rev[3 * p], s));
    batch_radix_upsweep_kernel<<<batch_grid(b, kGridRadixUp), BatchRadix::BLOCK, 0, s>>>(
        b->d_desc, b->d_jobs, b->d_sort_tile_job, p);
    BULKX_LAUNCHED();
    if (rev) BULKX_HIP(hipEventRecord(rev[3 * p + 1], s));
    batch_radix_scan_kernel<<<batch_grid(b, kGridRadixScan), kRadixRowBlock, 0, s>>>(
        b->d_desc, b->d_jobs, b->d_radix_jobs);
    BULKX_LAUNCHED();
    if (rev) BULKX_HIP(hipEventRecord(rev[3 * p + 2], s));
    batch_radix_downsweep_kernel<<<batch_grid(b, kGridRadixDown), BatchRadix::BLOCK, 0, s>>>(
        b->d_desc, b->d_jobs, b->d_sort_tile_job, p);
    BULKX_LAUNCHED();
  }
  if (rev) BULKX_HIP(hipEventRecord(rev[3 * kRadixPasses], s));
  return BULKX_SUCCESS;
}

static int batch_enqueue_kernels(bulkx_batch* b, hipStream_t s, bool timed) {
  batch_clear_kernel<<<256, 256, 0, s>>>(b->d_desc, b->d_status, b->d_job_ctr);
  BULKX_LAUNCHED();
  for (int f = 0; f < FAM_COUNT; ++f) {
    if (timed) BULKX_HIP(hipEventRecord(b->ev[f], s));
    BULKX_TRY(batch_enqueue_family(b, f, s));
  }
  if (timed) BULKX_HIP(hipEventRecord(b->ev[FAM_COUNT], s));
  BULKX_TRY(batch_enqueue_radix(b, s, timed ? b->rev : nullptr));
  if (timed) BULKX_HIP(hipEventRecord(b->ev[FAM_COUNT + 1], s));
  return BULKX_SUCCESS;
}

// The graph form: after the clear kernel the work forks into independent
// branches -- the sort chain (block sort of one-tile jobs, then the radix
// passes of the others), merge,
// reduce_by_key and scan -- which the device runs concurrently, so the
// compute-bound sort block and the look-back stalls of reduce_by_key overlap
// the memory-bound families.  Branches share only read-only tables and the
// error flag.
static int batch_enqueue_graph(bulkx_batch* b, hipStream_t s) {
  batch_clear_kernel<<<256, 256, 0, s>>>(b->d_desc, b->d_status, b->d_job_ctr);
  BULKX_LAUNCHED();
  BULKX_HIP(hipEventRecord(b->fork_ev, s));
  for (int br = 0; br < kBatchBranches; ++br) BULKX_HIP(hipStreamWaitEvent(b->side[br], b->fork_ev, 0));
  BULKX_TRY(batch_enqueue_family(b, FAM_SORT, b->side[0]));
  BULKX_TRY(batch_enqueue_radix(b, b->side[0], nullptr));
  BULKX_TRY(batch_enqueue_family(b, FAM_MERGE, b->side[1]));
  for (int f : {int(FAM_RBK_I32), int(FAM_RBK_F32), int(FAM_RBK_I64)})
    BULKX_TRY(batch_enqueue_family(b, f, b->side[2]));
  for (int f : {int(FAM_SCAN_I32), int(FAM_SCAN_F32)})
    BULKX_TRY(batch_enqueue_family(b, f, b->side[3]));
  for (int br = 0; br < kBatchBranches; ++br) {
    BULKX_HIP(hipEventRecord(b->join_ev[br], b->side[br]));
    BULKX_HIP(hipStreamWaitEvent(s, b->join_ev[br], 0));
  }
  return BULKX_SUCCESS;
}

static void batch_drop_graph(bulkx_batch* b);

extern "C" int bulkx_batch_create(int max_jobs, int64_t max_elements, void* stream,
                                  bulkx_batch** out) {
  if (!out || max_jobs <= 0 || max_elements <= 0) return BULKX_ERR_INVALID_ARGUMENT;
  DeviceProps props;
  BULKX_TRY(current_device_props(&props));
  bulkx_batch* b = new bulkx_batch();
  b->stream = static_cast<hipStream_t>(stream);
  b->max_jobs = max_jobs;
  b->max_elements = max_elements;
  for (int f = 0; f < FAM_COUNT; ++f) {
    const int algo = fam_tuning_algo(f);
    const int v = algo < 0 ? 0 : tuned_variant(algo, fam_tuning_dtype(f), 0);
    b->shape[f] = v >= 0 && v < kBatchShapes ? v : 0;
  }
  // every job can add one partial tile
  b->max_tiles = max_elements / kBatchMinTile + 2 * int64_t(max_jobs) + 1;
  // radix positions are 32-bit per job; a batch never holds a larger job
  if (max_elements > kRadixMaxN) {
    delete b;
    return BULKX_ERR_INVALID_ARGUMENT;
  }
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
  dmalloc(reinterpret_cast<void**>(&b->d_rbk_jobs), sizeof(uint32_t) * max_jobs);
  dmalloc(reinterpret_cast<void**>(&b->d_radix_jobs), sizeof(uint32_t) * max_jobs);
  dmalloc(reinterpret_cast<void**>(&b->d_merge_splits),
          sizeof(int64_t) * (b->max_tiles + int64_t(max_jobs)));
  dmalloc(reinterpret_cast<void**>(&b->d_job_ctr), sizeof(uint32_t) * max_jobs);
  dmalloc(reinterpret_cast<void**>(&b->d_status), sizeof(uint64_t) * b->status_words_cap);
  dmalloc(reinterpret_cast<void**>(&b->d_ws), b->ws_bytes);
  hmalloc(reinterpret_cast<void**>(&b->h_desc), sizeof(BatchDesc));
  hmalloc(reinterpret_cast<void**>(&b->h_jobs), sizeof(BatchJobDev) * max_jobs);
  hmalloc(reinterpret_cast<void**>(&b->h_tile_job), sizeof(uint32_t) * b->max_tiles);
  hmalloc(reinterpret_cast<void**>(&b->h_sort_tile_job), sizeof(uint32_t) * b->max_tiles);
  hmalloc(reinterpret_cast<void**>(&b->h_rbk_jobs), sizeof(uint32_t) * max_jobs);
  hmalloc(reinterpret_cast<void**>(&b->h_radix_jobs), sizeof(uint32_t) * max_jobs);
  hmalloc(reinterpret_cast<void**>(&b->h_err), sizeof(uint32_t));
  if (st != BULKX_SUCCESS) {
    bulkx_batch_destroy(b);
    return st;
  }
  memset(b->h_desc, 0, sizeof(BatchDesc));
  b->h_desc->guard.spin_limit = host_spin_limit();
  *b->h_err = 0;
  hipError_t r = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
  for (int i = 0; i <= kBatchStages && r == hipSuccess; ++i) r = hipEventCreate(&b->ev[i]);
  for (int i = 0; i <= 3 * kRadixPasses && r == hipSuccess; ++i) r = hipEventCreate(&b->rev[i]);
  for (int i = 0; i < kBatchBranches && r == hipSuccess; ++i) {
    r = hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking);
    if (r == hipSuccess) r = hipEventCreateWithFlags(&b->join_ev[i], hipEventDisableTiming);
  }
  if (r == hipSuccess) r = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming);
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
  batch_drop_graph(b);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  for (int i = 0; i < kBatchBranches; ++i) {
    if (b->side[i]) (void)hipStreamDestroy(b->side[i]);
    if (b->join_ev[i]) (void)hipEventDestroy(b->join_ev[i]);
  }
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->rev)
    if (e) (void)hipEventDestroy(e);
  for (void* p : {static_cast<void*>(b->d_desc), static_cast<void*>(b->d_jobs),
                  static_cast<void*>(b->d_tile_job), static_cast<void*>(b->d_sort_tile_job),
                  static_cast<void*>(b->d_radix_jobs), static_cast<void*>(b->d_job_ctr),
                  static_cast<void*>(b->d_status), static_cast<void*>(b->d_tmp),
                  static_cast<void*>(b->d_ws), static_cast<void*>(b->d_rbk_jobs),
                  static_cast<void*>(b->d_merge_splits)})
    if (p) (void)hipFree(p);
  for (void* p : {static_cast<void*>(b->h_desc), static_cast<void*>(b->h_jobs),
                  static_cast<void*>(b->h_tile_job), static_cast<void*>(b->h_sort_tile_job),
                  static_cast<void*>(b->h_rbk_jobs), static_cast<void*>(b->h_radix_jobs),
                  static_cast<void*>(b->h_err)})
    if (p) (void)hipHostFree(p);
  delete b;
  return BULKX_SUCCESS;
}

// Scratch of one radix sort job of n elements: ping-pong keys and values,
// the digit counts of its tiles and its per-digit totals.
static size_t radix_tmp_bytes(int64_t n) {
  const int64_t rt = (n + kBatchRadixTile - 1) / kBatchRadixTile;
  return 2 * align_up(size_t(n) * 4, 256) + align_up(size_t(rt) * kRadixDigits * sizeof(uint32_t), 256) +
         align_up(kRadixDigits * sizeof(uint32_t), 256);
}

// Scratch of one reduce_by_key job of t > 1 tiles of T elements with
// vs-byte values (compact form): the run slots of keys and values, TILE + 1
// per tile, and one meta word per tile.
static size_t rbk_tmp_bytes(int64_t t, int64_t T, size_t vs) {
  const size_t slots = size_t(t) * size_t(T + 1);
  return align_up(slots * 4, 256) + align_up(slots * vs, 256) + align_up(size_t(t) * 4, 256);
}

// Grows the sort ping-pong space and the per-call workspace to at least the
// given sizes (never inside a launch: the caller has synchronised the stream).
static int batch_grow(bulkx_batch* b, size_t tmp_bytes, size_t ws_need) {
  if (tmp_bytes > b->tmp_bytes_cap) {
    if (b->d_tmp) BULKX_HIP(hipFree(b->d_tmp));
    b->d_tmp = nullptr;
    b->tmp_bytes_cap = 0;
    const size_t want = tmp_bytes + tmp_bytes / 4;
    BULKX_HIP(hipMalloc(reinterpret_cast<void**>(&b->d_tmp), want));
    b->tmp_bytes_cap = want;
  }
  if (ws_need > b->ws_bytes) {
    if (b->d_ws) BULKX_HIP(hipFree(b->d_ws));
    b->d_ws = nullptr;
    b->ws_bytes = 0;
    BULKX_HIP(hipMalloc(reinterpret_cast<void**>(&b->d_ws), ws_need));
    b->ws_bytes = ws_need;
  }
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_reserve(bulkx_batch* b, int64_t elements) {
  if (!b || elements < 0 || elements > b->max_elements) return BULKX_ERR_INVALID_ARGUMENT;
  // Scratch of any job list of `elements` elements in all.  Sort jobs: 8
  // bytes per element (ping-pong keys and values), 1024 bytes of digit
  // counts per radix tile, and per job a partial tile's counts, the totals
  // and four alignments (< 4 KB).  reduce_by_key jobs: at most 12 bytes
  // per slot (4-byte keys, values of up to 8), TILE + 1 slots per tile
  // (tiles of at least batch_min_tile() elements, so at most
  // elements / batch_min_tile() + 1 tiles per job, and slots fewer than
  // elements + elements / batch_min_tile() + jobs * (8192 + 1)), a meta word
  // per tile and three alignments per job.  Any mix stays under the larger
  // per-element rate plus both per-job terms.
  const size_t per_job = 4096 + size_t(kRbkCompactMaxTile + 1) * 12 + 3 * 256 + 4;
  const size_t tmp = size_t(elements) * 12 + size_t(elements / batch_min_tile()) * (12 + 4) +
                     size_t(elements / kBatchRadixTile) * kRadixDigits * sizeof(uint32_t) +
                     size_t(b->max_jobs) * per_job;
  size_t ws = 256;
  const int64_t n = elements;
  for (int dt : {int(BULKX_INT32), int(BULKX_FLOAT32)}) {
    ws = std::max(ws, scan_workspace_bytes(dt, n));
    ws = std::max(ws, reduce_workspace_bytes(dt, n));
    ws = std::max(ws, sort_workspace_bytes(dt, BULKX_UINT32, n));
    ws = std::max(ws, merge_workspace_bytes(dt, BULKX_UINT32, n, 0));
  }
  for (int vt : {int(BULKX_INT32), int(BULKX_FLOAT32), int(BULKX_INT64)})
    ws = std::max(ws, reduce_by_key_workspace_bytes(BULKX_INT32, vt, n));
  BULKX_HIP(hipStreamSynchronize(b->stream));
  return batch_grow(b, tmp, ws);
}

extern "C" int bulkx_batch_set_jobs(bulkx_batch* b, const bulkx_job* jobs, int num_jobs) {
  if (!b || num_jobs < 0 || num_jobs > b->max_jobs || (num_jobs > 0 && !jobs))
    return BULKX_ERR_INVALID_ARGUMENT;
  // Validate and lay out on the host first; nothing is touched on failure.
  int64_t elements = 0, tiles0 = 0, sort_tiles = 0;
  size_t status_words = 0, tmp_bytes = 0, ws_need = 256;
  for (int i = 0; i < num_jobs; ++i) {
    const bulkx_job& j = jobs[i];
    BULKX_TRY(batch_check_job(j));
    const int op = batch_op(j);
    const int64_t n = j.n0 + (op == OP_MERGE_I32 || op == OP_MERGE_F32 ? j.n1 : 0);
    elements += n;
    ws_need = std::max(ws_need, per_call_workspace(j));
    if (n == 0 && j.kind != BULKX_JOB_REDUCE && j.kind != BULKX_JOB_REDUCE_BY_KEY) continue;
    const int fam = host_family(op);
    const int64_t T = fam_tile(fam, b->shape[fam]);
    int64_t t;
    switch (op) {
      case OP_SCAN_I32: case OP_SCAN_F32: case OP_REDUCE_I32: case OP_REDUCE_F32:
      case OP_RBK_I32_I32: case OP_RBK_I32_F32: case OP_RBK_I32_I64:
        t = (n + T - 1) / T;
        if (t > 1) {
          status_words += size_t(t) * kStatusWordsPerTile;
          if (op >= OP_RBK_I32_I32) tmp_bytes += rbk_tmp_bytes(t, T, dtype_size(j.value_dtype));
        }
        break;
      default:
        t = (n + T - 1) / T;
        if (op >= OP_SORT_I32) {
          t = 1;  // one block-sorted tile, or the radix passes (no phase-0 tile)
          if (n > kBatchSortRun) {
            const int64_t rt = (n + kBatchRadixTile - 1) / kBatchRadixTile;
            sort_tiles += rt;
            tmp_bytes += radix_tmp_bytes(n);
            t = 0;
          }
        }
        break;
    }
    if (t < 1 && op < OP_SORT_I32) t = 1;  // empty reduce / reduce_by_key: one tile writes the result
    tiles0 += t;
  }
  if (elements > b->max_elements || tiles0 > b->max_tiles || sort_tiles > b->max_tiles ||
      status_words > b->status_words_cap)
    return BULKX_ERR_BUDGET;

  // The staging buffers may still feed the previous upload.
  BULKX_HIP(hipStreamSynchronize(b->stream));
  BULKX_TRY(batch_grow(b, tmp_bytes, ws_need));
  b->jobs.assign(jobs, jobs + num_jobs);
  b->num_jobs = num_jobs;
  // Phase-0 tiles are grouped by family, jobs in list order within a family.
  auto job_tiles = [b](int op, int64_t n) {
    if (op >= OP_SORT_I32) return n <= kBatchSortRun ? int64_t(1) : int64_t(0);
    const int fam = host_family(op);
    const int64_t T = fam_tile(fam, b->shape[fam]);
    const int64_t t = (n + T - 1) / T;
    return t < 1 ? int64_t(1) : t;  // an empty reduce / reduce_by_key still writes its result
  };
  auto job_n = [](const bulkx_job& j, int op) {
    return j.n0 + (op == OP_MERGE_I32 || op == OP_MERGE_F32 ? j.n1 : 0);
  };
  auto skipped = [](const bulkx_job& j, int64_t n) {
    return n == 0 && !(j.kind == BULKX_JOB_REDUCE || j.kind == BULKX_JOB_REDUCE_BY_KEY);
  };
  int64_t fam_tiles[FAM_COUNT] = {};
  for (int i = 0; i < num_jobs; ++i) {
    const int op = batch_op(jobs[i]);
    const int64_t n = job_n(jobs[i], op);
    if (!skipped(jobs[i], n)) fam_tiles[host_family(op)] += job_tiles(op, n);
  }
  int64_t cursor[FAM_COUNT];
  memset(b->h_desc, 0, sizeof(BatchDesc));
  b->h_desc->guard.spin_limit = host_spin_limit();
  b->h_desc->fam_begin[0] = 0;
  for (int f = 0; f < FAM_COUNT; ++f) {
    cursor[f] = b->h_desc->fam_begin[f];
    b->h_desc->fam_begin[f + 1] = b->h_desc->fam_begin[f] + fam_tiles[f];
  }
  int64_t ts = 0, ms = 0;
  size_t sw = 0, tb = 0;
  std::vector<int> sort_jobs;
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
    const int64_t n = job_n(j, op);
    if (skipped(j, n)) {
      dj.op = OP_NONE;
      continue;
    }
    const int64_t t = job_tiles(op, n);
    const int f = host_family(op);
    const bool chained = op <= OP_RBK_I32_I64;
    dj.tiles0 = t;
    dj.tile_begin0 = cursor[f];
    if (chained && t > 1) {
      dj.status = b->d_status + sw;
      sw += size_t(t) * kStatusWordsPerTile;
    }
    if (op >= OP_RBK_I32_I32 && op <= OP_RBK_I32_I64 && t > 1) {  // the compact form's scratch
      const size_t slots = size_t(t) * size_t(fam_tile(f, b->shape[f]) + 1);
      dj.tmp_k = b->d_tmp + tb;
      tb += align_up(slots * 4, 256);
      dj.tmp_v = b->d_tmp + tb;
      tb += align_up(slots * dtype_size(j.value_dtype), 256);
      dj.counts = reinterpret_cast<uint32_t*>(b->d_tmp + tb);
      tb += align_up(size_t(t) * 4, 256);
    }
    if (op == OP_MERGE_I32 || op == OP_MERGE_F32) {
      dj.splits = b->d_merge_splits + ms;
      ms += t + 1;
    }
    if (op >= OP_SORT_I32 && n > kBatchSortRun) {
      dj.passes = kRadixPasses;
      dj.sort_tiles = (n + kBatchRadixTile - 1) / kBatchRadixTile;
      dj.tmp_k = b->d_tmp + tb;
      tb += align_up(size_t(n) * 4, 256);
      dj.tmp_v = b->d_tmp + tb;
      tb += align_up(size_t(n) * 4, 256);
      dj.counts = reinterpret_cast<uint32_t*>(b->d_tmp + tb);
      tb += align_up(size_t(dj.sort_tiles) * kRadixDigits * sizeof(uint32_t), 256);
      dj.totals = reinterpret_cast<uint32_t*>(b->d_tmp + tb);
      tb += align_up(kRadixDigits * sizeof(uint32_t), 256);
      sort_jobs.push_back(i);
    }
    for (int64_t k = 0; k < t; ++k) b->h_tile_job[cursor[f] + k] = uint32_t(i);
    cursor[f] += t;
  }
  // multi-tile reduce_by_key jobs by family, for the aggregate scans
  {
    int64_t c = 0;
    for (int r = 0; r < kRbkFamilies; ++r) {
      b->h_desc->rbk_begin[r] = c;
      for (int i = 0; i < num_jobs; ++i) {
        const BatchJobDev& dj = b->h_jobs[i];
        if (dj.tiles0 > 1 && host_family(dj.op) == FAM_RBK_I32 + r &&
            dj.op >= OP_RBK_I32_I32 && dj.op <= OP_RBK_I32_I64)
          b->h_rbk_jobs[c++] = uint32_t(i);
      }
      b->rbk_jobs[r] = c - b->h_desc->rbk_begin[r];
    }
    b->h_desc->rbk_begin[kRbkFamilies] = c;
    if (c > 0)
      BULKX_HIP(hipMemcpyAsync(b->d_rbk_jobs, b->h_rbk_jobs, sizeof(uint32_t) * c,
                               hipMemcpyHostToDevice, b->stream));
  }
  // the radix tiles of the sort jobs that run the radix passes, job by job
  int64_t nr = 0;
  for (int i : sort_jobs) {
    BatchJobDev& dj = b->h_jobs[i];
    dj.sort_tile_begin = ts;
    for (int64_t k = 0; k < dj.sort_tiles; ++k) b->h_sort_tile_job[ts + k] = uint32_t(i);
    ts += dj.sort_tiles;
    b->h_radix_jobs[nr++] = uint32_t(i);
  }
  b->num_radix_jobs = nr;
  b->h_desc->num_radix_jobs = nr;
  if (nr > 0)
    BULKX_HIP(hipMemcpyAsync(b->d_radix_jobs, b->h_radix_jobs, sizeof(uint32_t) * nr,
                             hipMemcpyHostToDevice, b->stream));
  const int64_t t0 = b->h_desc->fam_begin[FAM_COUNT];
  for (int f = 0; f < FAM_COUNT; ++f) b->fam_tiles[f] = fam_tiles[f];
  b->h_desc->total_sort_tiles = ts;
  b->h_desc->num_jobs = num_jobs;
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

static void batch_drop_one(BatchGraph& G) {
  if (G.exec) (void)hipGraphExecDestroy(G.exec);
  if (G.graph) (void)hipGraphDestroy(G.graph);
  G.exec = nullptr;
  G.graph = nullptr;
  G.sized_nodes.clear();
}

static void batch_drop_graph(bulkx_batch* b) {
  batch_drop_one(b->graph);
}

// Records clear + phase-0 + the radix passes once on the private
// capture stream and finds the kernel nodes whose grids follow the job list.
static int batch_capture(bulkx_batch* b, BatchGraph& G) {
  BULKX_HIP(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
  int st = batch_enqueue_graph(b, b->capture_stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->capture_stream, &g);
  if (st != BULKX_SUCCESS) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) {
    set_last_error("hipStreamEndCapture", e);
    return BULKX_ERR_HIP;
  }
  G.graph = g;
  size_t num = 0;
  BULKX_HIP(hipGraphGetNodes(g, nullptr, &num));
  std::vector<hipGraphNode_t> nodes(num);
  BULKX_HIP(hipGraphGetNodes(g, nodes.data(), &num));
  for (hipGraphNode_t node : nodes) {
    hipGraphNodeType ty;
    BULKX_HIP(hipGraphNodeGetType(node, &ty));
    if (ty != hipGraphNodeTypeKernel) continue;
    hipKernelNodeParams kp;
    BULKX_HIP(hipGraphKernelNodeGetParams(node, &kp));
    int kind = -1;
    for (int f = 0; f < FAM_COUNT; ++f)
      if (kPhase0[b->shape[f]][f] && kp.func == reinterpret_cast<void*>(kPhase0[b->shape[f]][f]))
        kind = f;
    if (kp.func == reinterpret_cast<void*>(kMergePartition[b->shape[FAM_MERGE]]))
      kind = kGridMergePartition;
    for (int r = 0; r < kRbkFamilies; ++r) {
      if (kp.func == reinterpret_cast<void*>(kBatchRbkCompact[b->shape[FAM_RBK_I32 + r]][r]) ||
          kp.func == reinterpret_cast<void*>(kBatchRbkMove[b->shape[FAM_RBK_I32 + r]][r]))
        kind = FAM_RBK_I32 + r;
      if (kp.func == reinterpret_cast<void*>(kRbkScan[r])) kind = grid_kind_rbk_scan(FAM_RBK_I32 + r);
    }
    if (kp.func == reinterpret_cast<void*>(batch_radix_upsweep_kernel)) kind = kGridRadixUp;
    if (kp.func == reinterpret_cast<void*>(batch_radix_scan_kernel)) kind = kGridRadixScan;
    if (kp.func == reinterpret_cast<void*>(batch_radix_downsweep_kernel)) kind = kGridRadixDown;
    if (kind >= 0) G.sized_nodes.emplace_back(node, kind);
  }
  BULKX_HIP(hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0));
  for (int k = 0; k < kGridKinds; ++k) G.graph_grid[k] = batch_grid(b, k);
  return BULKX_SUCCESS;
}

// The launch's error word into pinned host memory, behind the launch.
static int batch_copy_err(bulkx_batch* b) {
  BULKX_HIP(hipMemcpyAsync(b->h_err, &b->d_desc->guard.err, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, b->stream));
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_launch(bulkx_batch* b, int use_graph) {
  if (!b) return BULKX_ERR_INVALID_ARGUMENT;
  b->last_per_call = false;
  if (use_graph == 0) {
    b->timed = true;
    BULKX_TRY(batch_enqueue_kernels(b, b->stream, true));
    return batch_copy_err(b);
  }
  if (use_graph == 2) {  // the graph's branches as eager launches on the side streams
    b->timed = false;
    BULKX_TRY(batch_enqueue_graph(b, b->stream));
    return batch_copy_err(b);
  }
  BatchGraph& G = b->graph;
  if (G.exec) {  // resize the job-list-sized grids in place
    bool stale[kGridKinds];
    bool any = false;
    for (int k = 0; k < kGridKinds; ++k) {
      stale[k] = G.graph_grid[k] != batch_grid(b, k);
      any |= stale[k];
    }
    hipError_t e = hipSuccess;
    for (const auto& nk : G.sized_nodes) {
      if (!any || e != hipSuccess) break;
      if (!stale[nk.second]) continue;
      hipKernelNodeParams kp;
      e = hipGraphKernelNodeGetParams(nk.first, &kp);
      if (e == hipSuccess) {
        kp.gridDim = dim3(batch_grid(b, nk.second));
        e = hipGraphExecKernelNodeSetParams(G.exec, nk.first, &kp);
      }
    }
    if (e != hipSuccess) {  // a node could not be updated: record the graph again
      (void)hipGetLastError();
      batch_drop_one(G);
    } else {
      for (int k = 0; k < kGridKinds; ++k) G.graph_grid[k] = batch_grid(b, k);
    }
  }
  if (!G.exec) BULKX_TRY(batch_capture(b, G));
  BULKX_HIP(hipGraphLaunch(G.exec, b->stream));
  return batch_copy_err(b);
}

extern "C" int bulkx_batch_status(bulkx_batch* b) {
  if (!b) return BULKX_ERR_INVALID_ARGUMENT;
  // the plain entry points report through the device's guard
  if (b->last_per_call) return bulkx_stream_status(b->stream);
  BULKX_HIP(hipStreamSynchronize(b->stream));
  return *b->h_err ? BULKX_ERR_LOOKBACK : BULKX_SUCCESS;
}

extern "C" int bulkx_batch_launch_per_call(bulkx_batch* b) {
  if (!b) return BULKX_ERR_INVALID_ARGUMENT;
  b->last_per_call = true;
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
                                j.in3, j.n1, j.out0, j.out1, b->d_ws, b->ws_bytes, s);
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
  BULKX_HIP(hipEventSynchronize(b->ev[kBatchStages]));
  float a = 0.f, c = 0.f;
  BULKX_HIP(hipEventElapsedTime(&a, b->ev[0], b->ev[FAM_COUNT]));
  BULKX_HIP(hipEventElapsedTime(&c, b->ev[FAM_COUNT], b->ev[kBatchStages]));
  if (phase0_ms) *phase0_ms = a;
  if (passes_ms) *passes_ms = c;
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_prepare_graphs(bulkx_batch* b) {
  if (!b) return BULKX_ERR_INVALID_ARGUMENT;
  if (b->graph.exec) return BULKX_SUCCESS;
  return batch_capture(b, b->graph);
}

extern "C" int bulkx_batch_sort_geometry(int64_t* run_elements, int64_t* radix_tile) {
  if (!run_elements || !radix_tile) return BULKX_ERR_INVALID_ARGUMENT;
  *run_elements = kBatchSortRun;
  *radix_tile = kBatchRadixTile;
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_scratch_bytes(bulkx_batch* b, size_t* tmp_bytes, size_t* workspace_bytes) {
  if (!b) return BULKX_ERR_INVALID_ARGUMENT;
  if (tmp_bytes) *tmp_bytes = b->tmp_bytes_cap;
  if (workspace_bytes) *workspace_bytes = b->ws_bytes;
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_family_shapes(bulkx_batch* b, int* shape, int64_t* tile_elements) {
  if (!b) return BULKX_ERR_INVALID_ARGUMENT;
  for (int f = 0; f < FAM_COUNT; ++f) {
    if (shape) shape[f] = b->shape[f];
    if (tile_elements) tile_elements[f] = fam_tile(f, b->shape[f]);
  }
  return BULKX_SUCCESS;
}

extern "C" int bulkx_batch_stage_timing(bulkx_batch* b, float* stage_ms, int num_stages) {
  if (!b || !stage_ms || num_stages < 0) return BULKX_ERR_INVALID_ARGUMENT;
  if (!b->timed) return BULKX_ERR_NOT_READY;
  BULKX_HIP(hipEventSynchronize(b->ev[kBatchStages]));
  // stages [0, kBatchStages): the families and the radix passes; then the
  // passes' downsweep, upsweep and counts-scan kernels on their own
  float up_ms = 0.f, scan_ms = 0.f, down_ms = 0.f;
  for (int p = 0; p < kRadixPasses && num_stages > kBatchStages; ++p) {
    float a = 0.f, c = 0.f, e = 0.f;
    BULKX_HIP(hipEventElapsedTime(&a, b->rev[3 * p], b->rev[3 * p + 1]));
    BULKX_HIP(hipEventElapsedTime(&c, b->rev[3 * p + 1], b->rev[3 * p + 2]));
    BULKX_HIP(hipEventElapsedTime(&e, b->rev[3 * p + 2], b->rev[3 * p + 3]));
    up_ms += a;
    scan_ms += c;
    down_ms += e;
  }
  for (int k = 0; k < num_stages; ++k) {
    float ms = 0.f;
    if (k < kBatchStages) BULKX_HIP(hipEventElapsedTime(&ms, b->ev[k], b->ev[k + 1]));
    else if (k == kBatchStages) ms = down_ms;
    else if (k == kBatchStages + 1) ms = up_ms;
    else if (k == kBatchStages + 2) ms = scan_ms;
    stage_ms[k] = ms;
  }
  return BULKX_SUCCESS;
}
