// Tile shapes of the batch executor's op families, shared by the executor
// (batch.hip) and the tuning table (tuning.cpp).  Every family is compiled in
// kBatchShapes shapes; the tuning table (algos BULKX_ALGO_BATCH_*) picks one
// per family when a batch is created, and the autotuner times them on the
// mixed-job stream.
#pragma once

#include <stdint.h>

#include "../../include/bulkx.h"

namespace bulkx {

enum BatchFamily : int32_t {
  FAM_SCAN_I32 = 0, FAM_SCAN_F32, FAM_RBK_I32, FAM_RBK_F32, FAM_RBK_I64, FAM_MERGE, FAM_SORT,
  FAM_COUNT
};

constexpr int kBatchShapes = 2;
constexpr int kBatchBlock = 512;
constexpr int kBatchSortBlock = 512;
constexpr int kBatchSortItems = 16;

constexpr bool fam_is_rbk(int fam) {
  return fam == FAM_RBK_I32 || fam == FAM_RBK_F32 || fam == FAM_RBK_I64;
}

// Threads per workgroup of family `fam` in shape `sh`:
//   scan / reduce     shape 0: 512 x 32 (16384 elements), 1: 512 x 16
//   reduce_by_key     shape 0: 512 x 16 (8192 elements),  1: 256 x 32 (8192)
//   merge             shape 0: 512 x 8 (4096 outputs),    1: 256 x 16
//   sort (one tile)   512 x 16 in both
constexpr int fam_block(int fam, int sh) {
  return fam == FAM_SORT                         ? kBatchSortBlock
         : (fam == FAM_MERGE || fam_is_rbk(fam)) && sh == 1 ? 256
                                                            : kBatchBlock;
}
constexpr int fam_items(int fam, int sh) {
  return (fam == FAM_SCAN_I32 || fam == FAM_SCAN_F32) ? (sh == 0 ? 32 : 16)
         : fam_is_rbk(fam)                           ? (sh == 0 ? 16 : 32)
         : fam == FAM_MERGE                          ? (sh == 0 ? 8 : 16)
                                                     : kBatchSortItems;
}
constexpr int64_t fam_tile(int fam, int sh) { return int64_t(fam_block(fam, sh)) * fam_items(fam, sh); }

// The smallest phase-0 tile of any family and shape (sizes the tile tables).
constexpr int64_t batch_min_tile() {
  int64_t m = fam_tile(0, 0);
  for (int f = 0; f < FAM_COUNT; ++f)
    for (int s = 0; s < kBatchShapes; ++s) m = fam_tile(f, s) < m ? fam_tile(f, s) : m;
  return m;
}

// The tuning-table key of a family: (algo, dtype).
constexpr int fam_tuning_algo(int fam) {
  return (fam == FAM_SCAN_I32 || fam == FAM_SCAN_F32) ? BULKX_ALGO_BATCH_SCAN
         : fam_is_rbk(fam)                           ? BULKX_ALGO_BATCH_REDUCE_BY_KEY
         : fam == FAM_MERGE                          ? BULKX_ALGO_BATCH_MERGE
                                                     : -1;
}
constexpr int fam_tuning_dtype(int fam) {
  return fam == FAM_SCAN_F32 || fam == FAM_RBK_F32 ? BULKX_FLOAT32
         : fam == FAM_RBK_I64                      ? BULKX_INT64
                                                   : BULKX_INT32;
}

}  // namespace bulkx
