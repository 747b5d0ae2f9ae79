// Launch-shape table and the choose_sizes replacement.
#include <stdio.h>
#include <stdlib.h>

#include "host_common.hpp"
#include "merge_impl.hpp"
#include "reduce_by_key_impl.hpp"
#include "scan_config.hpp"
#include "tuning.hpp"

namespace bulkx {

int choose_shape(int algo, int dtype, int64_t n, LaunchShape* out) {
  size_t es = dtype_size(dtype);
  if (es == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  DeviceProps p;
  int cus = current_device_props(&p) == BULKX_SUCCESS ? p.num_cus : 256;
  LaunchShape s{1, 256, 1};
  switch (algo) {
    case BULKX_ALGO_INCLUSIVE_SCAN:
    case BULKX_ALGO_EXCLUSIVE_SCAN: {
      s.group_size = es == 4 ? ScanConfig<float>::BLOCK : ScanConfig<double>::BLOCK;
      s.grain_size = es == 4 ? ScanConfig<float>::ITEMS : ScanConfig<double>::ITEMS;
      int64_t tile = int64_t(s.group_size) * s.grain_size;
      s.num_groups = (n + tile - 1) / tile;
      break;
    }
    case BULKX_ALGO_REDUCE: {
      s.group_size = 256;
      s.grain_size = int(16 / es) * 4;
      int64_t tile = int64_t(s.group_size) * s.grain_size;
      int64_t want = (n + tile - 1) / tile;
      s.num_groups = want < int64_t(cus) * 8 ? want : int64_t(cus) * 8;
      break;
    }
    case BULKX_ALGO_REDUCE_BY_KEY: {
      s.group_size = RbkConfig<int, int>::BLOCK;
      s.grain_size = RbkConfig<int, int>::ITEMS;
      int64_t tile = int64_t(s.group_size) * s.grain_size;
      s.num_groups = (n + tile - 1) / tile;
      break;
    }
    case BULKX_ALGO_MERGE:
    case BULKX_ALGO_MERGE_BY_KEY:
    case BULKX_ALGO_SORT_BY_KEY: {
      s.group_size = es == 4 ? MergeConfig<float>::BLOCK : MergeConfig<double>::BLOCK;
      s.grain_size = es == 4 ? MergeConfig<float>::ITEMS : MergeConfig<double>::ITEMS;
      int64_t tile = int64_t(s.group_size) * s.grain_size;
      s.num_groups = (n + tile - 1) / tile;
      break;
    }
    case BULKX_ALGO_REDUCE_INTERVALS: {
      s.group_size = 256;
      s.grain_size = 1;
      s.num_groups = n;  // one workgroup per interval (capped at 65536 and looped)
      break;
    }
    default:
      return BULKX_ERR_INVALID_ARGUMENT;
  }
  if (s.num_groups < 1) s.num_groups = 1;
  *out = s;
  return BULKX_SUCCESS;
}

int tuned_variant(int, int, int64_t) { return -1; }

}  // namespace bulkx

using namespace bulkx;

extern "C" int bulkx_choose_sizes(int algo, int dtype, int64_t n, int64_t* num_groups,
                                  int* group_size, int* grain_size) {
  if (n < 0) return BULKX_ERR_INVALID_ARGUMENT;
  LaunchShape s;
  BULKX_TRY(choose_shape(algo, dtype, n, &s));
  if (num_groups) *num_groups = s.num_groups;
  if (group_size) *group_size = s.group_size;
  if (grain_size) *grain_size = s.grain_size;
  return BULKX_SUCCESS;
}

extern "C" int bulkx_load_tuning(const char*, int* entries) {
  if (entries) *entries = 0;
  return BULKX_SUCCESS;
}
