// stable_sort_by_key entry point.
#include <algorithm>

#include "sort_impl.hpp"
#include "tuning.hpp"

namespace bulkx {

size_t sort_workspace_bytes(int kdtype, int vdtype, int64_t n) {
  size_t ks = dtype_size(kdtype);
  if (ks == 0 || n <= 0) return 0;
  const int64_t tile = ks == 4 ? SortConfig<float>::TILE : SortConfig<double>::TILE;
  if (n <= tile) return 0;
  size_t vs = vdtype >= 0 ? dtype_size(vdtype) : 0;
  size_t bytes = SortWorkspace(ks, vs, n, (n + tile - 1) / tile).total;
  // any variant may run: the radix form's digit counts need more than the merge splits
  if (ks == 4 && n > kRadixTile && n <= kRadixMaxN)
    bytes = std::max(bytes, RadixWorkspace(ks, vs, n).total);
  return bytes;
}

}  // namespace bulkx

using namespace bulkx;

extern "C" int bulkx_stable_sort_by_key(int key_dtype, int value_dtype, int comp, void* keys,
                                        void* values, int64_t n, void* workspace,
                                        size_t workspace_bytes, void* stream_) {
  if (dtype_size(key_dtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  if (value_dtype >= 0 && dtype_size(value_dtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  if (!valid_cmp(comp)) return BULKX_ERR_UNSUPPORTED_OP;
  if (n < 0) return BULKX_ERR_INVALID_ARGUMENT;
  if (n == 0) return BULKX_SUCCESS;
  if (!keys || (value_dtype >= 0 && !values)) return BULKX_ERR_INVALID_ARGUMENT;
  if (n > (int64_t(1) << 40)) return BULKX_ERR_INVALID_ARGUMENT;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  Scratch scratch;
  BULKX_TRY(scratch.acquire(workspace, workspace_bytes,
                            sort_workspace_bytes(key_dtype, value_dtype, n), stream));
  const int v = tuned_variant(BULKX_ALGO_SORT_BY_KEY, key_dtype, n);
  int st;
  if (dtype_size(key_dtype) == 4)
    st = sort_dispatch_k4(key_dtype, value_dtype, comp, keys, values, n, scratch.ptr(), stream, v);
  else
    st = sort_dispatch_k8(key_dtype, value_dtype, comp, keys, values, n, scratch.ptr(), stream, v);
  int rs = scratch.release();
  return st != BULKX_SUCCESS ? st : rs;
}
