// Merge sort instantiations for 8-byte keys (int64, uint64, float64).
#include "sort_impl.hpp"

namespace bulkx {

int sort_dispatch_k8(int kdtype, int vdtype, int comp, void* keys, void* vals, int64_t n,
                     char* ws, hipStream_t stream, int variant) {
  switch (kdtype) {
    case BULKX_INT64:
      return sort_dispatch_key<int64_t>(vdtype, comp, keys, vals, n, ws, stream, variant);
    case BULKX_UINT64:
      return sort_dispatch_key<uint64_t>(vdtype, comp, keys, vals, n, ws, stream, variant);
    case BULKX_FLOAT64:
      return sort_dispatch_key<double>(vdtype, comp, keys, vals, n, ws, stream, variant);
    default:
      return BULKX_ERR_UNSUPPORTED_DTYPE;
  }
}

}  // namespace bulkx
