// Merge sort instantiations for 4-byte keys (int32, uint32, float32).
#include "sort_impl.hpp"

namespace bulkx {

int sort_dispatch_k4(int kdtype, int vdtype, int comp, void* keys, void* vals, int64_t n,
                     char* ws, hipStream_t stream, int variant) {
  switch (kdtype) {
    case BULKX_INT32:
      return sort_dispatch_key<int32_t>(vdtype, comp, keys, vals, n, ws, stream, variant);
    case BULKX_UINT32:
      return sort_dispatch_key<uint32_t>(vdtype, comp, keys, vals, n, ws, stream, variant);
    case BULKX_FLOAT32:
      return sort_dispatch_key<float>(vdtype, comp, keys, vals, n, ws, stream, variant);
    default:
      return BULKX_ERR_UNSUPPORTED_DTYPE;
  }
}

}  // namespace bulkx
