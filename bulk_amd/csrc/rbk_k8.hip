// reduce_by_key instantiations for 8-byte keys (int64, uint64, float64).
#include "reduce_by_key_impl.hpp"

namespace bulkx {

int rbk_dispatch_k8(int kdtype, int vdtype, int op, const void* keys, const void* vals, int64_t n,
                    void* kout, void* vout, int64_t* num_runs, char* ws, hipStream_t stream,
                    int variant) {
  switch (kdtype) {
    case BULKX_INT64:
      return rbk_dispatch_values<int64_t>(vdtype, op, keys, vals, n, kout, vout, num_runs, ws,
                                         stream, variant);
    case BULKX_UINT64:
      return rbk_dispatch_values<uint64_t>(vdtype, op, keys, vals, n, kout, vout, num_runs, ws,
                                         stream, variant);
    case BULKX_FLOAT64:
      return rbk_dispatch_values<double>(vdtype, op, keys, vals, n, kout, vout, num_runs, ws,
                                         stream, variant);
    default:
      return BULKX_ERR_UNSUPPORTED_DTYPE;
  }
}

}  // namespace bulkx
