// reduce_by_key instantiations for 4-byte keys (int32, uint32, float32).
#include "reduce_by_key_impl.hpp"

namespace bulkx {

int rbk_dispatch_k4(int kdtype, int vdtype, int op, const void* keys, const void* vals, int64_t n,
                    void* kout, void* vout, int64_t* num_runs, char* ws, hipStream_t stream,
                    int variant) {
  switch (kdtype) {
    case BULKX_INT32:
      return rbk_dispatch_values<int32_t>(vdtype, op, keys, vals, n, kout, vout, num_runs, ws,
                                         stream, variant);
    case BULKX_UINT32:
      return rbk_dispatch_values<uint32_t>(vdtype, op, keys, vals, n, kout, vout, num_runs, ws,
                                         stream, variant);
    case BULKX_FLOAT32:
      return rbk_dispatch_values<float>(vdtype, op, keys, vals, n, kout, vout, num_runs, ws,
                                         stream, variant);
    default:
      return BULKX_ERR_UNSUPPORTED_DTYPE;
  }
}

}  // namespace bulkx
