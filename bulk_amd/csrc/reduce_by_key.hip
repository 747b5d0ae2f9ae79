// reduce_by_key entry point, its preparation kernel, and the materialised
// head / tail flags (tail_flags.hpp:26-109, bulk/detail/head_flags.hpp:126-232).
#include "reduce_by_key_impl.hpp"
#include "tuning.hpp"

namespace bulkx {

__global__ void rbk_empty_kernel(int64_t* num_runs) {
  if (threadIdx.x == 0) *num_runs = 0;
}

__global__ void rbk_prepare_kernel(uint32_t* header, uint64_t* words, int64_t num_words) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0) {
    header[0] = 0;
    header[1] = 0;
  }
  for (; i < num_words; i += int64_t(gridDim.x) * blockDim.x) words[i] = 0;
}

size_t reduce_by_key_workspace_bytes(int kdtype, int vdtype, int64_t n) {
  size_t ks = dtype_size(kdtype), vs = dtype_size(vdtype);
  if (ks == 0 || vs == 0 || n <= 0) return 0;
  const int64_t tiles = (n + kRbkMinTile - 1) / kRbkMinTile;  // fits every tile shape
  if (tiles <= 1) return 0;
  // look-back records (single pass) or aggregates, totals and flags (three pass)
  const size_t words = kRbkHeaderBytes + size_t(tiles) * kRbkMaxStatusWords * sizeof(uint64_t);
  // the compact form's aggregates, totals, meta and scratch runs (T + 1
  // slots per tile of T), for each of its tile sizes, with the widest
  // aggregate and the shortest scan segment (so the most segment totals)
  size_t compact = 0;
  for (int64_t T : {int64_t(kRbkMinTile), int64_t(kRbkCompactMaxTile)}) {
    const RbkCompactLayout lay(ks, vs, kRbkMaxStatusWords * 4, n, T, rbk_scan_seg<Seg<dd>>());
    compact = lay.total > compact ? lay.total : compact;
  }
  return words > compact ? words : compact;
}

template <class K, bool TAIL>
__global__ void flags_kernel(const K* __restrict__ keys, int64_t n, uint8_t* __restrict__ flags) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    bool f;
    if (TAIL) f = (i == n - 1) || !(keys[i] == keys[i + 1]);
    else f = (i == 0) || !(keys[i] == keys[i - 1]);
    flags[i] = f ? 1 : 0;
  }
}

template <bool TAIL>
static int flags_entry(int kdtype, int pred, const void* keys, int64_t n, uint8_t* flags,
                       void* stream_) {
  if (dtype_size(kdtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  if (pred != BULKX_EQUAL_TO) return BULKX_ERR_UNSUPPORTED_OP;
  if (n < 0) return BULKX_ERR_INVALID_ARGUMENT;
  if (n == 0) return BULKX_SUCCESS;
  if (keys == nullptr || flags == nullptr) return BULKX_ERR_INVALID_ARGUMENT;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  BULKX_DISPATCH_DTYPE(kdtype, K, {
    flags_kernel<K, TAIL><<<int(blocks), 256, 0, stream>>>(static_cast<const K*>(keys), n, flags);
  });
  BULKX_LAUNCHED();
  return BULKX_SUCCESS;
}

}  // namespace bulkx

using namespace bulkx;

extern "C" int bulkx_reduce_by_key(int key_dtype, int value_dtype, int pred, int op,
                                   const void* keys, const void* values, int64_t n,
                                   void* keys_out, void* values_out, int64_t* num_runs,
                                   void* workspace, size_t workspace_bytes, void* stream_) {
  if (dtype_size(key_dtype) == 0 || dtype_size(value_dtype) == 0)
    return BULKX_ERR_UNSUPPORTED_DTYPE;
  if (pred != BULKX_EQUAL_TO || !valid_op(op)) return BULKX_ERR_UNSUPPORTED_OP;
  if (n < 0 || num_runs == nullptr) return BULKX_ERR_INVALID_ARGUMENT;
  if (n > 0 && (!keys || !values || !keys_out || !values_out)) return BULKX_ERR_INVALID_ARGUMENT;
  if (n > (int64_t(1) << 40)) return BULKX_ERR_INVALID_ARGUMENT;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  Scratch scratch;
  BULKX_TRY(scratch.acquire(workspace, workspace_bytes,
                            reduce_by_key_workspace_bytes(key_dtype, value_dtype, n), stream));
  const int v = tuned_variant(BULKX_ALGO_REDUCE_BY_KEY, value_dtype, n);
  int st;
  if (dtype_size(key_dtype) == 4)
    st = rbk_dispatch_k4(key_dtype, value_dtype, op, keys, values, n, keys_out, values_out,
                         num_runs, scratch.ptr(), stream, v);
  else
    st = rbk_dispatch_k8(key_dtype, value_dtype, op, keys, values, n, keys_out, values_out,
                         num_runs, scratch.ptr(), stream, v);
  int rs = scratch.release();
  return st != BULKX_SUCCESS ? st : rs;
}

extern "C" int bulkx_tail_flags(int key_dtype, int pred, const void* keys, int64_t n,
                                uint8_t* flags, void* stream) {
  return flags_entry<true>(key_dtype, pred, keys, n, flags, stream);
}

extern "C" int bulkx_head_flags(int key_dtype, int pred, const void* keys, int64_t n,
                                uint8_t* flags, void* stream) {
  return flags_entry<false>(key_dtype, pred, keys, n, flags, stream);
}
