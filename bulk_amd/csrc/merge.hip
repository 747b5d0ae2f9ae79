// merge / merge_by_key: replaces my_merge (merge.cu:127-162) and the
// device-wide use of bulk::merge_by_key (merge.hpp:507-607).  One launch:
// each workgroup finds its own tile boundaries with a wavefront search,
// stages the two input slices in LDS and merges them.
#include "merge_impl.hpp"
#include "tuning.hpp"

namespace bulkx {

template <class CMP, class K, class V, bool HAS_V, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void merge_kernel(const K* __restrict__ a,
                                                      const V* __restrict__ va, int64_t n1,
                                                      const K* __restrict__ b,
                                                      const V* __restrict__ vb, int64_t n2,
                                                      K* __restrict__ out, V* __restrict__ vout) {
  __shared__ MergeSmem<K, V, HAS_V, BLOCK, ITEMS> sm;
  merge_two_tile<CMP, K, V, HAS_V, BLOCK, ITEMS>(sm, a, va, n1, b, vb, n2, out, vout, blockIdx.x);
}

template <class CMP, class K, class V, bool HAS_V, int VAR>
static int merge_launch_v(const void* k1, const void* v1, int64_t n1, const void* k2,
                          const void* v2, int64_t n2, void* kout, void* vout, hipStream_t stream) {
  using C = MergeConfig<K, VAR>;
  const int64_t tiles = (n1 + n2 + C::TILE - 1) / C::TILE;
  merge_kernel<CMP, K, V, HAS_V, C::BLOCK, C::ITEMS><<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(
      static_cast<const K*>(k1), static_cast<const V*>(v1), n1, static_cast<const K*>(k2),
      static_cast<const V*>(v2), n2, static_cast<K*>(kout), static_cast<V*>(vout));
  BULKX_LAUNCHED();
  return BULKX_SUCCESS;
}

template <class CMP, class K, class V, bool HAS_V>
static int merge_launch(const void* k1, const void* v1, int64_t n1, const void* k2, const void* v2,
                        int64_t n2, void* kout, void* vout, hipStream_t stream, int variant) {
  if (variant == 1)
    return merge_launch_v<CMP, K, V, HAS_V, 1>(k1, v1, n1, k2, v2, n2, kout, vout, stream);
  return merge_launch_v<CMP, K, V, HAS_V, 0>(k1, v1, n1, k2, v2, n2, kout, vout, stream);
}

size_t merge_workspace_bytes(int, int, int64_t, int64_t) { return 0; }

static int merge_entry(int kdtype, int vdtype, int comp, const void* k1, const void* v1,
                       int64_t n1, const void* k2, const void* v2, int64_t n2, void* kout,
                       void* vout, void* stream_) {
  if (dtype_size(kdtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  const bool has_v = vdtype >= 0;
  if (has_v && dtype_size(vdtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  if (!valid_cmp(comp)) return BULKX_ERR_UNSUPPORTED_OP;
  if (n1 < 0 || n2 < 0) return BULKX_ERR_INVALID_ARGUMENT;
  if (n1 + n2 == 0) return BULKX_SUCCESS;
  if ((n1 > 0 && !k1) || (n2 > 0 && !k2) || !kout) return BULKX_ERR_INVALID_ARGUMENT;
  if (has_v && ((n1 > 0 && !v1) || (n2 > 0 && !v2) || !vout)) return BULKX_ERR_INVALID_ARGUMENT;
  if (n1 + n2 > (int64_t(1) << 40)) return BULKX_ERR_INVALID_ARGUMENT;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  const int v = tuned_variant(has_v ? BULKX_ALGO_MERGE_BY_KEY : BULKX_ALGO_MERGE, kdtype, n1 + n2);
  int st = BULKX_SUCCESS;
  // The payload is only moved, so it is dispatched by size.
  BULKX_DISPATCH_DTYPE(kdtype, K, BULKX_DISPATCH_CMP(comp, CMP, {
    if (!has_v)
      st = merge_launch<CMP, K, uint32_t, false>(k1, nullptr, n1, k2, nullptr, n2, kout, nullptr,
                                                 stream, v);
    else if (dtype_size(vdtype) == 4)
      st = merge_launch<CMP, K, uint32_t, true>(k1, v1, n1, k2, v2, n2, kout, vout, stream, v);
    else
      st = merge_launch<CMP, K, uint64_t, true>(k1, v1, n1, k2, v2, n2, kout, vout, stream, v);
  }));
  return st;
}

}  // namespace bulkx

using namespace bulkx;

extern "C" int bulkx_merge(int key_dtype, int comp, const void* keys1, int64_t n1,
                           const void* keys2, int64_t n2, void* keys_out, void* workspace,
                           size_t workspace_bytes, void* stream) {
  (void)workspace;
  (void)workspace_bytes;
  return merge_entry(key_dtype, -1, comp, keys1, nullptr, n1, keys2, nullptr, n2, keys_out,
                     nullptr, stream);
}

extern "C" int bulkx_merge_by_key(int key_dtype, int value_dtype, int comp, const void* keys1,
                                  const void* values1, int64_t n1, const void* keys2,
                                  const void* values2, int64_t n2, void* keys_out,
                                  void* values_out, void* workspace, size_t workspace_bytes,
                                  void* stream) {
  (void)workspace;
  (void)workspace_bytes;
  if (value_dtype < 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  return merge_entry(key_dtype, value_dtype, comp, keys1, values1, n1, keys2, values2, n2,
                     keys_out, values_out, stream);
}
