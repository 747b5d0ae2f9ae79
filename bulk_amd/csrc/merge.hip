// merge / merge_by_key: replaces my_merge (merge.cu:127-162) and the
// device-wide use of bulk::merge_by_key (merge.hpp:507-607).  Two launches,
// as the reference's locate_merge_path + merge_kernel (merge.cu:100-124,
// 39-97): a partition kernel finds every tile boundary (one lane per
// boundary, binary search), then each workgroup stages its two input slices
// in LDS and merges them.  Splitting the search out keeps its chain of
// dependent loads off the merge's critical path.
#include "merge_impl.hpp"
#include "tuning.hpp"

namespace bulkx {

// splits[t] = merge path of diagonal min(t * TILE, n1 + n2), t in [0, tiles].
template <class CMP, class K, int TILE>
__global__ __launch_bounds__(256) void merge_split_kernel(const K* __restrict__ a, int64_t n1,
                                                          const K* __restrict__ b, int64_t n2,
                                                          int64_t tiles,
                                                          int64_t* __restrict__ splits) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t > tiles) return;
  const int64_t n = n1 + n2;
  const int64_t d = t * TILE < n ? t * TILE : n;
  splits[t] = lane_merge_path<CMP>(a, n1, b, n2, d);
}

template <class CMP, class K, class V, bool HAS_V, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void merge_kernel(const K* __restrict__ a,
                                                      const V* __restrict__ va, int64_t n1,
                                                      const K* __restrict__ b,
                                                      const V* __restrict__ vb, int64_t n2,
                                                      K* __restrict__ out, V* __restrict__ vout,
                                                      const int64_t* __restrict__ splits) {
  __shared__ MergeSmem<K, V, HAS_V, BLOCK, ITEMS> sm;
  constexpr int TILE = BLOCK * ITEMS;
  const int64_t tile = blockIdx.x;
  const int64_t n = n1 + n2;
  const int64_t d0 = tile * TILE;
  const int64_t d1 = d0 + TILE < n ? d0 + TILE : n;
  // a single-tile merge runs without splits: the tile is all of A and B
  const int64_t a0 = splits ? splits[tile] : 0;
  const int64_t a1 = splits ? splits[tile + 1] : n1;
  const int64_t b0 = d0 - a0, b1 = d1 - a1;
  merge_tile<CMP, K, V, HAS_V, BLOCK, ITEMS>(a + a0, HAS_V ? va + a0 : nullptr, int(a1 - a0),
                                             b + b0, HAS_V ? vb + b0 : nullptr, int(b1 - b0),
                                             out + d0, HAS_V ? vout + d0 : nullptr, sm.k, sm.v);
}

template <class CMP, class K, class V, bool HAS_V, int VAR>
static int merge_launch_v(const void* k1, const void* v1, int64_t n1, const void* k2,
                          const void* v2, int64_t n2, void* kout, void* vout, char* ws,
                          hipStream_t stream) {
  using C = MergeConfig<K, VAR>;
  const int64_t tiles = (n1 + n2 + C::TILE - 1) / C::TILE;
  int64_t* splits = nullptr;
  if (tiles > 1) {
    splits = reinterpret_cast<int64_t*>(ws);
    const int64_t groups = (tiles + 1 + 255) / 256;
    merge_split_kernel<CMP, K, C::TILE><<<dim3(uint32_t(groups)), 256, 0, stream>>>(
        static_cast<const K*>(k1), n1, static_cast<const K*>(k2), n2, tiles, splits);
    BULKX_LAUNCHED();
  }
  merge_kernel<CMP, K, V, HAS_V, C::BLOCK, C::ITEMS><<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(
      static_cast<const K*>(k1), static_cast<const V*>(v1), n1, static_cast<const K*>(k2),
      static_cast<const V*>(v2), n2, static_cast<K*>(kout), static_cast<V*>(vout), splits);
  BULKX_LAUNCHED();
  return BULKX_SUCCESS;
}

template <class CMP, class K, class V, bool HAS_V>
static int merge_launch(const void* k1, const void* v1, int64_t n1, const void* k2, const void* v2,
                        int64_t n2, void* kout, void* vout, char* ws, hipStream_t stream,
                        int variant) {
  if (variant == 1)
    return merge_launch_v<CMP, K, V, HAS_V, 1>(k1, v1, n1, k2, v2, n2, kout, vout, ws, stream);
  return merge_launch_v<CMP, K, V, HAS_V, 0>(k1, v1, n1, k2, v2, n2, kout, vout, ws, stream);
}

// One int64 split per tile boundary (every variant has the same tile).
size_t merge_workspace_bytes(int kdtype, int, int64_t n1, int64_t n2) {
  const size_t ks = dtype_size(kdtype);
  if (ks == 0 || n1 < 0 || n2 < 0) return 0;
  const int64_t tile = ks == 4 ? MergeConfig<float>::TILE : MergeConfig<double>::TILE;
  const int64_t tiles = (n1 + n2 + tile - 1) / tile;
  return tiles > 1 ? align_up(size_t(tiles + 1) * sizeof(int64_t), 256) : 0;
}

static int merge_entry(int kdtype, int vdtype, int comp, const void* k1, const void* v1,
                       int64_t n1, const void* k2, const void* v2, int64_t n2, void* kout,
                       void* vout, void* workspace, size_t workspace_bytes, void* stream_) {
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
  Scratch scratch;
  BULKX_TRY(scratch.acquire(workspace, workspace_bytes, merge_workspace_bytes(kdtype, vdtype, n1, n2),
                            stream));
  char* ws = scratch.ptr();
  int st = BULKX_SUCCESS;
  // The payload is only moved, so it is dispatched by size.
  BULKX_DISPATCH_DTYPE(kdtype, K, BULKX_DISPATCH_CMP(comp, CMP, {
    if (!has_v)
      st = merge_launch<CMP, K, uint32_t, false>(k1, nullptr, n1, k2, nullptr, n2, kout, nullptr,
                                                 ws, stream, v);
    else if (dtype_size(vdtype) == 4)
      st = merge_launch<CMP, K, uint32_t, true>(k1, v1, n1, k2, v2, n2, kout, vout, ws, stream, v);
    else
      st = merge_launch<CMP, K, uint64_t, true>(k1, v1, n1, k2, v2, n2, kout, vout, ws, stream, v);
  }));
  int rs = scratch.release();
  return st != BULKX_SUCCESS ? st : rs;
}

}  // namespace bulkx

using namespace bulkx;

extern "C" int bulkx_merge(int key_dtype, int comp, const void* keys1, int64_t n1,
                           const void* keys2, int64_t n2, void* keys_out, void* workspace,
                           size_t workspace_bytes, void* stream) {
  return merge_entry(key_dtype, -1, comp, keys1, nullptr, n1, keys2, nullptr, n2, keys_out,
                     nullptr, workspace, workspace_bytes, stream);
}

extern "C" int bulkx_merge_by_key(int key_dtype, int value_dtype, int comp, const void* keys1,
                                  const void* values1, int64_t n1, const void* keys2,
                                  const void* values2, int64_t n2, void* keys_out,
                                  void* values_out, void* workspace, size_t workspace_bytes,
                                  void* stream) {
  if (value_dtype < 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  return merge_entry(key_dtype, value_dtype, comp, keys1, values1, n1, keys2, values2, n2,
                     keys_out, values_out, workspace, workspace_bytes, stream);
}
