// Stable merge sort by key: replaces stable_merge_sort_by_key
// (merge_sort_by_key.cu:150-209).
//
// Phase 1 (block sort, as stable_sort_each_kernel merge_sort_by_key.cu:12-25
// -> stable_merge_sort.hpp:125-206): each workgroup loads a tile of
// BLOCK*ITEMS keys, every thread sorts its grain by odd-even transposition
// (sort.hpp:33-75, stable with a strict comparator), then log2(BLOCK)
// rounds of LDS merge-path merges double the sorted runs.  Only (key, tile
// index) pairs move between rounds, packed into one LDS element so that the
// merge reads the index with the key; values wait in registers and are
// staged and gathered once at the end.  A barrier separates every read of
// the stage from the next write to it, so the reference's shared key/value
// stage hazard (stable_merge_sort.hpp:97-115) cannot occur.
//
// Phase 2 (merge passes, as merge_by_key_kernel merge_sort_by_key.cu:61-87
// and locate_merge_path :90-131): runs of width W = TILE, 2*TILE, ... are
// merged pairwise, one output tile per workgroup, ping-ponging between the
// caller's arrays and the workspace; the block sort writes to whichever
// buffer makes the last pass land in the caller's arrays, so no copy-back
// (the reference's K12, merge_sort_by_key.cu:204-208) is needed.
#pragma once

#include "merge_impl.hpp"
#include "radix_impl.hpp"

namespace bulkx {

template <class K, int VAR = 0>
using SortConfig = MergeConfig<K, VAR>;

// Sort shapes: the merge passes use the merge tile of MergeConfig (variant
// 0 or 1); the block sort writes runs of RUN0 elements.  Variants 2 and 3
// sort 8192- and 16384-element runs (512 and 1024 threads x 16) in LDS, so
// one or two fewer passes stream the whole array through HBM; they are
// compiled for 4-byte keys and values (LDS and registers), other types use
// variant 0.
template <class K, class V, int VAR>
struct SortShape {
  static constexpr bool kBig = VAR >= 2 && sizeof(K) == 4 && sizeof(V) == 4;
  using Pass = MergeConfig<K, VAR == 1 ? 1 : 0>;
  static constexpr int BBLOCK = kBig ? (VAR == 2 ? 512 : 1024) : Pass::BLOCK;
  static constexpr int BITEMS = kBig ? 16 : Pass::ITEMS;
  static constexpr int RUN0 = BBLOCK * BITEMS;
  static_assert(RUN0 % Pass::TILE == 0, "runs are whole pass tiles");
};
// Variant 4: the radix form (radix_impl.hpp) for 4-byte keys; 8-byte keys
// run variant 0.
constexpr int kSortVariants = 5;
constexpr int kSortRadixVariant = 4;

// A key and its index in the tile: 8 bytes for 4-byte keys (one ds_read_b64
// per merge step), 16 for 8-byte keys.
template <class K>
struct alignas(2 * sizeof(K)) KeyIdx {
  K k;
  uint32_t i;
};

// One LDS stage, used in turn for the key transposes of the load, the
// (key, index) rounds, the value gather and the transposes of the store:
// 35 KB for 4096 4-byte keys and values, so four workgroups share a CU.
template <class K, class V, bool HAS_V, int BLOCK, int ITEMS>
struct SortSmem {
  using P = KeyIdx<K>;
  static constexpr int cmax(int a, int b) { return a > b ? a : b; }
  static constexpr int PBYTES = lpad_elems<P>(BLOCK * ITEMS) * int(sizeof(P));
  static constexpr int KBYTES = MergeLds<K, BLOCK, ITEMS>::ELEMS * int(sizeof(K));
  static constexpr int VBYTES = HAS_V ? MergeLds<V, BLOCK, ITEMS>::ELEMS * int(sizeof(V)) : 0;
  __align__(16) char kp[cmax(cmax(PBYTES, KBYTES), VBYTES)];
};

// merge_path over padded (key, index) runs in LDS: A = s[base, base + n1),
// B = s[base + b_off, base + b_off + n2) (logical indices); reads keys only.
template <class CMP, class K>
__device__ __forceinline__ int pair_merge_path(const KeyIdx<K>* s, int base, int n1, int b_off,
                                               int n2, int diag) {
  using P = KeyIdx<K>;
  int lo = diag > n2 ? diag - n2 : 0;
  int hi = diag < n1 ? diag : n1;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (CMP::apply(s[lpad<P>(base + b_off + diag - 1 - m)].k, s[lpad<P>(base + m)].k)) hi = m;
    else lo = m + 1;
  }
  return lo;
}

// lds_serial_merge over (key, index) runs: the index travels with its key,
// so no gather by LDS position follows the merge.
template <class CMP, class K, int ITEMS, int LAST>
__device__ __forceinline__ void pair_serial_merge(const KeyIdx<K>* s, int i1, int e1, int i2,
                                                  int e2, K (&rk)[ITEMS], int (&ridx)[ITEMS]) {
  using P = KeyIdx<K>;
  P pa = s[lpad<P>(i1 < LAST ? i1 : LAST)];
  P pb = s[lpad<P>(i2 < LAST ? i2 : LAST)];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const bool has1 = i1 < e1, has2 = i2 < e2;
    const bool p = !has2 || (has1 && !CMP::apply(pb.k, pa.k));
    rk[i] = p ? pa.k : pb.k;
    ridx[i] = int(p ? pa.i : pb.i);
    if (i + 1 < ITEMS) {
      const int nx = (p ? i1 : i2) + 1;
      const P pn = s[lpad<P>(nx < LAST ? nx : LAST)];
      if (p) {
        i1 = nx;
        pa = pn;
      } else {
        i2 = nx;
        pb = pn;
      }
    }
  }
}

// Block sort of tile `tile` of keys/vals[0, n) into kout/vout (may alias).
template <class CMP, class K, class V, bool HAS_V, int BLOCK, int ITEMS>
__device__ __forceinline__ void block_sort_tile(SortSmem<K, V, HAS_V, BLOCK, ITEMS>& sm,
                                                const K* keys, const V* vals, int64_t n, K* kout,
                                                V* vout, int64_t tile) {
  constexpr int TILE = BLOCK * ITEMS;
  using WK = WaveTile<K, ITEMS>;
  using WV = WaveTile<V, ITEMS>;
  using P = KeyIdx<K>;
  K* s_k = reinterpret_cast<K*>(sm.kp);  // key transposes of the load and the store
  P* s_p = reinterpret_cast<P*>(sm.kp);  // the merge rounds
  V* s_v = reinterpret_cast<V*>(sm.kp);  // value gather and transposes

  const int64_t base = tile * TILE;
  const int cnt = int(n - base < TILE ? n - base : TILE);
  const int mine = threadIdx.x * ITEMS;
  int vt = cnt - mine;  // valid items of this thread
  vt = vt < 0 ? 0 : (vt > ITEMS ? ITEMS : vt);
  const int wid = threadIdx.x / kWave;
  const int wb = wid * WK::N;
  const int wcount = cnt - wb <= 0 ? 0 : (cnt - wb >= WK::N ? WK::N : cnt - wb);

  K k[ITEMS];
  int idx[ITEMS];
  // Striped value loads first, so they are in flight with the keys; the
  // values stay in registers until the gather.
  V lv[ITEMS];
  if constexpr (HAS_V) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int j = threadIdx.x + i * BLOCK;
      if (j < cnt) lv[i] = vals[base + j];
    }
  }
  // Coalesced key load through the wave's transpose region (inside kp, which
  // the merge rounds only reuse after their first barrier).
  wave_load_blocked<K, ITEMS>(keys + base + (wcount > 0 ? wb : 0), wcount, k, K(),
                              s_k + wid * WK::LDS_ELEMS);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) idx[i] = mine + i;

  // Odd-even transposition sort of the grain; items at or past vt never move.
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
#pragma unroll
    for (int j = r & 1; j < ITEMS - 1; j += 2) {
      if (j + 1 < vt && CMP::apply(k[j + 1], k[j])) {
        K tk = k[j];
        k[j] = k[j + 1];
        k[j + 1] = tk;
        int ti = idx[j];
        idx[j] = idx[j + 1];
        idx[j + 1] = ti;
      }
    }
  }

  // Merge rounds inside the tile: runs of w items -> runs of 2w.
  for (int w = ITEMS; w < TILE; w *= 2) {
    __syncthreads();  // previous round's LDS reads are done
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if (i < vt) s_p[lpad<P>(mine + i)] = P{k[i], uint32_t(idx[i])};
    }
    __syncthreads();
    if (vt > 0) {
      const int ps = (mine / (2 * w)) * (2 * w);
      int na = cnt - ps;
      na = na < 0 ? 0 : (na > w ? w : na);
      int nb = cnt - ps - w;
      nb = nb < 0 ? 0 : (nb > w ? w : nb);
      const int diag = mine - ps;
      const int ai = pair_merge_path<CMP>(s_p, ps, na, w, nb, diag);
      const int bi = diag - ai;
      K rk[ITEMS];
      int ri[ITEMS];
      pair_serial_merge<CMP, K, ITEMS, TILE - 1>(s_p, ps + ai, ps + na, ps + w + bi, ps + w + nb,
                                                 rk, ri);
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        if (i < vt) {
          k[i] = rk[i];
          idx[i] = ri[i];
        }
      }
    }
  }

  V rv[ITEMS];
  if constexpr (HAS_V) {
    __syncthreads();  // the last round's reads of the stage are done
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int j = threadIdx.x + i * BLOCK;
      if (j < cnt) s_v[lpad<V>(j)] = lv[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) rv[i] = i < vt ? s_v[lpad<V>(idx[i])] : V();
  }
  __syncthreads();  // all LDS reads done; reuse the stage as transpose regions
  if (wcount > 0)
    wave_store_blocked<K, ITEMS>(kout + base + wb, wcount, k, s_k + wid * WK::LDS_ELEMS);
  if constexpr (HAS_V) {
    // a wave's value region is its key region when K and V have one size;
    // otherwise it overlaps other waves' key regions
    if constexpr (sizeof(K) == sizeof(V)) wave_lds_sync();
    else __syncthreads();
    if (wcount > 0)
      wave_store_blocked<V, ITEMS>(vout + base + wb, wcount, rv, s_v + wid * WV::LDS_ELEMS);
  }
}

template <class CMP, class K, class V, bool HAS_V, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void block_sort_kernel(const K* keys, const V* vals, int64_t n,
                                                           K* kout, V* vout) {
  __shared__ SortSmem<K, V, HAS_V, BLOCK, ITEMS> sm;
  block_sort_tile<CMP, K, V, HAS_V, BLOCK, ITEMS>(sm, keys, vals, n, kout, vout, blockIdx.x);
}

// Geometry of output tile `tile` of a merge pass over runs of `width`: the
// pair of runs it belongs to starts at ps, its runs hold na and nb elements,
// and the tile covers diagonals [d0, d1) of the pair.
struct PassTile {
  int64_t ps, na, nb, d0, d1;
};

template <int TILE>
__device__ __forceinline__ PassTile pass_tile(int64_t n, int64_t width, int64_t tile) {
  PassTile g;
  const int64_t out0 = tile * TILE;
  g.ps = (out0 / (2 * width)) * (2 * width);
  g.na = n - g.ps;
  g.na = g.na < 0 ? 0 : (g.na > width ? width : g.na);
  g.nb = n - g.ps - width;
  g.nb = g.nb < 0 ? 0 : (g.nb > width ? width : g.nb);
  g.d0 = out0 - g.ps;
  g.d1 = g.d0 + TILE < g.na + g.nb ? g.d0 + TILE : g.na + g.nb;
  return g;
}

// Merge path of a tile's first diagonal in a merge pass (the reference's
// locate_merge_path, merge_sort_by_key.cu:90-131), searched by one lane.
template <class CMP, class K, int TILE>
__device__ __forceinline__ int64_t pass_split(const K* src, int64_t n, int64_t width,
                                              int64_t tile) {
  const PassTile g = pass_tile<TILE>(n, width, tile);
  return lane_merge_path<CMP>(src + g.ps, g.na, src + g.ps + width, g.nb, g.d0);
}

// One tile of a merge pass: pairs of sorted runs of width `width` (a
// multiple of TILE) from (src_k, src_v) into (dst_k, dst_v).  splits[t] is
// the merge path of tile t's first diagonal (pass_split); the tile's last
// diagonal is the pair's end or the next tile's first.
template <class CMP, class K, class V, bool HAS_V, int BLOCK, int ITEMS>
__device__ __forceinline__ void merge_pass_tile(MergeSmem<K, V, HAS_V, BLOCK, ITEMS>& sm,
                                                const K* src_k, const V* src_v, int64_t n,
                                                int64_t width, K* dst_k, V* dst_v, int64_t tile,
                                                const int64_t* splits) {
  constexpr int TILE = BLOCK * ITEMS;
  const PassTile g = pass_tile<TILE>(n, width, tile);
  const int64_t a0 = splits[tile];
  const int64_t a1 = g.d1 == g.na + g.nb ? g.na : splits[tile + 1];
  const int64_t b0 = g.d0 - a0, b1 = g.d1 - a1;
  const int64_t pb = g.ps + width;
  merge_tile<CMP, K, V, HAS_V, BLOCK, ITEMS>(
      src_k + g.ps + a0, HAS_V ? src_v + g.ps + a0 : nullptr, int(a1 - a0), src_k + pb + b0,
      HAS_V ? src_v + pb + b0 : nullptr, int(b1 - b0), dst_k + g.ps + g.d0,
      HAS_V ? dst_v + g.ps + g.d0 : nullptr, sm.k, sm.v);
}

// One lane per tile: splits[t] for every tile of a merge pass.
template <class CMP, class K, int TILE>
__global__ __launch_bounds__(256) void merge_partition_kernel(const K* __restrict__ src,
                                                              int64_t n, int64_t width,
                                                              int64_t tiles,
                                                              int64_t* __restrict__ splits) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= tiles) return;
  splits[t] = pass_split<CMP, K, TILE>(src, n, width, t);
}

template <class CMP, class K, class V, bool HAS_V, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void merge_pass_kernel(const K* __restrict__ src_k,
                                                           const V* __restrict__ src_v, int64_t n,
                                                           int64_t width, K* __restrict__ dst_k,
                                                           V* __restrict__ dst_v,
                                                           const int64_t* __restrict__ splits) {
  __shared__ MergeSmem<K, V, HAS_V, BLOCK, ITEMS> sm;
  merge_pass_tile<CMP, K, V, HAS_V, BLOCK, ITEMS>(sm, src_k, src_v, n, width, dst_k, dst_v,
                                                  blockIdx.x, splits);
}

inline int sort_num_passes(int64_t tiles) {
  int p = 0;
  while ((int64_t(1) << p) < tiles) ++p;
  return p;
}

// Workspace of a sort: ping-pong keys and values, then the pass splits.
struct SortWorkspace {
  size_t keys, vals, splits, total;
  SortWorkspace(size_t ks, size_t vs, int64_t n, int64_t tiles) {
    keys = 0;
    vals = align_up(size_t(n) * ks, 256);
    splits = vals + align_up(size_t(n) * vs, 256);
    total = splits + size_t(tiles + 1) * sizeof(int64_t);
  }
};

template <class CMP, class K, class V, bool HAS_V, int VAR>
int sort_launch(K* keys, V* vals, int64_t n, char* ws, hipStream_t stream) {
  using S = SortShape<K, V, VAR>;
  using C = typename S::Pass;
  const int64_t tiles = (n + C::TILE - 1) / C::TILE;   // merge-pass tiles
  const int64_t runs = (n + S::RUN0 - 1) / S::RUN0;    // block-sorted runs
  const int passes = sort_num_passes(runs);
  const SortWorkspace w(sizeof(K), HAS_V ? sizeof(V) : 0, n, tiles);
  K* tk = reinterpret_cast<K*>(ws + w.keys);
  V* tv = HAS_V ? reinterpret_cast<V*>(ws + w.vals) : nullptr;
  int64_t* splits = reinterpret_cast<int64_t*>(ws + w.splits);
  // An odd pass count starts from the workspace so the last pass ends in place.
  K* ck = (passes & 1) ? tk : keys;
  V* cv = (passes & 1) ? tv : vals;
  K* ok = (passes & 1) ? keys : tk;
  V* ov = (passes & 1) ? vals : tv;
  block_sort_kernel<CMP, K, V, HAS_V, S::BBLOCK, S::BITEMS>
      <<<dim3(uint32_t(runs)), S::BBLOCK, 0, stream>>>(keys, vals, n, ck, cv);
  BULKX_LAUNCHED();
  int64_t width = S::RUN0;
  const uint32_t pgrid = uint32_t((tiles + 255) / 256);
  for (int p = 0; p < passes; ++p, width *= 2) {
    merge_partition_kernel<CMP, K, C::TILE><<<dim3(pgrid), 256, 0, stream>>>(ck, n, width, tiles,
                                                                             splits);
    BULKX_LAUNCHED();
    merge_pass_kernel<CMP, K, V, HAS_V, C::BLOCK, C::ITEMS>
        <<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(ck, cv, n, width, ok, ov, splits);
    BULKX_LAUNCHED();
    K* t1 = ck;
    ck = ok;
    ok = t1;
    V* t2 = cv;
    cv = ov;
    ov = t2;
  }
  return BULKX_SUCCESS;
}

template <class K, int VAR>
int sort_dispatch_variant(int vdtype, int comp, void* keys, void* vals, int64_t n, char* ws,
                          hipStream_t stream) {
  int st = BULKX_SUCCESS;
  BULKX_DISPATCH_CMP(comp, CMP, {
    if (vdtype < 0)
      st = sort_launch<CMP, K, uint32_t, false, VAR>(static_cast<K*>(keys), nullptr, n, ws, stream);
    else if (dtype_size(vdtype) == 4)
      st = sort_launch<CMP, K, uint32_t, true, VAR>(static_cast<K*>(keys),
                                                    static_cast<uint32_t*>(vals), n, ws, stream);
    else
      st = sort_launch<CMP, K, uint64_t, true, VAR>(static_cast<K*>(keys),
                                                    static_cast<uint64_t*>(vals), n, ws, stream);
  });
  return st;
}

// The radix form over 4-byte keys; one tile is sorted by the block sort
// alone (variant 2 writes 8192-element runs), and n past 32-bit positions
// takes the merge form.
template <class K>
int sort_dispatch_radix(int vdtype, int comp, void* keys, void* vals, int64_t n, char* ws,
                        hipStream_t stream) {
  if (n <= kRadixTile) return sort_dispatch_variant<K, 2>(vdtype, comp, keys, vals, n, ws, stream);
  if (n > kRadixMaxN) return sort_dispatch_variant<K, 3>(vdtype, comp, keys, vals, n, ws, stream);
  int st = BULKX_SUCCESS;
  BULKX_DISPATCH_CMP(comp, CMP, {
    K* k = static_cast<K*>(keys);
    if (vdtype < 0)
      st = radix_sort_launch<CMP, K, uint32_t, false>(k, nullptr, n, ws, stream);
    else if (dtype_size(vdtype) == 4)
      st = radix_sort_launch<CMP, K, uint32_t, true>(k, static_cast<uint32_t*>(vals), n, ws, stream);
    else
      st = radix_sort_launch<CMP, K, uint64_t, true>(k, static_cast<uint64_t*>(vals), n, ws, stream);
  });
  return st;
}

template <class K>
int sort_dispatch_key(int vdtype, int comp, void* keys, void* vals, int64_t n, char* ws,
                      hipStream_t stream, int variant) {
  if constexpr (sizeof(K) == 4) {
    if (variant == kSortRadixVariant)
      return sort_dispatch_radix<K>(vdtype, comp, keys, vals, n, ws, stream);
  }
  switch (variant) {
    case 1: return sort_dispatch_variant<K, 1>(vdtype, comp, keys, vals, n, ws, stream);
    case 2: return sort_dispatch_variant<K, 2>(vdtype, comp, keys, vals, n, ws, stream);
    case 3: return sort_dispatch_variant<K, 3>(vdtype, comp, keys, vals, n, ws, stream);
    default: return sort_dispatch_variant<K, 0>(vdtype, comp, keys, vals, n, ws, stream);
  }
}

int sort_dispatch_k4(int kdtype, int vdtype, int comp, void* keys, void* vals, int64_t n,
                     char* ws, hipStream_t stream, int variant);
int sort_dispatch_k8(int kdtype, int vdtype, int comp, void* keys, void* vals, int64_t n,
                     char* ws, hipStream_t stream, int variant);

}  // namespace bulkx
