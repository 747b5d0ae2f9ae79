// Fused single-pass reduce_by_key: head flags, segmented scan and
// compaction in one kernel with decoupled look-back.
//
// Replaces my_reduce_by_key (reduce_by_key.cu:176-272), which counts tail
// flags per interval (reduce_intervals, reduce_by_key.cu:240), scans the
// counts (:243), runs bulk::reduce_by_key per interval
// (reduce_by_key_kernel :31-94 -> reduce_by_key.hpp:115-216), then fixes up
// the carries with inclusive_scan_by_key + transform_if (:255-267) and reads
// the count back (:269): five launches and three reads of the keys.  Here
// each key and value is read once and each run written once.
//
// The scan operator is the reference's scan_head_flags_functor
// (reduce_by_key.hpp:41-62) on (count, value) pairs:
//   (c1, v1) . (c2, v2) = (c1 + c2, c2 ? v2 : v1 op v2),
// where c counts run heads.  A run's output slot is (heads up to and
// including its head) - 1 and its key is the run's last key, as the
// reference's scatter_tails_n writes it (reduce_by_key.hpp:65-108).
#pragma once

#include "device_common.hpp"
#include "host_common.hpp"

namespace bulkx {

template <class A>
struct Seg {
  uint64_t c;
  A v;
};

template <class AOp, class A>
struct SegOp {
  __device__ __forceinline__ static Seg<A> apply(Seg<A> a, Seg<A> b) {
    Seg<A> r;
    r.c = a.c + b.c;
    r.v = b.c ? b.v : AOp::apply(a.v, b.v);
    return r;
  }
  template <class X>
  __device__ __forceinline__ static Seg<A> identity() {
    Seg<A> r;
    r.c = 0;
    r.v = AOp::template identity<A>();
    return r;
  }
};

template <class K, class V>
struct RbkConfig {
  static constexpr int BLOCK = 256;
  static constexpr int ITEMS = (sizeof(K) + sizeof(V) <= 8) ? 16 : 8;
  static constexpr int TILE = BLOCK * ITEMS;
};

constexpr size_t kRbkHeaderBytes = 256;
constexpr int kRbkMaxStatusWords = 6;  // Seg<dd> = 24 bytes

template <class K, class V, class Op, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void reduce_by_key_kernel(
    const K* __restrict__ keys, const V* __restrict__ vals, int64_t n, K* __restrict__ kout,
    V* __restrict__ vout, int64_t* __restrict__ num_runs, uint32_t* header, uint64_t* words,
    bool single_tile) {
  using AccP = Acc<V, Op>;
  using A = typename AccP::type;
  using AOp = typename AccP::op;
  using S = Seg<A>;
  using SOp = SegOp<AOp, A>;
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int NW = BLOCK / kWave;
  using WK = WaveTile<K, ITEMS>;
  using WV = WaveTile<V, ITEMS>;
  static_assert(NW * WK::LDS_ELEMS >= TILE && NW * WV::LDS_ELEMS >= TILE, "staging fits");

  __shared__ S s_wave[NW];
  __shared__ S s_prefix;
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_count;
  __shared__ uint32_t s_first_head;
  // Per-wave transpose regions; after the block scan they are reused as the
  // tile's compaction stage (s_k / s_v).
  __shared__ K s_trk[NW][WK::LDS_ELEMS];
  __shared__ V s_trv[NW][WV::LDS_ELEMS];
  K* s_k = &s_trk[0][0];
  V* s_v = &s_trv[0][0];
  const TileStatus<S> st{words};

  int64_t tile;
  if (single_tile) {
    tile = 0;
  } else {
    if (threadIdx.x == 0) s_tile = atomicAdd(header, 1u);
    __syncthreads();
    tile = s_tile;
  }
  const int wid = threadIdx.x / kWave;
  const int lane = lane_id();
  const int64_t chunk_base = tile * TILE + int64_t(wid) * WK::N;
  const int64_t rem = n - chunk_base;
  const int count = rem <= 0 ? 0 : (rem >= WK::N ? WK::N : int(rem));
  const int64_t my_base = chunk_base + int64_t(lane) * ITEMS;
  const int64_t valid = count - lane * ITEMS;

  K k[ITEMS];
  V v[ITEMS];
  const int64_t cb = count > 0 ? chunk_base : 0;
  wave_load_blocked<K, ITEMS>(keys + cb, count, k, K(0), s_trk[wid]);
  wave_load_blocked<V, ITEMS>(vals + cb, count, v, V(0), s_trv[wid]);
  // Neighbouring keys: inside the wave by DPP lane shifts, across wave
  // chunks from global memory (one load per wave edge).
  K gprev = K(0), gnext = K(0);
  if (lane == 0 && count > 0 && chunk_base > 0) gprev = keys[chunk_base - 1];
  if (lane == kWave - 1 && chunk_base + WK::N < n) gnext = keys[chunk_base + WK::N];
  const K kprev = dpp_move<kWaveShr1>(gprev, k[ITEMS - 1]);
  const K knext = dpp_move<kWaveShl1>(gnext, k[0]);

  // Head / tail flags of the owned elements (head_flags.hpp, tail_flags.hpp).
  uint32_t heads = 0, tails = 0;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (i < valid) {
      const int64_t g = my_base + i;
      const K before = i ? k[i - 1] : kprev;
      const K after = (i + 1 < ITEMS) ? k[i + 1] : knext;
      if (g == 0 || !(k[i] == before)) heads |= 1u << i;
      if (g == n - 1 || !(k[i] == after)) tails |= 1u << i;
    }
  }
  if (threadIdx.x == 0) {
    s_count = 0;
    s_first_head = heads & 1u;
  }

  // Thread aggregate in the (count, value) monoid.
  S agg = SOp::template identity<S>();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (i < valid) {
      const A x = AccP::from(v[i]);
      if (heads & (1u << i)) {
        agg.c += 1;
        agg.v = x;
      } else {
        agg.v = AOp::apply(agg.v, x);
      }
    }
  }

  S tile_agg;
  const S thread_prefix = block_exclusive_scan<BLOCK, SOp>(agg, s_wave, &tile_agg);

  S tile_prefix = SOp::template identity<S>();
  if (!single_tile) {
    if (tile == 0) {
      if (threadIdx.x == 0) st.publish(0, kStatusPrefix, tile_agg);
    } else {
      if (threadIdx.x == 0) st.publish(tile, kStatusAggregate, tile_agg);
      if (threadIdx.x < kWave) {
        S excl = lookback_exclusive_prefix<SOp>(st, tile, header + 1);
        if (threadIdx.x == 0) {
          st.publish(tile, kStatusPrefix, SOp::apply(excl, tile_agg));
          s_prefix = excl;
        }
      }
      __syncthreads();
      tile_prefix = s_prefix;
    }
  }

  // Output slots of this tile start at `base`: the run open at the tile's
  // first element (its head may lie in an earlier tile) gets slot base.
  const uint64_t base = tile_prefix.c - (s_first_head ? 0 : 1);
  S p = SOp::apply(tile_prefix, thread_prefix);
  uint64_t c = p.c;
  A s = p.v;
  uint32_t my_count = 0;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (i < valid) {
      const A x = AccP::from(v[i]);
      if (heads & (1u << i)) {
        c += 1;
        s = x;
      } else {
        s = AOp::apply(s, x);
      }
      if (tails & (1u << i)) {
        const uint32_t slot = uint32_t(c - 1 - base);
        if (slot < uint32_t(TILE)) {
          s_k[slot] = k[i];
          s_v[slot] = AccP::to(s);
          my_count = slot + 1;
        }
      }
    }
  }
  if (my_count) atomicMax(&s_count, my_count);
  if (my_base < n && n - 1 < my_base + ITEMS) *num_runs = int64_t(c);  // owner of element n-1
  __syncthreads();
  // Bounds guard: a tile never writes past TILE slots nor past n outputs.
  uint64_t cnt = s_count < uint32_t(TILE) ? s_count : uint32_t(TILE);
  if (base >= uint64_t(n)) cnt = 0;
  else if (base + cnt > uint64_t(n)) cnt = uint64_t(n) - base;
  for (uint32_t j = threadIdx.x; j < cnt; j += BLOCK) {
    kout[base + j] = s_k[j];
    vout[base + j] = s_v[j];
  }
}

__global__ void rbk_empty_kernel(int64_t* num_runs);
__global__ void rbk_prepare_kernel(uint32_t* header, uint64_t* words, int64_t num_words);

template <class K, class V>
inline int64_t rbk_num_tiles(int64_t n) {
  constexpr int TILE = RbkConfig<K, V>::TILE;
  return (n + TILE - 1) / TILE;
}

template <class K, class V, class Op>
int rbk_launch(const K* keys, const V* vals, int64_t n, K* kout, V* vout, int64_t* num_runs,
               char* ws, hipStream_t stream) {
  using C = RbkConfig<K, V>;
  using S = Seg<typename Acc<V, Op>::type>;
  if (n == 0) {
    rbk_empty_kernel<<<1, 64, 0, stream>>>(num_runs);
    BULKX_LAUNCHED();
    return BULKX_SUCCESS;
  }
  const int64_t tiles = rbk_num_tiles<K, V>(n);
  if (tiles == 1) {
    reduce_by_key_kernel<K, V, Op, C::BLOCK, C::ITEMS><<<1, C::BLOCK, 0, stream>>>(
        keys, vals, n, kout, vout, num_runs, nullptr, nullptr, true);
    BULKX_LAUNCHED();
    return BULKX_SUCCESS;
  }
  uint32_t* header = reinterpret_cast<uint32_t*>(ws);
  uint64_t* words = reinterpret_cast<uint64_t*>(ws + kRbkHeaderBytes);
  const int64_t num_words = tiles * TileStatus<S>::W;
  const int64_t pb = (num_words + 255) / 256;
  rbk_prepare_kernel<<<int(pb < 1024 ? pb : 1024), 256, 0, stream>>>(header, words, num_words);
  BULKX_LAUNCHED();
  reduce_by_key_kernel<K, V, Op, C::BLOCK, C::ITEMS><<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(
      keys, vals, n, kout, vout, num_runs, header, words, false);
  BULKX_LAUNCHED();
  return BULKX_SUCCESS;
}

// Dispatch over value type and op for one key type K.
template <class K>
int rbk_dispatch_values(int vdtype, int op, const void* keys, const void* vals, int64_t n,
                        void* kout, void* vout, int64_t* num_runs, char* ws, hipStream_t stream) {
  int st = BULKX_SUCCESS;
  BULKX_DISPATCH_DTYPE(vdtype, V, BULKX_DISPATCH_OP(op, OP, {
    st = rbk_launch<K, V, OP>(static_cast<const K*>(keys), static_cast<const V*>(vals), n,
                              static_cast<K*>(kout), static_cast<V*>(vout), num_runs, ws, stream);
  }));
  return st;
}

int rbk_dispatch_k4(int kdtype, int vdtype, int op, const void* keys, const void* vals, int64_t n,
                    void* kout, void* vout, int64_t* num_runs, char* ws, hipStream_t stream);
int rbk_dispatch_k8(int kdtype, int vdtype, int op, const void* keys, const void* vals, int64_t n,
                    void* kout, void* vout, int64_t* num_runs, char* ws, hipStream_t stream);

}  // namespace bulkx
