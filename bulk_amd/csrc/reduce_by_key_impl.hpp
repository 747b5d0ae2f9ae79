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

// Row-level pair for the short wavefront scans: 32-bit head count (a row
// holds at most 64*VE elements) and the value in its own type.  It is
// lifted to Seg<A> (64-bit count, carry accumulator) for row, wave, tile and
// look-back carries.
template <class V>
struct RowSeg {
  uint32_t c;
  V v;
};

template <class Op, class V>
struct RowSegOp {
  __device__ __forceinline__ static RowSeg<V> apply(RowSeg<V> a, RowSeg<V> b) {
    RowSeg<V> r;
    r.c = a.c + b.c;
    r.v = b.c ? b.v : Op::apply(a.v, b.v);
    return r;
  }
  template <class X>
  __device__ __forceinline__ static RowSeg<V> identity() {
    RowSeg<V> r;
    r.c = 0;
    r.v = Op::template identity<V>();
    return r;
  }
};

// Tile shapes: variant 0 = 512 x 16 (default), 1 = 512 x 8, 2 = 256 x 16,
// 3 = 1024 x 16, 4 = 1024 x 8, 5 = 256 x 32, all single-pass with
// decoupled look-back; variants 6, 7 and 8 are the three-pass form
// (rbk_upsweep_kernel) of shapes 0, 2 and 5.
template <class K, class V, int VAR = 0>
struct RbkConfig {
  static constexpr int BLOCK = (VAR == 2 || VAR == 5) ? 256 : (VAR == 3 || VAR == 4) ? 1024 : 512;
  static constexpr int ITEMS = (VAR == 1 || VAR == 4) ? 8 : VAR == 5 ? 32 : 16;
  static constexpr int TILE = BLOCK * ITEMS;
};
constexpr int kRbkShapes = 6;
constexpr int kRbkVariants = 12;
// tile shape of a variant: 0-5 single pass, 6-8 three-pass and 9-11 compact
// forms of shapes 0, 2 and 5
constexpr int rbk_shape(int variant) {
  return variant == 6 || variant == 9 ? 0 : variant == 7 || variant == 10 ? 2
         : variant == 8 || variant == 11 ? 5 : variant;
}
enum RbkForm { kRbkSinglePass = 0, kRbkThreePass = 1, kRbkCompact = 2 };
constexpr int kRbkCompactMaxTile = 8192;  // largest tile of a compact variant
constexpr int kRbkMinTile = 4096;  // smallest tile of any variant (workspace sizing)

// Look-back window of 64*Q tiles (Q tiles per lane).  Q = 1 measured
// fastest on MI355X at 2^28 elements (Q = 2, 4, 8: 8 %, 90 %, 130 % slower).
#ifndef BULKX_RBK_LOOKBACK_Q
#define BULKX_RBK_LOOKBACK_Q 1
#endif
constexpr int kRbkLookbackQ = BULKX_RBK_LOOKBACK_Q;
constexpr size_t kRbkHeaderBytes = 256;
constexpr int kRbkMaxStatusWords = 6;  // Seg<dd> = 24 bytes

// Per-wave LDS region, used in turn for the padded transpose of the keys,
// that of the values, the stage of the wave's output keys and that of its
// output values: one region instead of two halves the LDS of a workgroup,
// so twice as many workgroups (and their loads) fit on a CU.
template <class K, class V, int ITEMS>
struct RbkLds {
  static constexpr int KBYTES = WaveTile<K, ITEMS>::LDS_ELEMS * int(sizeof(K));
  static constexpr int VBYTES = WaveTile<V, ITEMS>::LDS_ELEMS * int(sizeof(V));
  static constexpr int BYTES = ((KBYTES > VBYTES ? KBYTES : VBYTES) + 15) / 16 * 16;
  static_assert(kWave * ITEMS * int(sizeof(K)) <= BYTES, "key stage fits");
  static_assert(kWave * ITEMS * int(sizeof(V)) <= BYTES, "value stage fits");
};

// Layout: a wavefront owns a contiguous chunk of 64*ITEMS elements, loaded
// with coalesced 16-byte accesses and transposed through LDS so that lane l
// holds the ITEMS consecutive elements [l*ITEMS, (l+1)*ITEMS).  Flags, the
// lane's run aggregate and the final walk are sequential per lane (the
// bounded-agent body of reduce_by_key.hpp:159-208); only one DPP scan per
// wave chunk crosses lanes.  The chunk's tails occupy consecutive output
// slots, so they are staged in the wave's LDS region and written with
// coalesced stores.
template <class K, class V, class Op, int BLOCK, int ITEMS>
struct RbkSmem {
  using S = Seg<typename Acc<V, Op>::type>;
  S wave[BLOCK / kWave];
  S prefix;
  uint32_t ntails[BLOCK / kWave];  // compact form: run ends per wavefront
  __align__(16) char region[BLOCK / kWave][RbkLds<K, V, ITEMS>::BYTES];
};

// Aggregates scanned per workgroup of the three-pass form's pass 2: 256
// threads x 8 (x 4 for double-double) -- all of a thread's loads in flight at
// once, no spills.
template <class S>
constexpr int rbk_scan_items() {
  return sizeof(S) <= 16 ? 8 : 4;
}
template <class S>
constexpr int rbk_scan_seg() {
  return 256 * rbk_scan_items<S>();
}

// One wavefront's chunk of a tile while its loads are in flight: the keys
// and values as coalesced 16-byte vectors (full aligned chunks; a partial
// chunk is read element by element by rbk_lane_pass), plus the keys just
// before and after the chunk.
template <class K, class V, int ITEMS>
struct RbkChunk {
  u32x4 tk[WaveTile<K, ITEMS>::NV];
  u32x4 tv[WaveTile<V, ITEMS>::NV];
  K gprev, gnext;
  int64_t base;  // index of the chunk's first element
  int count;     // elements of the chunk inside [0, n)
};

// Issues the key loads of the chunk at `chunk_base` (no wait).
template <class K, class V, int ITEMS>
__device__ __forceinline__ void rbk_load_keys(const K* keys, int64_t n, int64_t chunk_base,
                                              RbkChunk<K, V, ITEMS>& c, int lane) {
  using WK = WaveTile<K, ITEMS>;
  constexpr int N = WK::N;
  const int64_t rem = n - chunk_base;
  c.base = chunk_base;
  c.count = rem <= 0 ? 0 : (rem >= N ? N : int(rem));
  const int64_t cb = c.count > 0 ? chunk_base : 0;
  if (c.count >= N && aligned16(keys + cb)) {
    const u32x4* q = reinterpret_cast<const u32x4*>(keys + cb);
#pragma unroll
    for (int j = 0; j < WK::NV; ++j) c.tk[j] = __builtin_nontemporal_load(q + j * kWave + lane);
  }
  c.gprev = K(0);
  c.gnext = K(0);
  if (lane == 0 && c.count > 0 && chunk_base > 0) c.gprev = keys[chunk_base - 1];
  if (lane == kWave - 1 && c.count >= N && chunk_base + N < n) c.gnext = keys[chunk_base + N];
}

// Issues the value loads of the chunk whose keys rbk_load_keys requested.
template <class K, class V, int ITEMS>
__device__ __forceinline__ void rbk_load_values(const V* vals, RbkChunk<K, V, ITEMS>& c, int lane) {
  using WV = WaveTile<V, ITEMS>;
  const int64_t cb = c.count > 0 ? c.base : 0;
  if (c.count >= WV::N && aligned16(vals + cb)) {
    const u32x4* q = reinterpret_cast<const u32x4*>(vals + cb);
#pragma unroll
    for (int j = 0; j < WV::NV; ++j) c.tv[j] = __builtin_nontemporal_load(q + j * kWave + lane);
  }
}

// The lane's part of a chunk after its loads: keys and values transposed
// through the wave's LDS region to ITEMS consecutive elements per lane,
// head flags h and tail flags t (bit i: element i), each v[i] replaced by
// the fold since its run head (for a tail, the run's output value once the
// lane's prefix is folded in where no head precedes it in the lane), and the
// lane aggregate (heads, value since the last head).  Values past the end
// of the array only extend the last run's aggregate, which nothing reads.
template <class K, class V, class Op, int ITEMS>
struct RbkLane {
  K k[ITEMS];
  V v[ITEMS];
  uint32_t h, t;
  RowSeg<V> a;
};

template <class K, class V, class Op, int ITEMS>
__device__ __forceinline__ void rbk_lane_pass(const K* keys, const V* vals, int64_t n,
                                              const RbkChunk<K, V, ITEMS>& c, char* region,
                                              RbkLane<K, V, Op, ITEMS>& L, int lane) {
  using WK = WaveTile<K, ITEMS>;
  using WV = WaveTile<V, ITEMS>;
  constexpr int N = WK::N;
  const int count = c.count;
  const int64_t cb = count > 0 ? c.base : 0;
  K* lk = reinterpret_cast<K*>(region);
  V* lv = reinterpret_cast<V*>(region);
  if (count >= N && aligned16(keys + cb)) {
#pragma unroll
    for (int j = 0; j < WK::NV; ++j) {
      const K* e = reinterpret_cast<const K*>(&c.tk[j]);
      const int base = WK::pos((j * kWave + lane) * WK::VE);
#pragma unroll
      for (int q = 0; q < WK::VE; ++q) lk[base + q] = e[q];
    }
  } else {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const int e = j * kWave + lane;
      lk[WK::pos(e)] = e < count ? keys[cb + e] : K(0);
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) L.k[i] = lk[WK::pos(lane * ITEMS + i)];
  const K kprev = lane > 0 ? lk[WK::pos(lane * ITEMS - 1)] : c.gprev;
  const K knext = lane < kWave - 1 ? lk[WK::pos(lane * ITEMS + ITEMS)] : c.gnext;
  wave_lds_sync();  // key reads done before the values reuse the region
  if (count >= N && aligned16(vals + cb)) {
#pragma unroll
    for (int j = 0; j < WV::NV; ++j) {
      const V* e = reinterpret_cast<const V*>(&c.tv[j]);
      const int base = WV::pos((j * kWave + lane) * WV::VE);
#pragma unroll
      for (int q = 0; q < WV::VE; ++q) lv[base + q] = e[q];
    }
  } else {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const int e = j * kWave + lane;
      lv[WV::pos(e)] = e < count ? vals[cb + e] : V(0);
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) L.v[i] = lv[WV::pos(lane * ITEMS + i)];
  wave_lds_sync();  // the region is free again (used as the output stages)

  // Head / tail flags (head_flags.hpp, tail_flags.hpp): one key comparison
  // per element; a tail is the element before a head.
  int valid = count - lane * ITEMS;
  valid = valid < 0 ? 0 : (valid > ITEMS ? ITEMS : valid);
  const uint32_t vmask = valid >= 32 ? ~0u : ((1u << valid) - 1);
  uint32_t h = 0;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) h |= uint32_t(!(L.k[i] == (i ? L.k[i - 1] : kprev))) << i;
  uint32_t t = (h >> 1) | (uint32_t(!(L.k[ITEMS - 1] == knext)) << (ITEMS - 1));
  if (c.base == 0 && lane == 0) h |= 1u;  // the array's first element
  // The array's last element ends a run wherever it falls in the lane (the
  // element after it in the lane's slice is fill, which may equal its key).
  if (valid > 0 && c.base + int64_t(lane) * ITEMS + valid == n) t |= 1u << (valid - 1);
  L.h = h & vmask;
  L.t = t & vmask;
  L.a.c = uint32_t(__builtin_popcount(L.h));
  L.a.v = Op::template identity<V>();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    L.a.v = (L.h >> i) & 1u ? L.v[i] : Op::apply(L.a.v, L.v[i]);
    L.v[i] = L.a.v;
  }
}

template <class V, class Op>
__device__ __forceinline__ Seg<typename Acc<V, Op>::type> rbk_lift(RowSeg<V> r) {
  Seg<typename Acc<V, Op>::type> x;
  x.c = r.c;
  x.v = Acc<V, Op>::from(r.v);
  return x;
}

// The walk, first half: a tail's run has slot (heads so far) - 1.  The
// chunk's tails hold consecutive slots from base (the run open at the
// chunk's first element), so slot - base < 64*ITEMS is a 32-bit stage
// index.  Slots are kept relative to the tile (base_rel) until the tile's
// prefix arrives.  This half stages the output keys.  It has no branches
// and no chain: a tail's slot is sbase + (heads up to it in the lane),
// masked into the stage; an element that is no tail writes to a spare slot
// past the N staged ones (one per lane, or lane pair).  The lane's one
// tail that may precede its first head closes the run carried into the
// lane; its value needs the tile's prefix and is patched in the second half.
template <class K, class V, class Op, int ITEMS>
struct RbkWalk {
  Seg<typename Acc<V, Op>::type> q;  // in-tile prefix of the lane
  int64_t base_rel;
  uint32_t ntails, sbase;
  bool pre;
  V pre_local;
};

template <class K, class V, class Op, int ITEMS>
__device__ __forceinline__ RbkWalk<K, V, Op, ITEMS> rbk_stage_keys(
    const RbkLane<K, V, Op, ITEMS>& L, Seg<typename Acc<V, Op>::type> q, char* region, int lane) {
  constexpr int N = WaveTile<K, ITEMS>::N;
  RbkWalk<K, V, Op, ITEMS> w;
  w.q = q;
  w.base_rel = readlane(int64_t(q.c) + int64_t(L.h & 1u) - 1, 0);
  w.ntails = wave_reduce<OpPlus>(uint32_t(__builtin_popcount(L.t)));
  K* sk = reinterpret_cast<K*>(region);
  constexpr uint32_t KSPARE = uint32_t(RbkLds<K, V, ITEMS>::BYTES / int(sizeof(K)) - N);
  static_assert(KSPARE >= 1, "the region has spare slots past the stage");
  static_assert((N & (N - 1)) == 0, "the stage index is masked with N - 1");
  const uint32_t kspare = uint32_t(N) + uint32_t(lane) % KSPARE;
  const uint32_t e0 = uint32_t(int64_t(q.c) - w.base_rel);
  w.sbase = e0 - 1;  // slot of a tail with no head before it in the lane
  w.pre = false;     // the lane has such a tail
  w.pre_local = Op::template identity<V>();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t upto = L.h & (i == 31 ? ~0u : ((2u << i) - 1u));  // heads up to element i
    const uint32_t slot = (uint32_t(__builtin_popcount(upto)) + w.sbase) & uint32_t(N - 1);
    const bool ti = (L.t >> i) & 1u;
    sk[ti ? slot : kspare] = L.k[i];
    const bool p = ti && upto == 0;
    w.pre_local = p ? L.v[i] : w.pre_local;
    w.pre |= p;
  }
  return w;
}

// The walk, second half, once the tile's exclusive prefix is known: the
// staged keys go out, the values are staged (the pre-head tail's value with
// the prefix folded in) and go out, all with coalesced stores.
template <class K, class V, class Op, int ITEMS>
__device__ __forceinline__ void rbk_write_runs(const RbkLane<K, V, Op, ITEMS>& L,
                                               const RbkWalk<K, V, Op, ITEMS>& w,
                                               Seg<typename Acc<V, Op>::type> tile_prefix,
                                               int64_t n, K* kout, V* vout, char* region,
                                               int lane) {
  using AccP = Acc<V, Op>;
  using AOp = typename AccP::op;
  constexpr int N = WaveTile<K, ITEMS>::N;
  K* sk = reinterpret_cast<K*>(region);
  V* sv = reinterpret_cast<V*>(region);
  constexpr uint32_t VSPARE = uint32_t(RbkLds<K, V, ITEMS>::BYTES / int(sizeof(V)) - N);
  static_assert(VSPARE >= 1, "the region has spare slots past the stage");
  const uint32_t vspare = uint32_t(N) + uint32_t(lane) % VSPARE;
  // The pre-head tail's value: the lane's in-tile prefix folded in, and the
  // tile's prefix too when no head precedes the lane in the tile (same
  // association as folding the full prefix first: (tile_prefix . q) . local).
  V pre_val = Op::template identity<V>();
  if (w.pre)
    pre_val = w.q.c > 0
                  ? AccP::to(AOp::apply(w.q.v, AccP::from(w.pre_local)))
                  : AccP::to(AOp::apply(AOp::apply(tile_prefix.v, w.q.v), AccP::from(w.pre_local)));
  const int64_t base = int64_t(tile_prefix.c) + w.base_rel;
  // bounds guard: slots stay below n when the scan is consistent
  uint32_t cnt = w.ntails < uint32_t(N) ? w.ntails : uint32_t(N);
  if (base < 0 || base >= n) cnt = 0;
  else if (base + int64_t(cnt) > n) cnt = uint32_t(n - base);
  for (uint32_t x = lane; x < cnt; x += kWave) kout[base + x] = sk[x];
  wave_lds_sync();  // key stage read before the value stage overwrites it
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t upto = L.h & (i == 31 ? ~0u : ((2u << i) - 1u));
    const uint32_t slot = (uint32_t(__builtin_popcount(upto)) + w.sbase) & uint32_t(N - 1);
    sv[(L.t >> i) & 1u ? slot : vspare] = L.v[i];
  }
  if (w.pre) sv[w.sbase & uint32_t(N - 1)] = pre_val;
  wave_lds_sync();
  for (uint32_t x = lane; x < cnt; x += kWave) vout[base + x] = sv[x];
}

// Compact form, pass 1: where a tile leaves its runs and what it reports.
// Every tile writes the runs that end in it to its own slots of a scratch
// array (slot s of tile t at t * (TILE + 1) + s, the run's tail being the
// s-th head of the tile counting a run carried in as slot 0), its aggregate
// (head count, fold since its last head) and meta = number of runs ending
// in it | (its first element is a head) << 31.
template <class S>
struct RbkCompactOut {
  S* aggs;
  uint32_t* meta;
};

// One tile of reduce_by_key (shared by the device-wide kernels and the
// batch executor).
template <class K, class V, class Op, int BLOCK, int ITEMS>
__device__ __forceinline__ void rbk_tile(RbkSmem<K, V, Op, BLOCK, ITEMS>& sm, const K* keys,
                                         const V* vals, int64_t n, K* kout, V* vout,
                                         int64_t* num_runs, int64_t tile, uint64_t* words,
                                         LookbackGuard* guard, bool single_tile,
                                         const Seg<typename Acc<V, Op>::type>* prefixes = nullptr,
                                         const Seg<typename Acc<V, Op>::type>* group_prefixes = nullptr,
                                         const uint8_t* skip = nullptr, int64_t skip_tiles = 0,
                                         const RbkCompactOut<Seg<typename Acc<V, Op>::type>>* co = nullptr) {
  using AccP = Acc<V, Op>;
  using A = typename AccP::type;
  using S = Seg<A>;
  using SOp = SegOp<typename AccP::op, A>;
  using RS = RowSeg<V>;
  using RSOp = RowSegOp<Op, V>;
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int NW = BLOCK / kWave;
  constexpr int N = WaveTile<K, ITEMS>::N;
  const TileStatus<S> st{words};
  const int wid = threadIdx.x / kWave;
  // Three-pass form: prefix of the tile's scan segment, then the tile's
  // exclusive prefix inside it, read before the tile's loads so that the
  // round trip overlaps them.
  S known = SOp::template identity<S>();
  if (prefixes)
    known = group_prefixes ? SOp::apply(group_prefixes[tile / rbk_scan_seg<S>()], prefixes[tile])
                           : prefixes[tile];
  char* region = sm.region[wid];
  // Coalesced 16-byte loads of keys and values, all in flight before the
  // first LDS write.
  const int lane = lane_id();
  RbkChunk<K, V, ITEMS> c;
  rbk_load_keys(keys, n, tile * TILE + int64_t(wid) * N, c, lane);
  // Three-pass form with `skip` (the upsweep's flags): a tile in which no
  // run ends writes nothing, so it stops here, while its key loads are in
  // flight and before its values are read.
  if (skip != nullptr && tile + 1 < skip_tiles && !(skip[tile] & 2u) && !(skip[tile + 1] & 1u))
    return;
  rbk_load_values(vals, c, lane);
  RbkLane<K, V, Op, ITEMS> L;
  rbk_lane_pass(keys, vals, n, c, region, L, lane);

  // Lane aggregates: DPP scan across the wave, LDS combine across waves.
  const RS inc = wave_inclusive_scan<RSOp>(L.a);
  const RS lane_prefix = wave_shift_right<RSOp>(inc);
  const RS wave_total = readlane(inc, kWave - 1);
  if (lane == 0) sm.wave[wid] = rbk_lift<V, Op>(wave_total);
  __syncthreads();
  S wave_prefix = SOp::template identity<S>(), tile_agg = SOp::template identity<S>();
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const S x = sm.wave[w];
    if (w < wid) wave_prefix = SOp::apply(wave_prefix, x);
    tile_agg = SOp::apply(tile_agg, x);
  }

  // Publish this tile's aggregate (or, for tile 0, its inclusive prefix),
  // then let wavefront 0 look back while the other wavefronts already stage
  // their keys: that only needs the in-tile prefix.  With `prefixes` (the
  // three-pass form) the tile's exclusive prefix is already known: nothing
  // is published and nothing looked back.
  const bool chained = !single_tile && tile > 0 && prefixes == nullptr && co == nullptr;
  if (!single_tile && prefixes == nullptr && co == nullptr && threadIdx.x == 0)
    st.publish(tile, tile == 0 ? kStatusPrefix : kStatusAggregate, tile_agg);
  if (chained && threadIdx.x < kWave) {
    S excl = lookback_exclusive_prefix<SOp, kRbkLookbackQ>(st, tile, guard);
    if (threadIdx.x == 0) {
      st.publish(tile, kStatusPrefix, SOp::apply(excl, tile_agg));
      sm.prefix = excl;
    }
  }
  const RbkWalk<K, V, Op, ITEMS> w =
      rbk_stage_keys(L, SOp::apply(wave_prefix, rbk_lift<V, Op>(lane_prefix)), region, lane);
  if (co && lane == 0) sm.ntails[wid] = w.ntails;
  __syncthreads();  // the tile's prefix is in sm.prefix
  if (co) {
    // slots relative to the tile: a prefix of one head puts the run carried
    // in at slot 0 and the tile's k-th head's run at slot k
    S one = SOp::template identity<S>();
    one.c = 1;
    if (threadIdx.x == 0) {
      uint32_t tails = 0;
#pragma unroll
      for (int q = 0; q < NW; ++q) tails += sm.ntails[q];
      co->aggs[tile] = tile_agg;
      co->meta[tile] = tails | ((L.h & 1u) << 31);  // lane 0 of wave 0 holds the first element
    }
    const int64_t off = tile * int64_t(TILE + 1);
    rbk_write_runs(L, w, one, int64_t(TILE + 1), kout + off, vout + off, region, lane);
    return;
  }
  const S tile_prefix = prefixes ? known : chained ? sm.prefix : SOp::template identity<S>();
  // The tile holding element n-1 reports the run count.
  if (threadIdx.x == 0 && tile * TILE + TILE >= n)
    *num_runs = int64_t(tile_prefix.c + tile_agg.c);
  rbk_write_runs(L, w, tile_prefix, n, kout, vout, region, lane);
}


template <class K, class V, class Op, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void reduce_by_key_kernel(
    const K* __restrict__ keys, const V* __restrict__ vals, int64_t n, K* __restrict__ kout,
    V* __restrict__ vout, int64_t* __restrict__ num_runs, uint32_t* header, uint64_t* words,
    LookbackGuard* guard, bool single_tile) {
  __shared__ RbkSmem<K, V, Op, BLOCK, ITEMS> sm;
  __shared__ uint32_t s_tile;
  if (single_tile) {
    rbk_tile<K, V, Op, BLOCK, ITEMS>(sm, keys, vals, n, kout, vout, num_runs, 0, words, guard,
                                     true);
    return;
  }
  if (threadIdx.x == 0) s_tile = atomicAdd(header, 1u);
  __syncthreads();
  rbk_tile<K, V, Op, BLOCK, ITEMS>(sm, keys, vals, n, kout, vout, num_runs,
                                   int64_t(wave_uniform(s_tile)), words, guard, false);
}

// Three-pass form (variants kRbkShapes and up): the tiles never wait on
// one another.  Pass 1 writes every tile's aggregate (its head count and
// the fold of the values from its last head, or from its start when it has
// none, to its end), reading the keys and only those values; pass 2 scans
// the aggregates in one workgroup; pass 3 is rbk_tile with the tile's
// exclusive prefix read from pass 2.  The aggregates are those the
// scan_head_flags operator (reduce_by_key.hpp:41-62) folds, so pass 3
// produces what the single-pass kernel does.  It reads the keys twice
// (12 bytes per element for 4-byte keys and values instead of 8) but has
// no look-back, whose waits on the slowest predecessor of a window bound
// the single-pass kernel on large inputs.
template <class K, class V, class Op, int BLOCK, int ITEMS>
__device__ __forceinline__ void rbk_upsweep_tile(const K* __restrict__ keys,
                                                 const V* __restrict__ vals, int64_t n,
                                                 Seg<typename Acc<V, Op>::type>* __restrict__ aggs,
                                                 int64_t tile, uint8_t* __restrict__ flags = nullptr) {
  using AccP = Acc<V, Op>;
  using A = typename AccP::type;
  using AOp = typename AccP::op;
  using WK = WaveTile<K, ITEMS>;
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int NW = BLOCK / kWave;
  constexpr int VE = WK::VE;         // keys per 16-byte vector
  constexpr int NV = ITEMS / VE;     // vectors per thread
  constexpr int NVEC = BLOCK * NV;   // vectors per tile
  // Striped layout: thread t holds vectors j*BLOCK + t (coalesced 16-byte
  // loads, 1 KB per wavefront instruction).  A vector's first key is
  // compared with the last key of the vector before it, exchanged in LDS.
  __shared__ K s_tail[NVEC];
  __shared__ uint32_t s_cnt[NW];
  __shared__ int s_last[NW];
  __shared__ A s_val[NW];
  const int wid = threadIdx.x / kWave;
  const int64_t base = tile * TILE;
  const int cnt = int(n - base < TILE ? n - base : TILE);
  const bool full = cnt == TILE && aligned16(keys + base);
  K k[NV][VE];
  if (full) {
    const u32x4* q = reinterpret_cast<const u32x4*>(keys + base);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const u32x4 t = __builtin_nontemporal_load(q + j * BLOCK + threadIdx.x);
#pragma unroll
      for (int c = 0; c < VE; ++c) k[j][c] = reinterpret_cast<const K*>(&t)[c];
    }
  } else {
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int c = 0; c < VE; ++c) {
        const int e = (j * BLOCK + int(threadIdx.x)) * VE + c;
        k[j][c] = e < cnt ? keys[base + e] : K();
      }
  }
  // The last vectors of the tile (the last 64*VE keys) load their values
  // with the keys: the tile's last head is nearly always among them, which
  // saves a second memory round trip after the head search.
  // (whole 16-byte value vectors only: not for 8-byte keys with 4-byte values)
  constexpr bool kSpecOk = (VE * sizeof(V)) % 16 == 0;
  const bool spec =
      kSpecOk && full && int(threadIdx.x) >= BLOCK - kWave && aligned16(vals + base);
  constexpr int SV = kSpecOk ? int(VE * sizeof(V)) / 16 : 1;  // value vectors per key vector
  V sv[VE];
  if (spec) {
    const int e0 = ((NV - 1) * BLOCK + int(threadIdx.x)) * VE;
    const u32x4* q = reinterpret_cast<const u32x4*>(vals + base + e0);
#pragma unroll
    for (int u = 0; u < SV; ++u) {
      const u32x4 t = __builtin_nontemporal_load(q + u);
#pragma unroll
      for (int c = 0; c < VE / SV; ++c) sv[u * (VE / SV) + c] = reinterpret_cast<const V*>(&t)[c];
    }
  }
  const K gprev = (threadIdx.x == 0 && base > 0) ? keys[base - 1] : K();
#pragma unroll
  for (int j = 0; j < NV; ++j) s_tail[j * BLOCK + threadIdx.x] = k[j][VE - 1];
  __syncthreads();
  uint32_t heads = 0;
  int last = -1;
  bool head0 = false;  // the tile's first element is a head (thread 0)
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const int vi = j * BLOCK + int(threadIdx.x);
    const K before = vi > 0 ? s_tail[vi - 1] : gprev;
#pragma unroll
    for (int c = 0; c < VE; ++c) {
      const int e = vi * VE + c;
      const bool head = e < cnt && (base + e == 0 || !(k[j][c] == (c ? k[j][c - 1] : before)));
      heads += head;
      last = head ? e : last;  // e grows with j and c
      if (j == 0 && c == 0) head0 = head && threadIdx.x == 0;
    }
  }
  const uint32_t wc = wave_reduce<OpPlus>(heads);
  const int wl = wave_reduce<OpMax>(last);
  if (lane_id() == 0) {
    s_cnt[wid] = wc;
    s_last[wid] = wl;
  }
  __syncthreads();
  uint64_t tc = 0;
  int from = -1;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    tc += s_cnt[w];
    from = s_last[w] > from ? s_last[w] : from;
  }
  from = from < 0 ? 0 : from;  // no head: the whole tile continues the open run
  // fold of values [from, cnt): each thread folds its elements of that range
  // in element order, then lanes and waves combine in a fixed order
  A part = AOp::template identity<A>();
  // Values of the key vectors that reach [from, cnt): in a full aligned tile
  // as whole 16-byte vectors striped like the keys (the last wave's last
  // vector is already loaded), all in flight before the fold; runs longer
  // than a tile make this the whole tile's values.
  const bool vec = kSpecOk && full && aligned16(vals + base);
  if (vec) {
    V w[NV][VE];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const int e0 = (j * BLOCK + int(threadIdx.x)) * VE;
      if (j == NV - 1 && spec) {
#pragma unroll
        for (int c = 0; c < VE; ++c) w[j][c] = sv[c];
      } else if (e0 + VE > from) {
        const u32x4* q = reinterpret_cast<const u32x4*>(vals + base + e0);
#pragma unroll
        for (int u = 0; u < SV; ++u) {
          const u32x4 t = __builtin_nontemporal_load(q + u);
#pragma unroll
          for (int c = 0; c < VE / SV; ++c)
            w[j][u * (VE / SV) + c] = reinterpret_cast<const V*>(&t)[c];
        }
      } else {
#pragma unroll
        for (int c = 0; c < VE; ++c) w[j][c] = V();  // before the range: not folded
      }
    }
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int c = 0; c < VE; ++c) {
        const int e = (j * BLOCK + int(threadIdx.x)) * VE + c;
        if (e >= from) part = AOp::apply(part, AccP::from(w[j][c]));
      }
  } else {
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int c = 0; c < VE; ++c) {
        const int e = (j * BLOCK + int(threadIdx.x)) * VE + c;
        if (e >= from && e < cnt) part = AOp::apply(part, AccP::from(vals[base + e]));
      }
  }
  const A wv = wave_reduce<AOp>(part);
  if (lane_id() == 0) s_val[wid] = wv;
  __syncthreads();
  if (threadIdx.x == 0) {
    A v = AOp::template identity<A>();
#pragma unroll
    for (int w = 0; w < NW; ++w) v = AOp::apply(v, s_val[w]);
    Seg<A> g;
    g.c = tc;
    g.v = v;
    aggs[tile] = g;
    // bit 0: the tile starts a run; bit 1: a run starts after its first
    // element (so one ends inside it)
    if (flags) flags[tile] = uint8_t(head0) | uint8_t((tc - uint64_t(head0)) > 0) << 1;
  }
}

template <class K, class V, class Op, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void rbk_upsweep_kernel(
    const K* __restrict__ keys, const V* __restrict__ vals, int64_t n,
    Seg<typename Acc<V, Op>::type>* __restrict__ aggs, uint8_t* __restrict__ flags) {
  rbk_upsweep_tile<K, V, Op, BLOCK, ITEMS>(keys, vals, n, aggs, blockIdx.x, flags);
}

// Pass 2: exclusive scan of the tile aggregates in place.  With one
// workgroup it scans all `count` of them, rbk_scan_seg() per step; with
// several, workgroup b scans segment b only and writes the segment's total
// to totals[b] (then scanned by one workgroup, and added by pass 3).
// Exclusive scan of aggs[begin, end) in place by one 256-thread workgroup;
// returns the total (in every thread).
template <class V, class Op>
__device__ __forceinline__ Seg<typename Acc<V, Op>::type> rbk_scan_range(
    Seg<typename Acc<V, Op>::type>* __restrict__ aggs, int64_t begin, int64_t end) {
  using AccP = Acc<V, Op>;
  using A = typename AccP::type;
  using S = Seg<A>;
  using SOp = SegOp<typename AccP::op, A>;
  constexpr int ITEMS = rbk_scan_items<S>();
  constexpr int SEG = rbk_scan_seg<S>();
  constexpr int NW = 256 / kWave;
  __shared__ S s_wave[NW];
  const int wid = threadIdx.x / kWave;
  const S id = SOp::template identity<S>();
  S carry = id;
  for (int64_t c0 = begin; c0 < end; c0 += SEG) {
    const int64_t i0 = c0 + int64_t(threadIdx.x) * ITEMS;
    S x[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) x[j] = i0 + j < end ? aggs[i0 + j] : id;
    S t = id;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) t = SOp::apply(t, x[j]);
    const S inc = wave_inclusive_scan<SOp>(t);
    const S lane_excl = wave_shift_right<SOp>(inc);
    if (lane_id() == kWave - 1) s_wave[wid] = inc;
    __syncthreads();
    S wave_excl = id, total = id;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      if (w < wid) wave_excl = SOp::apply(wave_excl, s_wave[w]);
      total = SOp::apply(total, s_wave[w]);
    }
    S pre = SOp::apply(carry, SOp::apply(wave_excl, lane_excl));
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      if (i0 + j < end) aggs[i0 + j] = pre;
      pre = SOp::apply(pre, x[j]);
    }
    carry = SOp::apply(carry, total);
    __syncthreads();  // s_wave is rewritten by the next step
  }
  return carry;
}

template <class V, class Op>
__global__ __launch_bounds__(256) void rbk_scan_aggs_kernel(
    Seg<typename Acc<V, Op>::type>* __restrict__ aggs, int64_t count,
    Seg<typename Acc<V, Op>::type>* __restrict__ totals) {
  using S = Seg<typename Acc<V, Op>::type>;
  constexpr int SEG = rbk_scan_seg<S>();
  const int64_t begin = int64_t(blockIdx.x) * SEG;
  const int64_t end = gridDim.x == 1 ? count : (begin + SEG < count ? begin + SEG : count);
  const S total = rbk_scan_range<V, Op>(aggs, begin, end);
  if (totals && threadIdx.x == 0) totals[blockIdx.x] = total;
}

// Pass 3: one tile per workgroup, in any order.
template <class K, class V, class Op, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void rbk_downsweep_kernel(
    const K* __restrict__ keys, const V* __restrict__ vals, int64_t n, K* __restrict__ kout,
    V* __restrict__ vout, int64_t* __restrict__ num_runs,
    const Seg<typename Acc<V, Op>::type>* __restrict__ prefixes,
    const Seg<typename Acc<V, Op>::type>* __restrict__ group_prefixes,
    const uint8_t* __restrict__ flags, int64_t tiles) {
  __shared__ RbkSmem<K, V, Op, BLOCK, ITEMS> sm;
  const int64_t tile = blockIdx.x;
  // A tile in which no run ends writes nothing (its runs' sums reach later
  // tiles through the aggregates): with 8-byte values it is skipped without
  // reading its keys or values, which runs spanning many tiles (Zipf-like
  // lengths) make the common case (measured at 2^28: -27 % time for Zipf
  // runs, unchanged for runs of ~8).  With 4-byte values a flag read ahead
  // of the tile's loads cost 9 % on runs of ~8, so the tile reads its flags
  // behind its key loads and skips only its values.  The last tile always
  // holds a run's end.
  if constexpr (sizeof(V) >= 8) {
    if (tile + 1 < tiles && !(flags[tile] & 2u) && !(flags[tile + 1] & 1u)) return;
  }
  rbk_tile<K, V, Op, BLOCK, ITEMS>(sm, keys, vals, n, kout, vout, num_runs, tile, nullptr,
                                   nullptr, false, prefixes, group_prefixes,
                                   sizeof(V) >= 8 ? nullptr : flags, tiles);
}

// Compact form (variants 9-11): keys and values are read once.  Pass 1
// (rbk_compact_kernel) is the tile above writing the runs that end in a
// tile to the tile's own scratch slots with the tile-local value of a run
// carried in, plus the tile's aggregate; pass 2 scans the aggregates (as the
// three-pass form); pass 3 (rbk_move_kernel) moves each tile's runs to
// their places, folding the carried-in value into its first one.  Bytes
// per element for 4-byte keys and values: 8 + 24 * runs / n (11 for runs of
// 8 elements, ~8 for long runs) against the three-pass form's 12-13.
template <class K, class V, class Op, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void rbk_compact_kernel(
    const K* __restrict__ keys, const V* __restrict__ vals, int64_t n, K* __restrict__ tk,
    V* __restrict__ tv, Seg<typename Acc<V, Op>::type>* __restrict__ aggs,
    uint32_t* __restrict__ meta) {
  __shared__ RbkSmem<K, V, Op, BLOCK, ITEMS> sm;
  const RbkCompactOut<Seg<typename Acc<V, Op>::type>> co{aggs, meta};
  rbk_tile<K, V, Op, BLOCK, ITEMS>(sm, keys, vals, n, tk, tv, nullptr, blockIdx.x, nullptr,
                                   nullptr, false, nullptr, nullptr, nullptr, 0, &co);
}

// Pass 3 of the compact form: tile t's runs, scratch slots [h0, h0 + m),
// go to output slots prefix.c + slot - 1; slot 0 (the run carried into the
// tile) gets the prefix's value folded in.  The last tile reports the run
// count.  One workgroup of 256 lanes per tile (shared with the batch
// executor).
template <class K, class V, class Op, int TILE>
__device__ __forceinline__ void rbk_move_tile(const K* __restrict__ tk, const V* __restrict__ tv,
                                              const Seg<typename Acc<V, Op>::type>* __restrict__ prefixes,
                                              const Seg<typename Acc<V, Op>::type>* __restrict__ group_prefixes,
                                              const uint32_t* __restrict__ meta, int64_t tiles,
                                              int64_t n, K* __restrict__ kout, V* __restrict__ vout,
                                              int64_t* __restrict__ num_runs, int64_t tile) {
  using AccP = Acc<V, Op>;
  using S = Seg<typename AccP::type>;
  using SOp = SegOp<typename AccP::op, typename AccP::type>;
  const uint32_t m = meta[tile];
  const uint32_t cnt = m & 0x7fffffffu, h0 = m >> 31;
  const S pre = group_prefixes ? SOp::apply(group_prefixes[tile / rbk_scan_seg<S>()], prefixes[tile])
                               : prefixes[tile];
  const int64_t src = tile * int64_t(TILE + 1) + h0;
  const int64_t dst = int64_t(pre.c) + h0 - 1;
  // bounds guard: slots stay inside [0, n) when the scan is consistent
  const uint32_t c = dst < 0 || dst + int64_t(cnt) > n ? 0u : cnt;
  for (uint32_t x = threadIdx.x; x < c; x += blockDim.x) {
    kout[dst + x] = tk[src + x];
    V v = tv[src + x];
    if (h0 == 0 && x == 0) v = AccP::to(AccP::op::apply(pre.v, AccP::from(v)));
    vout[dst + x] = v;
  }
  if (tile == tiles - 1 && threadIdx.x == 0) *num_runs = int64_t(pre.c) + h0 + cnt - 1;
}

template <class K, class V, class Op, int TILE>
__global__ __launch_bounds__(256) void rbk_move_kernel(
    const K* __restrict__ tk, const V* __restrict__ tv,
    const Seg<typename Acc<V, Op>::type>* __restrict__ prefixes,
    const Seg<typename Acc<V, Op>::type>* __restrict__ group_prefixes,
    const uint32_t* __restrict__ meta, int64_t tiles, int64_t n, K* __restrict__ kout,
    V* __restrict__ vout, int64_t* __restrict__ num_runs) {
  rbk_move_tile<K, V, Op, TILE>(tk, tv, prefixes, group_prefixes, meta, tiles, n, kout, vout,
                                num_runs, int64_t(blockIdx.x));
}

// Workspace of the compact form of tile T over n elements: aggregates,
// segment totals and meta per tile, then the scratch keys and values.
struct RbkCompactLayout {
  size_t aggs, totals, meta, tk, tv, total;
  RbkCompactLayout(size_t ks, size_t vs, size_t ss, int64_t n, int64_t T, int64_t seg) {
    const int64_t tiles = (n + T - 1) / T;
    const int64_t groups = (tiles + seg - 1) / seg;
    const size_t slots = size_t(tiles) * size_t(T + 1);
    aggs = kRbkHeaderBytes;
    totals = aggs + size_t(tiles) * ss;
    meta = totals + size_t(groups) * ss;
    tk = align_up(meta + size_t(tiles) * sizeof(uint32_t), 256);
    tv = align_up(tk + slots * ks, 256);
    total = tv + slots * vs;
  }
};

__global__ void rbk_empty_kernel(int64_t* num_runs);
__global__ void rbk_prepare_kernel(uint32_t* header, uint64_t* words, int64_t num_words);

template <class C>
inline int64_t rbk_num_tiles(int64_t n) {
  return (n + C::TILE - 1) / C::TILE;
}

template <class K, class V, class Op, int VAR, int FORM = kRbkSinglePass>
int rbk_launch(const K* keys, const V* vals, int64_t n, K* kout, V* vout, int64_t* num_runs,
               char* ws, hipStream_t stream) {
  using C = RbkConfig<K, V, VAR>;
  using S = Seg<typename Acc<V, Op>::type>;
  if (n == 0) {
    rbk_empty_kernel<<<1, 64, 0, stream>>>(num_runs);
    BULKX_LAUNCHED();
    return BULKX_SUCCESS;
  }
  const int64_t tiles = rbk_num_tiles<C>(n);
  if (tiles == 1) {
    reduce_by_key_kernel<K, V, Op, C::BLOCK, C::ITEMS><<<1, C::BLOCK, 0, stream>>>(
        keys, vals, n, kout, vout, num_runs, nullptr, nullptr, nullptr, true);
    BULKX_LAUNCHED();
    return BULKX_SUCCESS;
  }
  if constexpr (FORM == kRbkCompact) {
    static_assert(C::TILE <= kRbkCompactMaxTile, "scratch sized for the largest compact tile");
    constexpr int64_t SEG = rbk_scan_seg<S>();
    const RbkCompactLayout lay(sizeof(K), sizeof(V), sizeof(S), n, C::TILE, SEG);
    S* aggs = reinterpret_cast<S*>(ws + lay.aggs);
    const int64_t groups = (tiles + SEG - 1) / SEG;
    S* totals = groups > 1 ? reinterpret_cast<S*>(ws + lay.totals) : nullptr;
    uint32_t* meta = reinterpret_cast<uint32_t*>(ws + lay.meta);
    K* tk = reinterpret_cast<K*>(ws + lay.tk);
    V* tv = reinterpret_cast<V*>(ws + lay.tv);
    rbk_compact_kernel<K, V, Op, C::BLOCK, C::ITEMS><<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(
        keys, vals, n, tk, tv, aggs, meta);
    BULKX_LAUNCHED();
    rbk_scan_aggs_kernel<V, Op><<<dim3(uint32_t(groups)), 256, 0, stream>>>(aggs, tiles, totals);
    BULKX_LAUNCHED();
    if (groups > 1) {
      rbk_scan_aggs_kernel<V, Op><<<1, 256, 0, stream>>>(totals, groups, nullptr);
      BULKX_LAUNCHED();
    }
    rbk_move_kernel<K, V, Op, C::TILE><<<dim3(uint32_t(tiles)), 256, 0, stream>>>(
        tk, tv, aggs, totals, meta, tiles, n, kout, vout, num_runs);
    BULKX_LAUNCHED();
    return BULKX_SUCCESS;
  }
  if constexpr (FORM == kRbkThreePass) {
    S* aggs = reinterpret_cast<S*>(ws + kRbkHeaderBytes);
    // the segment totals follow the aggregates, then one flag byte per tile
    const int64_t groups = (tiles + rbk_scan_seg<S>() - 1) / rbk_scan_seg<S>();
    S* totals = groups > 1 ? aggs + tiles : nullptr;
    uint8_t* flags = reinterpret_cast<uint8_t*>(aggs + tiles + groups);
    // per tile (tiles >= 2): an aggregate, a flag byte and less than half an
    // aggregate of totals, inside the status words' 2 * sizeof(S)
    static_assert(sizeof(S) + 1 + sizeof(S) / 2 <= TileStatus<S>::W * sizeof(uint64_t),
                  "aggregates, totals and flags fit the status words");
    rbk_upsweep_kernel<K, V, Op, C::BLOCK, C::ITEMS><<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(
        keys, vals, n, aggs, flags);
    BULKX_LAUNCHED();
    // pass 2: segments of rbk_scan_seg() aggregates in parallel, then their
    // totals in one workgroup
    rbk_scan_aggs_kernel<V, Op><<<dim3(uint32_t(groups)), 256, 0, stream>>>(aggs, tiles, totals);
    BULKX_LAUNCHED();
    if (groups > 1) {
      rbk_scan_aggs_kernel<V, Op><<<1, 256, 0, stream>>>(totals, groups, nullptr);
      BULKX_LAUNCHED();
    }
    rbk_downsweep_kernel<K, V, Op, C::BLOCK, C::ITEMS>
        <<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(keys, vals, n, kout, vout, num_runs, aggs,
                                                          totals, flags, tiles);
    BULKX_LAUNCHED();
    return BULKX_SUCCESS;
  }
  LookbackGuard* guard = nullptr;
  BULKX_TRY(lookback_guard(&guard));
  uint32_t* header = reinterpret_cast<uint32_t*>(ws);
  uint64_t* words = reinterpret_cast<uint64_t*>(ws + kRbkHeaderBytes);
  const int64_t num_words = tiles * TileStatus<S>::W;
  const int64_t pb = (num_words + 255) / 256;
  rbk_prepare_kernel<<<int(pb < 1024 ? pb : 1024), 256, 0, stream>>>(header, words, num_words);
  BULKX_LAUNCHED();
  reduce_by_key_kernel<K, V, Op, C::BLOCK, C::ITEMS><<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(
      keys, vals, n, kout, vout, num_runs, header, words, guard, false);
  BULKX_LAUNCHED();
  return BULKX_SUCCESS;
}

// Dispatch over value type and op for one key type K.
template <class K>
int rbk_dispatch_values(int vdtype, int op, const void* keys, const void* vals, int64_t n,
                        void* kout, void* vout, int64_t* num_runs, char* ws, hipStream_t stream,
                        int variant) {
  int st = BULKX_SUCCESS;
  BULKX_DISPATCH_DTYPE(vdtype, V, BULKX_DISPATCH_OP(op, OP, {
    const K* k = static_cast<const K*>(keys);
    const V* v = static_cast<const V*>(vals);
    K* ko = static_cast<K*>(kout);
    V* vo = static_cast<V*>(vout);
    // the other shapes are instantiated for plus only (min / max use the default)
    if constexpr (std::is_same<OP, OpPlus>::value) {
      switch (variant) {
        case 1: st = rbk_launch<K, V, OP, 1>(k, v, n, ko, vo, num_runs, ws, stream); break;
        case 2: st = rbk_launch<K, V, OP, 2>(k, v, n, ko, vo, num_runs, ws, stream); break;
        // 1024-lane tiles fit the LDS only with 4-byte keys and values
        case 3:
          if constexpr (sizeof(K) == 4 && sizeof(V) == 4)
            st = rbk_launch<K, V, OP, 3>(k, v, n, ko, vo, num_runs, ws, stream);
          else
            st = rbk_launch<K, V, OP, 0>(k, v, n, ko, vo, num_runs, ws, stream);
          break;
        case 4:
          if constexpr (sizeof(K) == 4 && sizeof(V) == 4)
            st = rbk_launch<K, V, OP, 4>(k, v, n, ko, vo, num_runs, ws, stream);
          else
            st = rbk_launch<K, V, OP, 0>(k, v, n, ko, vo, num_runs, ws, stream);
          break;
        case 5: st = rbk_launch<K, V, OP, 5>(k, v, n, ko, vo, num_runs, ws, stream); break;
        // three-pass forms of shapes 0, 2 and 5
        case 6: st = rbk_launch<K, V, OP, 0, kRbkThreePass>(k, v, n, ko, vo, num_runs, ws, stream); break;
        case 7: st = rbk_launch<K, V, OP, 2, kRbkThreePass>(k, v, n, ko, vo, num_runs, ws, stream); break;
        case 8: st = rbk_launch<K, V, OP, 5, kRbkThreePass>(k, v, n, ko, vo, num_runs, ws, stream); break;
        // compact forms of shapes 0, 2 and 5
        case 9: st = rbk_launch<K, V, OP, 0, kRbkCompact>(k, v, n, ko, vo, num_runs, ws, stream); break;
        case 10: st = rbk_launch<K, V, OP, 2, kRbkCompact>(k, v, n, ko, vo, num_runs, ws, stream); break;
        case 11: st = rbk_launch<K, V, OP, 5, kRbkCompact>(k, v, n, ko, vo, num_runs, ws, stream); break;
        default: st = rbk_launch<K, V, OP, 0>(k, v, n, ko, vo, num_runs, ws, stream); break;
      }
    } else {
      st = rbk_launch<K, V, OP, 0>(k, v, n, ko, vo, num_runs, ws, stream);
    }
  }));
  return st;
}

int rbk_dispatch_k4(int kdtype, int vdtype, int op, const void* keys, const void* vals, int64_t n,
                    void* kout, void* vout, int64_t* num_runs, char* ws, hipStream_t stream,
                    int variant);
int rbk_dispatch_k8(int kdtype, int vdtype, int op, const void* keys, const void* vals, int64_t n,
                    void* kout, void* vout, int64_t* num_runs, char* ws, hipStream_t stream,
                    int variant);

}  // namespace bulkx
