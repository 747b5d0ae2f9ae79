// Single-pass scan tile body (decoupled look-back), shared by the
// device-wide scan kernel and the batch executor.
#pragma once

#include "device_common.hpp"

namespace bulkx {

template <class T, class Op, int BLOCK>
struct ScanSmem {
  using A = typename Acc<T, Op>::type;
  using R = typename Acc<T, Op>::row_t;
  R wave[BLOCK / kWave];
  A prefix;
};

// Scans tile `tile` of in[0, n) into out.  WRITE = false computes only the
// tile chain and stores the grand total (init op x[0] op ... op x[n-1]) to
// *total from the last tile (the batch executor's reduce jobs).
template <class T, class Op, int BLOCK, int ITEMS, bool EXCLUSIVE, bool WRITE = true>
__device__ __forceinline__ void scan_tile(ScanSmem<T, Op, BLOCK>& sm, const T* in, T* out,
                                          int64_t n, T init, bool has_init, int64_t tile,
                                          uint64_t* status_words, LookbackGuard* guard,
                                          bool single_tile, T* total) {
  using AccP = Acc<T, Op>;
  using A = typename AccP::type;
  using AOp = typename AccP::op;
  using SL = Striped<T, ITEMS>;
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int NW = BLOCK / kWave;
  const TileStatus<A> st{status_words};
  const int wid = threadIdx.x / kWave;
  const int64_t chunk_base = tile * TILE + int64_t(wid) * SL::N;
  const int64_t rem = n - chunk_base;
  const int count = rem <= 0 ? 0 : (rem >= SL::N ? SL::N : int(rem));
  const T id = Op::template identity<T>();
  const A aid = AOp::template identity<A>();

  T x[ITEMS];
  // Nontemporal loads and stores: every element is read once and written
  // once, so neither stream is worth a place in the caches.
  striped_load<T, ITEMS, SL::VE, true>(in + (count > 0 ? chunk_base : 0), count, x, id);

  // Lane-local inclusive scan of each 16-byte vector (the bounded-agent
  // body of bulk::inclusive_scan, scan.hpp:35-55, over VE elements).
#pragma unroll
  for (int j = 0; j < SL::NV; ++j)
#pragma unroll
    for (int q = 1; q < SL::VE; ++q)
      x[j * SL::VE + q] = Op::apply(x[j * SL::VE + q - 1], x[j * SL::VE + q]);

  // DPP scan across each row, then the rows and the waves of the tile are
  // combined in the row accumulator (double for float and float64 input);
  // only the chain from tile to tile uses the carry accumulator (the
  // double-double of float64), whose error must not grow with the number of
  // tiles.  Inside a tile an element's prefix passes through at most
  // 1 + 6 + 7 + 15 + 1 additions (lane, row, rows, waves, tile prefix), so
  // its error stays below 4 * ceil(log2 n) ulps of sum |x| at every n.
  using R = typename AccP::row_t;
  using ROp = typename AccP::row_op;
  const R rid = ROp::template identity<R>();
  R rowp[SL::NV];
  R rowt[SL::NV];  // row totals (wave-uniform)
  R carry = rid;
#pragma unroll
  for (int j = 0; j < SL::NV; ++j) {
    const R inc = wave_inclusive_scan<ROp>(AccP::row(x[j * SL::VE + SL::VE - 1]));
    rowp[j] = wave_shift_right<ROp>(inc);
    rowt[j] = readlane(inc, kWave - 1);
    carry = ROp::apply(carry, rowt[j]);
  }

  // Combine the waves of the tile: lane w of every wavefront takes wave w's
  // total and one DPP scan over the lanes gives the wave's exclusive prefix
  // and the tile's aggregate (one scan instead of 2 * NW additions per lane).
  static_assert(NW <= kWave, "one lane per wavefront of the tile");
  if (lane_id() == 0) sm.wave[wid] = carry;
  __syncthreads();
  const R winc = wave_inclusive_scan<ROp>(lane_id() < NW ? sm.wave[lane_id()] : rid);
  const R wave_prefix = wid > 0 ? readlane(winc, wid - 1) : rid;
  const A tile_agg = AccP::lift(readlane(winc, NW - 1));

  A tile_prefix;  // everything before this tile, init included
  if (tile == 0) {
    tile_prefix = has_init ? AccP::from(init) : aid;
    if (!single_tile && threadIdx.x == 0)
      st.publish(0, kStatusPrefix, AOp::apply(tile_prefix, tile_agg));
  } else {
    if (threadIdx.x == 0) st.publish(tile, kStatusAggregate, tile_agg);
    if (threadIdx.x < kWave) {
      A excl = lookback_exclusive_prefix<AOp>(st, tile, guard);
      if (threadIdx.x == 0) {
        st.publish(tile, kStatusPrefix, AOp::apply(excl, tile_agg));
        sm.prefix = excl;
      }
    }
    __syncthreads();
    tile_prefix = sm.prefix;
  }

  R rc = wave_prefix;  // the tile's elements before row j
#pragma unroll
  for (int j = 0; j < SL::NV; ++j) {
    // prefix of this lane's vector in row j: the tiles before, then the
    // tile's own elements before it
    const A p = AOp::apply(tile_prefix, AccP::lift(ROp::apply(rc, rowp[j])));
    rc = ROp::apply(rc, rowt[j]);
    if constexpr (EXCLUSIVE) {
#pragma unroll
      for (int q = SL::VE - 1; q > 0; --q)
        x[j * SL::VE + q] = AccP::finish(p, AccP::row(x[j * SL::VE + q - 1]));
      x[j * SL::VE] = AccP::to(p);
    } else {
#pragma unroll
      for (int q = 0; q < SL::VE; ++q)
        x[j * SL::VE + q] = AccP::finish(p, AccP::row(x[j * SL::VE + q]));
    }
  }
  if constexpr (WRITE) {
    if (count > 0) striped_store<T, ITEMS, SL::VE, true>(out + chunk_base, count, x);
  } else {
    if (threadIdx.x == 0 && tile * TILE + TILE >= n)
      *total = AccP::to(AOp::apply(tile_prefix, tile_agg));
  }
}


}  // namespace bulkx
