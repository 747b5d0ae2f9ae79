// Single-pass scan tile body (decoupled look-back), shared by the
// device-wide scan kernel and the batch executor.
#pragma once

#include "device_common.hpp"

namespace bulkx {

template <class T, class Op, int BLOCK>
struct ScanSmem {
  using A = typename Acc<T, Op>::type;
  A wave[BLOCK / kWave];
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

  // DPP scan across each row (row accumulator), carried from row to row
  // (carry accumulator).
  using R = typename AccP::row_t;
  using ROp = typename AccP::row_op;
  R rowp[SL::NV];
  R rowt[SL::NV];  // row totals (wave-uniform)
  A carry = aid;
#pragma unroll
  for (int j = 0; j < SL::NV; ++j) {
    const R inc = wave_inclusive_scan<ROp>(AccP::row(x[j * SL::VE + SL::VE - 1]));
    rowp[j] = wave_shift_right<ROp>(inc);
    rowt[j] = readlane(inc, kWave - 1);
    carry = AOp::apply(carry, AccP::lift(rowt[j]));
  }

  // Combine the waves of the tile.
  if (lane_id() == 0) sm.wave[wid] = carry;
  __syncthreads();
  A wave_prefix = aid, tile_agg = aid;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const A t = sm.wave[w];
    if (w < wid) wave_prefix = AOp::apply(wave_prefix, t);
    tile_agg = AOp::apply(tile_agg, t);
  }

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

  A rc = AOp::apply(tile_prefix, wave_prefix);  // carry into row j
#pragma unroll
  for (int j = 0; j < SL::NV; ++j) {
    // prefix of this lane's vector in row j: carry level + row level
    const A p = AOp::apply(rc, AccP::lift(rowp[j]));
    rc = AOp::apply(rc, AccP::lift(rowt[j]));
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
