// Merge-path machinery shared by merge, merge_by_key and the merge sort.
//
// bulk::merge_path (merge.hpp:35-60) finds, for a diagonal d of the merge
// matrix, how many of the first d outputs come from range A; ties go to A
// because the probe is comp(B[d-1-m], A[m]).  Here the global search of a
// tile boundary is done by one wavefront probing 64 positions per step
// (about 5 steps for 2^27 elements instead of 27 dependent loads), so no
// separate partition launch is needed (the reference's thrust::tabulate,
// merge.cu:152-154).  The tile itself is staged through LDS like the
// reference's staged_merge (merge.cu:12-36): each thread finds its own
// diagonal inside the tile and runs the sequential agent merge
// (merge.hpp:63-170) over its grain.
#pragma once

#include "device_common.hpp"
#include "host_common.hpp"

namespace bulkx {

// Tile shapes (same tile, two thread counts): variant 0 = 256 threads x
// 16 (4-byte keys) / 8 items, variant 1 = 512 threads x 8 / 4 items (shorter
// sequential merges per lane, twice the lanes).
template <class K, int VAR = 0>
struct MergeConfig {
  static constexpr int BLOCK = VAR == 0 ? 256 : 512;
  static constexpr int ITEMS = (sizeof(K) == 4 ? 16 : 8) / (VAR == 0 ? 1 : 2);
  static constexpr int TILE = BLOCK * ITEMS;
};
constexpr int kMergeVariants = 2;

// Wave-cooperative merge_path over global memory.  Every lane of the
// calling wavefront must participate; every lane returns the same result.
template <class CMP, class K>
__device__ int64_t wave_merge_path(const K* a, int64_t n1, const K* b, int64_t n2, int64_t diag) {
  const int lane = lane_id();
  int64_t lo = diag > n2 ? diag - n2 : 0;
  int64_t hi = diag < n1 ? diag : n1;
  while (lo < hi) {
    const int64_t len = hi - lo;
    int64_t m;
    bool inside;
    if (len <= kWave) {
      m = lo + lane;
      inside = m < hi;
    } else {
      m = lo + ((len * lane) >> 6);
      inside = true;
    }
    // predicate is monotone in m: false below the answer, true from it on
    bool p = true;
    if (inside) p = CMP::apply(b[diag - 1 - m], a[m]);
    const uint64_t mask = __ballot(p);
    if (len <= kWave) {
      lo = mask ? readlane(m, __builtin_ctzll(mask)) : hi;
      break;
    }
    if (mask == 0) {
      lo = readlane(m, kWave - 1) + 1;
    } else {
      const int f = __builtin_ctzll(mask);
      const int64_t mf = readlane(m, f);
      if (f > 0) lo = readlane(m, f - 1) + 1;
      hi = mf;
    }
  }
  return lo;
}

// merge_path searched by one lane: a plain binary search.  Used where many
// boundaries are searched at once (the partition kernels): each search
// touches about 2*log2(n) cache lines, against ~128 per step of the
// wave-cooperative search, and the dependent loads of different lanes
// overlap in flight.
template <class CMP, class K>
__device__ __forceinline__ int64_t lane_merge_path(const K* a, int64_t n1, const K* b, int64_t n2,
                                                   int64_t diag) {
  int64_t lo = diag > n2 ? diag - n2 : 0;
  int64_t hi = diag < n1 ? diag : n1;
  while (lo < hi) {
    const int64_t m = (lo + hi) >> 1;
    if (CMP::apply(b[diag - 1 - m], a[m])) hi = m;
    else lo = m + 1;
  }
  return lo;
}

// Padded LDS index: one pad element per 128 bytes, so that lanes reading
// elements a multiple of 32 words apart (blocked per-thread layouts, merge
// paths of neighbouring lanes) fall into different banks.
// Indices are non-negative, so the division is a shift.
template <class T>
__device__ __forceinline__ int lpad(int e) {
  return e + int(unsigned(e) / unsigned(128 / sizeof(T)));
}
// LDS elements a padded array of n logical elements needs.
template <class T>
constexpr int lpad_elems(int n) {
  return n + n / (128 / int(sizeof(T)));
}

// Sequential merge_path over padded LDS: A = s[base, base + n1),
// B = s[base + b_off, base + b_off + n2) (logical indices).
template <class CMP, class K>
__device__ __forceinline__ int lds_merge_path(const K* s, int base, int n1, int b_off, int n2,
                                              int diag) {
  int lo = diag > n2 ? diag - n2 : 0;
  int hi = diag < n1 ? diag : n1;
  while (lo < hi) {
    int m = (lo + hi) >> 1;
    if (CMP::apply(s[lpad<K>(base + b_off + diag - 1 - m)], s[lpad<K>(base + m)])) hi = m;
    else lo = m + 1;
  }
  return lo;
}

// Sequential agent merge of up to ITEMS outputs from padded LDS runs
// A = [i1, e1), B = [i2, e2) (logical indices; the merge.hpp:109-129 body:
// take A unless B's head compares strictly less).  Produces keys and the
// logical LDS index they came from.  Both heads stay in registers, so each
// step reads one element (the successor of the head it consumed): LDS reads
// at data-dependent addresses are the bank-conflict-bound part of a merge.
// A read past the end of its run is clamped to LAST (the stage's last
// logical index) and never used, so it needs no branch; outputs past the
// caller's count are computed the same way and ignored by the caller.
template <class CMP, class K, int ITEMS, int LAST>
__device__ __forceinline__ void lds_serial_merge(const K* s, int i1, int e1, int i2, int e2,
                                                 K (&rk)[ITEMS], int (&ri)[ITEMS]) {
  K ka = s[lpad<K>(i1 < LAST ? i1 : LAST)];
  K kb = s[lpad<K>(i2 < LAST ? i2 : LAST)];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const bool has1 = i1 < e1, has2 = i2 < e2;
    const bool p = !has2 || (has1 && !CMP::apply(kb, ka));
    rk[i] = p ? ka : kb;
    ri[i] = p ? i1 : i2;
    if (i + 1 < ITEMS) {
      const int nx = (p ? i1 : i2) + 1;
      const K kn = s[lpad<K>(nx < LAST ? nx : LAST)];
      if (p) {
        i1 = nx;
        ka = kn;
      } else {
        i2 = nx;
        kb = kn;
      }
    }
  }
}

// LDS elements the staging arrays of merge_tile need: the tile itself, and
// the padded per-wave transpose regions used for the output stores.
template <class T, int BLOCK, int ITEMS>
struct MergeLds {
  static constexpr int ELEMS = (BLOCK / kWave) * WaveTile<T, ITEMS>::LDS_ELEMS;
  static_assert(ELEMS >= lpad_elems<T>(BLOCK * ITEMS), "room for the padded stage");
};

// Merge of one tile: A[0, na) and B[0, nb) (na + nb <= TILE) into
// out[0, na + nb).  s_k / s_v hold MergeLds<...>::ELEMS elements.  Ends with
// a workgroup barrier, so it can be called again on the same LDS.
template <class CMP, class K, class V, bool HAS_V, int BLOCK, int ITEMS>
__device__ __forceinline__ void merge_tile(const K* a, const V* va, int na, const K* b,
                                           const V* vb, int nb, K* out, V* vout, K* s_k,
                                           V* s_v) {
  // Bounds guard (only reachable with a comparator that is not a strict weak
  // order, e.g. NaN keys): never stage or write more than one tile.
  constexpr int TILE = BLOCK * ITEMS;
  na = na < 0 ? 0 : (na > TILE ? TILE : na);
  nb = nb < 0 ? 0 : (nb > TILE - na ? TILE - na : nb);
  const int cnt = na + nb;
  // Stage A ++ B: all loads of the thread are issued before the first LDS
  // write, so they are in flight together.
  {
    K lk[ITEMS];
    V lv[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int j = threadIdx.x + i * BLOCK;
      if (j < cnt) {
        const bool fa = j < na;
        lk[i] = __builtin_nontemporal_load(fa ? a + j : b + (j - na));
        if constexpr (HAS_V) lv[i] = __builtin_nontemporal_load(fa ? va + j : vb + (j - na));
      }
    }
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int j = threadIdx.x + i * BLOCK;
      if (j < cnt) {
        s_k[lpad<K>(j)] = lk[i];
        if constexpr (HAS_V) s_v[lpad<V>(j)] = lv[i];
      }
    }
  }
  __syncthreads();
  const int d = threadIdx.x * ITEMS < cnt ? threadIdx.x * ITEMS : cnt;
  const int ai = lds_merge_path<CMP>(s_k, 0, na, na, nb, d);
  const int bi = na + (d - ai);
  K rk[ITEMS];
  V rv[ITEMS];
  int ri[ITEMS];
  const int count = cnt - d < ITEMS ? cnt - d : ITEMS;
  lds_serial_merge<CMP, K, ITEMS, TILE - 1>(s_k, ai, na, bi, cnt, rk, ri);
  if constexpr (HAS_V) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) rv[i] = i < count ? s_v[lpad<V>(ri[i])] : V();
  }
  __syncthreads();  // every read of the stage is done; reuse it for the transposes
  const int wid = threadIdx.x / kWave;
  const int wb = wid * WaveTile<K, ITEMS>::N;
  const int wcount = cnt - wb <= 0 ? 0 : (cnt - wb >= WaveTile<K, ITEMS>::N ? WaveTile<K, ITEMS>::N : cnt - wb);
  if (wcount > 0) {
    wave_store_blocked<K, ITEMS>(out + wb, wcount, rk, s_k + wid * WaveTile<K, ITEMS>::LDS_ELEMS);
    if constexpr (HAS_V)
      wave_store_blocked<V, ITEMS>(vout + wb, wcount, rv, s_v + wid * WaveTile<V, ITEMS>::LDS_ELEMS);
  }
  __syncthreads();
}

// Tile boundaries of a two-range merge: wave 0 searches the tile's first
// diagonal, wave 1 its last; results land in s_split[0..1].
template <class CMP, class K>
__device__ __forceinline__ void tile_splits(const K* a, int64_t n1, const K* b, int64_t n2,
                                            int64_t d0, int64_t d1, int64_t* s_split) {
  const int wid = threadIdx.x / kWave;
  if (wid < 2) {
    const int64_t r = wave_merge_path<CMP>(a, n1, b, n2, wid == 0 ? d0 : d1);
    if (lane_id() == 0) s_split[wid] = r;
  }
  __syncthreads();
}

// LDS of a merge tile.
template <class K, class V, bool HAS_V, int BLOCK, int ITEMS>
struct MergeSmem {
  K k[MergeLds<K, BLOCK, ITEMS>::ELEMS];
  V v[HAS_V ? MergeLds<V, BLOCK, ITEMS>::ELEMS : 1];
  int64_t split[2];
};

// Output tile `tile` of the stable merge of a[0, n1) and b[0, n2).
template <class CMP, class K, class V, bool HAS_V, int BLOCK, int ITEMS>
__device__ __forceinline__ void merge_two_tile(MergeSmem<K, V, HAS_V, BLOCK, ITEMS>& sm,
                                               const K* a, const V* va, int64_t n1, const K* b,
                                               const V* vb, int64_t n2, K* out, V* vout,
                                               int64_t tile, const int64_t* splits = nullptr) {
  constexpr int TILE = BLOCK * ITEMS;
  const int64_t n = n1 + n2;
  const int64_t d0 = tile * TILE;
  const int64_t d1 = d0 + TILE < n ? d0 + TILE : n;
  // the tile's two boundaries: precomputed (splits[tile], splits[tile + 1])
  // or searched here
  int64_t a0, a1;
  if (splits) {
    a0 = splits[tile];
    a1 = splits[tile + 1];
  } else {
    tile_splits<CMP>(a, n1, b, n2, d0, d1, sm.split);
    a0 = sm.split[0];
    a1 = sm.split[1];
  }
  const int64_t b0 = d0 - a0, b1 = d1 - a1;
  merge_tile<CMP, K, V, HAS_V, BLOCK, ITEMS>(a + a0, HAS_V ? va + a0 : nullptr, int(a1 - a0),
                                             b + b0, HAS_V ? vb + b0 : nullptr, int(b1 - b0),
                                             out + d0, HAS_V ? vout + d0 : nullptr, sm.k, sm.v);
}

}  // namespace bulkx
