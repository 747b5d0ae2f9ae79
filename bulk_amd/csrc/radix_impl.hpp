// LSD radix sort of 4-byte keys (int32, uint32, float32), stable, with an
// optional 4- or 8-byte payload: the MI355X form of stable_merge_sort_by_key
// (merge_sort_by_key.cu:150-209) for keys whose order is a bit order.
//
// The reference sorts tiles of 896 elements and then merges runs pairwise,
// log2(n / 896) passes over the whole array (17 at 2^26).  A radix sort over
// 8-bit digits reads and writes the array 4 times, whatever n is.  Each pass
// p (digit p, least significant first) is three launches:
//   upsweep   every tile counts its keys per digit into counts[d * T + t]
//             (digit-major), reading the keys only;
//   scan      an exclusive scan of counts[] turns every (digit, tile) count
//             into the global position of the tile's first key of that digit
//             (the batch executor keeps the counts tile-major instead, scans
//             each digit's row over the tiles, radix_scan_rows, and adds the
//             smaller digits' totals in the downsweep);
//   downsweep every tile ranks its keys by digit in input order (stable),
//             reorders keys and values through LDS into digit order, and
//             writes each digit's run to its global position -- consecutive
//             lanes write consecutive addresses.
//
// Key image: an order-preserving unsigned image of the key under `less`
// (int32: sign bit flipped; float32: sign-magnitude to two's complement
// order, with -0.0 folded onto +0.0 so the two keep their input order, as
// they compare equal); `greater` sorts by the complemented image.  Keys that
// compare equal have one image, distinct keys distinct images in key order,
// so the output equals a stable comparison sort's bit for bit.  NaN keys are
// not ordered by `less` (std::stable_sort's result is unspecified for them);
// here they land by their bits: positive NaNs after +inf, negative NaNs
// before -inf.
//
// Ranking (the agent-level counterpart of the reference's odd-even grain
// sort and merge-path rounds, sort.hpp:33-149, stable_merge_sort.hpp:52-206):
// a wavefront owns a chunk of 64 * ITEMS keys loaded row by row, key
// (row i, lane l) = chunk[i * 64 + l], so rows in order and lanes in order
// are input order.  For one row, eight ballots of the digit's bits give
// every lane the mask of lanes holding its digit; its rank is the popcount
// of that mask below it plus the wave's running count of the digit (an LDS
// counter that the digit's lowest lane advances).  Wave prefixes per digit
// and the tile's digit starts follow from the counters.
#pragma once

#include "device_common.hpp"

namespace bulkx {

constexpr int kRadixBits = 8;
constexpr int kRadixDigits = 1 << kRadixBits;
constexpr int kRadixPasses = 32 / kRadixBits;

// Tile shape of the radix passes: 512 lanes x 16 keys = 8192 keys.
template <class V>
struct RadixConfig {
  static constexpr int BLOCK = 512;
  static constexpr int ITEMS = 16;
  static constexpr int TILE = BLOCK * ITEMS;
};
constexpr int64_t kRadixTile = 8192;
static_assert(RadixConfig<uint32_t>::TILE == kRadixTile, "one tile size for every payload");

template <class K>
__device__ __forceinline__ uint32_t radix_image(K k) {
  if constexpr (std::is_same<K, uint32_t>::value) {
    return k;
  } else if constexpr (std::is_same<K, int32_t>::value) {
    return uint32_t(k) ^ 0x80000000u;
  } else {
    static_assert(std::is_same<K, float>::value, "4-byte keys");
    uint32_t b = bitcast<uint32_t>(k);
    if (b == 0x80000000u) b = 0u;  // -0.0 sorts as +0.0
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  }
}

template <class K, bool GREATER>
__device__ __forceinline__ uint32_t radix_digit(K k, int shift) {
  uint32_t u = radix_image(k);
  if constexpr (GREATER) u = ~u;
  return (u >> shift) & uint32_t(kRadixDigits - 1);
}

template <class K, class V, bool HAS_V, int BLOCK, int ITEMS>
struct RadixSmem {
  static constexpr int NW = BLOCK / kWave;
  static constexpr int TILE = BLOCK * ITEMS;
  static constexpr int cmax(int a, int b) { return a > b ? a : b; }
  // 4-byte keys with a 4-byte payload stage as one 8-byte element each (one
  // round through LDS instead of two)
  static constexpr bool PACK = HAS_V && sizeof(K) == 4 && sizeof(V) == 4;
  static constexpr int STAGE =
      TILE * (PACK ? 8 : cmax(int(sizeof(K)), HAS_V ? int(sizeof(V)) : 0));
  __align__(16) char stage[STAGE];          // keys, then values (or both packed), in tile digit order
  uint32_t wcnt[NW][kRadixDigits];          // per wave: running count, then wave prefix
  uint32_t tstart[kRadixDigits];            // tile position of each digit's first key
  uint32_t delta[kRadixDigits];             // global position - tile position, per digit
  uint32_t tcount[kRadixDigits];
  uint32_t gtotal[kRadixDigits];            // row-scanned counts: the sort's keys per digit
};

// Tile of workgroup b when every XCD takes a contiguous range of tiles:
// workgroups are dispatched to the 8 XCDs round robin (b % 8), and a tile's
// digit runs share their edge cache lines with the neighbouring tiles' runs
// of the same digits, so neighbours on one XCD merge those lines in its L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t tiles) {
  const int64_t x = b % kXcds, i = b / kXcds, q = tiles / kXcds, r = tiles % kXcds;
  return x * q + (x < r ? x : r) + i;
}

// Histogram copies of the upsweep: lane l counts into copy l % 32, so one
// LDS atomic instruction puts at most two lanes on any counter, whatever
// the digits are.  Rows are 257 words apart, so the copies of one digit
// fall in different banks.  With one histogram per wavefront, keys whose
// digit is mostly one value (float32 in [0, 1) in the top byte, keys with
// few distinct values, presorted runs) serialise up to 64 lanes on one
// counter: 2^26 keys count in 44 us per pass with the copies on every
// distribution, against 132 us (float32 in [0, 1), top byte) to 238 us
// (all keys equal) with one histogram per wavefront.
constexpr int kRadixCountCopies = 32;
constexpr int kRadixCountStride = kRadixDigits + 1;
constexpr int kRadixCountWords = kRadixCountCopies * kRadixCountStride;

// Upsweep of tile `tile`: counts[d * num_tiles + tile] = keys of digit d in
// the tile (digit-major), or with TILE_MAJOR counts[tile * 256 + d] (each
// tile's 256 counts one coalesced 1 KB row).  s_hist holds
// kRadixCountWords words.
template <class K, bool GREATER, int BLOCK, int ITEMS, bool TILE_MAJOR = false>
__device__ __forceinline__ void radix_count_tile(uint32_t* s_hist, const K* keys, int64_t n,
                                                 int shift, uint32_t* counts, int64_t num_tiles,
                                                 int64_t tile) {
  constexpr int TILE = BLOCK * ITEMS;
  using SL = Striped<K, ITEMS>;
  const int wid = threadIdx.x / kWave;
  const int lane = lane_id();
  for (int i = threadIdx.x; i < kRadixCountWords; i += BLOCK) s_hist[i] = 0;
  K x[ITEMS];
  const int64_t cb = tile * TILE + int64_t(wid) * SL::N;
  const int64_t rem = n - cb;
  const int count = rem <= 0 ? 0 : (rem >= SL::N ? SL::N : int(rem));
  striped_load<K, ITEMS, SL::VE, true>(keys + (count > 0 ? cb : 0), count, x, K());
  __syncthreads();
  uint32_t* h = s_hist + (lane % kRadixCountCopies) * kRadixCountStride;
  if (count == SL::N) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) atomicAdd(&h[radix_digit<K, GREATER>(x[i], shift)], 1u);
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if (SL::offset(i, lane) < count) atomicAdd(&h[radix_digit<K, GREATER>(x[i], shift)], 1u);
  }
  __syncthreads();
  static_assert(BLOCK >= kRadixDigits, "one lane per digit");
  if (threadIdx.x < kRadixDigits && tile < num_tiles) {
    const int d = threadIdx.x;
    uint32_t c = 0;
#pragma unroll
    for (int r = 0; r < kRadixCountCopies; ++r) c += s_hist[r * kRadixCountStride + d];
    if constexpr (TILE_MAJOR) counts[tile * kRadixDigits + d] = c;
    else counts[int64_t(d) * num_tiles + tile] = c;
  }
}

// Lanes of the wavefront holding digit d (8-bit): the ballots of the
// digit's bits, each compared with the lane's own bit (s = 0 or all ones
// from the bit: lanes that agree with this lane in bit b are ~(ballot ^ s)),
// ANDed in two 32-bit halves, so each bit costs one extract, one compare and
// two three-input bit operations (v_bitop3 on gfx950).
__device__ __forceinline__ uint64_t radix_peers(uint32_t d, uint64_t active) {
  uint32_t lo = uint32_t(active), hi = uint32_t(active >> 32);
#pragma unroll
  for (int b = 0; b < kRadixBits; ++b) {
    const uint32_t sb = uint32_t(-int32_t((d >> b) & 1u));
    const uint64_t m = __ballot(sb != 0u);
    lo &= ~(uint32_t(m) ^ sb);
    hi &= ~(uint32_t(m >> 32) ^ sb);
  }
  return (uint64_t(hi) << 32) | lo;
}

// Downsweep of tile `tile`: keys (and values) of the tile to their global
// positions, offsets[d * num_tiles + tile] being the global position of the
// tile's first key of digit d.  With `digit_totals` the offsets are instead
// tile-major and row-scanned -- offsets[tile * 256 + d] counts the keys of
// digit d in the tiles before this one -- and the keys of smaller digits
// (the exclusive scan of digit_totals) are added here.  FULL tiles (all but
// the last) run without per-element bounds.
template <class K, class V, bool HAS_V, bool GREATER, int BLOCK, int ITEMS, bool FULL>
__device__ __forceinline__ void radix_scatter_tile_impl(RadixSmem<K, V, HAS_V, BLOCK, ITEMS>& sm,
                                                        const K* __restrict__ kin,
                                                        const V* __restrict__ vin, int64_t n,
                                                        int shift,
                                                        const uint32_t* __restrict__ offsets,
                                                        int64_t num_tiles, K* __restrict__ kout,
                                                        V* __restrict__ vout, int64_t tile,
                                                        const uint32_t* __restrict__ digit_totals) {
  constexpr int NW = BLOCK / kWave;
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int CHUNK = kWave * ITEMS;
  static_assert(BLOCK >= kRadixDigits, "one lane per digit in the prefix step");
  const int wid = threadIdx.x / kWave;
  const int lane = lane_id();
  const int64_t tbase = tile * TILE;
  const int tcnt = FULL ? TILE : int(n - tbase < TILE ? n - tbase : TILE);
  const int cb = wid * CHUNK;  // the wave's chunk within the tile
  const int ccount = FULL ? CHUNK : (tcnt - cb <= 0 ? 0 : (tcnt - cb >= CHUNK ? CHUNK : tcnt - cb));

  // the global base of each digit: its load is in flight during the ranking
  uint32_t gbase = 0, gtotal = 0;
  if (threadIdx.x < kRadixDigits) {
    if (digit_totals) {
      gbase = offsets[tile * kRadixDigits + threadIdx.x];
      gtotal = digit_totals[threadIdx.x];
    } else {
      gbase = offsets[int64_t(threadIdx.x) * num_tiles + tile];
    }
  }
  for (int i = threadIdx.x; i < NW * kRadixDigits; i += BLOCK) (&sm.wcnt[0][0])[i] = 0;

  // rows of the wave's chunk: key (i, lane) = chunk[i * 64 + lane]; values
  // load with the keys and wait in registers
  K k[ITEMS];
  V v[ITEMS];
  const K* kc = kin + tbase + cb + lane;
  const V* vc = HAS_V ? vin + tbase + cb + lane : nullptr;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (FULL) k[i] = __builtin_nontemporal_load(kc + i * kWave);
    else k[i] = i * kWave + lane < ccount ? __builtin_nontemporal_load(kc + i * kWave) : K();
  }
  if constexpr (HAS_V) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if (FULL) v[i] = __builtin_nontemporal_load(vc + i * kWave);
      else v[i] = i * kWave + lane < ccount ? __builtin_nontemporal_load(vc + i * kWave) : V();
    }
  }
  __syncthreads();  // counters cleared

  // rank within the wave, row by row
  uint32_t pos[ITEMS];  // rank among the wave's keys of the same digit
  uint32_t* wc = sm.wcnt[wid];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const bool valid = FULL || i * kWave + lane < ccount;
    const uint32_t d = radix_digit<K, GREATER>(k[i], shift);
    const uint64_t peers = radix_peers(d, FULL ? ~uint64_t(0) : __ballot(valid));
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(
        uint32_t(peers >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(peers), 0u));
    const uint32_t prev = wc[d];
    pos[i] = prev + below;
    if (valid && below == 0) wc[d] = prev + uint32_t(__popcll(peers));
  }
  __syncthreads();

  // per digit: wave prefixes and the tile's count
  if (threadIdx.x < kRadixDigits) {
    const int d = threadIdx.x;
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t c = sm.wcnt[w][d];
      sm.wcnt[w][d] = run;
      run += c;
    }
    sm.tcount[d] = run;
    sm.delta[d] = gbase;
    sm.gtotal[d] = gtotal;
  }
  __syncthreads();
  // the tile's digit starts (exclusive scan over 256 digits, 4 per lane of wave 0)
  if (wid == 0) {
    const uint32_t c0 = sm.tcount[4 * lane], c1 = sm.tcount[4 * lane + 1];
    const uint32_t c2 = sm.tcount[4 * lane + 2], c3 = sm.tcount[4 * lane + 3];
    const uint32_t inc = wave_inclusive_scan<OpPlus>(c0 + c1 + c2 + c3);
    const uint32_t e0 = inc - (c0 + c1 + c2 + c3);
    const uint32_t s0 = e0, s1 = e0 + c0, s2 = s1 + c1, s3 = s2 + c2;
    sm.tstart[4 * lane] = s0;
    sm.tstart[4 * lane + 1] = s1;
    sm.tstart[4 * lane + 2] = s2;
    sm.tstart[4 * lane + 3] = s3;
    sm.delta[4 * lane] -= s0;
    sm.delta[4 * lane + 1] -= s1;
    sm.delta[4 * lane + 2] -= s2;
    sm.delta[4 * lane + 3] -= s3;
    if (digit_totals) {  // keys of the smaller digits, over the whole sort
      const uint32_t g0 = sm.gtotal[4 * lane], g1 = sm.gtotal[4 * lane + 1];
      const uint32_t g2 = sm.gtotal[4 * lane + 2], g3 = sm.gtotal[4 * lane + 3];
      const uint32_t ginc = wave_inclusive_scan<OpPlus>(g0 + g1 + g2 + g3);
      const uint32_t b0 = ginc - (g0 + g1 + g2 + g3), b1 = b0 + g0, b2 = b1 + g1, b3 = b2 + g2;
      sm.delta[4 * lane] += b0;
      sm.delta[4 * lane + 1] += b1;
      sm.delta[4 * lane + 2] += b2;
      sm.delta[4 * lane + 3] += b3;
    }
  }
  __syncthreads();

  if constexpr (RadixSmem<K, V, HAS_V, BLOCK, ITEMS>::PACK) {
    // keys and values to their tile positions in digit order, packed; then
    // each digit's run to its global position (lane t of the group writes
    // tile positions t, t + BLOCK, ...)
    uint64_t* sp = reinterpret_cast<uint64_t*>(sm.stage);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if (FULL || i * kWave + lane < ccount) {
        const uint32_t d = radix_digit<K, GREATER>(k[i], shift);
        pos[i] += sm.tstart[d] + wc[d];
        sp[pos[i]] = (uint64_t(bitcast<uint32_t>(v[i])) << 32) | bitcast<uint32_t>(k[i]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int j = i * BLOCK + int(threadIdx.x);
      if (FULL || j < tcnt) {
        const uint64_t e = sp[j];
        const K key = bitcast<K>(uint32_t(e));
        const uint32_t dst = uint32_t(j) + sm.delta[radix_digit<K, GREATER>(key, shift)];
        if (dst < n) {  // always true when the counts are consistent
          kout[dst] = key;
          vout[dst] = bitcast<V>(uint32_t(e >> 32));
        }
      }
    }
    return;
  }
  // keys to their tile positions in digit order
  K* sk = reinterpret_cast<K*>(sm.stage);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (FULL || i * kWave + lane < ccount) {
      const uint32_t d = radix_digit<K, GREATER>(k[i], shift);
      pos[i] += sm.tstart[d] + wc[d];
      sk[pos[i]] = k[i];
    }
  }
  __syncthreads();
  // each digit's run to its global position: lane t of the group writes
  // tile positions t, t + BLOCK, ...
  uint32_t dst[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const int j = i * BLOCK + int(threadIdx.x);
    if (FULL || j < tcnt) {
      const K key = sk[j];
      dst[i] = uint32_t(j) + sm.delta[radix_digit<K, GREATER>(key, shift)];
      if (dst[i] < n) kout[dst[i]] = key;  // always true when the counts are consistent
    }
  }
  if constexpr (HAS_V) {
    V* sv = reinterpret_cast<V*>(sm.stage);
    __syncthreads();  // key reads of the stage are done
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if (FULL || i * kWave + lane < ccount) sv[pos[i]] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int j = i * BLOCK + int(threadIdx.x);
      if ((FULL || j < tcnt) && dst[i] < n) vout[dst[i]] = sv[j];
    }
  }
}

template <class K, class V, bool HAS_V, bool GREATER, int BLOCK, int ITEMS>
__device__ __forceinline__ void radix_scatter_tile(RadixSmem<K, V, HAS_V, BLOCK, ITEMS>& sm,
                                                   const K* __restrict__ kin,
                                                   const V* __restrict__ vin, int64_t n, int shift,
                                                   const uint32_t* __restrict__ offsets,
                                                   int64_t num_tiles, K* __restrict__ kout,
                                                   V* __restrict__ vout, int64_t tile,
                                                   const uint32_t* __restrict__ digit_totals = nullptr) {
  if ((tile + 1) * int64_t(BLOCK * ITEMS) <= n)
    radix_scatter_tile_impl<K, V, HAS_V, GREATER, BLOCK, ITEMS, true>(
        sm, kin, vin, n, shift, offsets, num_tiles, kout, vout, tile, digit_totals);
  else
    radix_scatter_tile_impl<K, V, HAS_V, GREATER, BLOCK, ITEMS, false>(
        sm, kin, vin, n, shift, offsets, num_tiles, kout, vout, tile, digit_totals);
}

// Row scan of tile-major counts (counts[t * 256 + d], T tiles): digits
// [16 * group, 16 * group + 16) of one sort, by one workgroup of
// 16 * SLICES lanes.  Lane (s, dl) = (tid / 16, tid % 16) sums digit
// 16 * group + dl over its SLICES-th of the tiles (each wavefront load
// instruction reads four tiles' 64-byte runs of the group's 16 counts), the
// slices' sums are scanned in LDS, and a second sweep replaces every count
// by the keys of that digit in the tiles before it; totals[d] receives the
// digit's keys over the sort.  A lane keeps 32 loads in flight.  One
// 2^26-key sort (8192 tiles) scans in 61 us per pass with 16 slices (72 us
// with 8 loads in flight); 64 slices of 1024 lanes take it to 34 us but
// double the time of a batch of 311 sorts of 2^13..2^22 keys (26 -> 55 us),
// whose small jobs leave most of those lanes idle.
template <int SLICES>
__device__ __forceinline__ void radix_scan_rows(uint32_t* __restrict__ counts, int64_t T,
                                                uint32_t* __restrict__ totals, int group,
                                                uint32_t (*s_part)[16] /* [SLICES][16] */) {
  const int s = threadIdx.x / 16, dl = threadIdx.x % 16;
  const int d = group * 16 + dl;
  const int64_t lo = T * s / SLICES, hi = T * (s + 1) / SLICES;
  constexpr int U = 32;
  uint32_t sum = 0;
  for (int64_t t0 = lo; t0 < hi; t0 += U) {
    uint32_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = t0 + u < hi ? counts[(t0 + u) * kRadixDigits + d] : 0u;
#pragma unroll
    for (int u = 0; u < U; ++u) sum += c[u];
  }
  s_part[s][dl] = sum;
  __syncthreads();
  uint32_t run = 0, total = 0;
#pragma unroll 16
  for (int q = 0; q < SLICES; ++q) {
    const uint32_t p = s_part[q][dl];
    run += q < s ? p : 0u;
    total += p;
  }
  if (s == 0) totals[d] = total;
  for (int64_t t0 = lo; t0 < hi; t0 += U) {
    uint32_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = t0 + u < hi ? counts[(t0 + u) * kRadixDigits + d] : 0u;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (t0 + u < hi) counts[(t0 + u) * kRadixDigits + d] = run;
      run += c[u];
    }
  }
}

}  // namespace bulkx

// ---------------------------------------------------------------------------
// Device-wide radix sort (the plain entry point)
// ---------------------------------------------------------------------------
namespace bulkx {

// One upsweep workgroup per tile, tiles XCD-contiguous (xcd_tile): the
// digit-major counts of neighbouring tiles share 64-byte lines, and a line
// whose counts all come from one XCD's L2 leaves it as one full write
// instead of one partial write per XCD that touched it (2^26 keys: 59 us
// per pass with the plain block order, 43 us with this one).
template <class K, bool GREATER, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void radix_upsweep_kernel(const K* __restrict__ keys,
                                                              int64_t n, int shift,
                                                              uint32_t* __restrict__ counts,
                                                              int64_t num_tiles) {
  __shared__ uint32_t s_hist[kRadixCountWords];
  radix_count_tile<K, GREATER, BLOCK, ITEMS>(s_hist, keys, n, shift, counts, num_tiles,
                                             xcd_tile(blockIdx.x, num_tiles));
}

template <class K, class V, bool HAS_V, bool GREATER, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void radix_downsweep_kernel(
    const K* __restrict__ kin, const V* __restrict__ vin, int64_t n, int shift,
    const uint32_t* __restrict__ offsets, int64_t num_tiles, K* __restrict__ kout,
    V* __restrict__ vout) {
  __shared__ RadixSmem<K, V, HAS_V, BLOCK, ITEMS> sm;
  radix_scatter_tile<K, V, HAS_V, GREATER, BLOCK, ITEMS>(sm, kin, vin, n, shift, offsets,
                                                         num_tiles, kout, vout,
                                                         xcd_tile(blockIdx.x, num_tiles));
}

// Workspace of a radix sort: ping-pong keys and values, the digit counts of
// every tile, and the counts scan's own look-back space.
struct RadixWorkspace {
  size_t keys, vals, counts, scan, total;
  RadixWorkspace(size_t ks, size_t vs, int64_t n) {
    const int64_t tiles = (n + kRadixTile - 1) / kRadixTile;
    keys = 0;
    vals = align_up(size_t(n) * ks, 256);
    counts = vals + align_up(size_t(n) * vs, 256);
    scan = counts + align_up(size_t(tiles) * kRadixDigits * sizeof(uint32_t), 256);
    total = scan + scan_workspace_bytes(BULKX_UINT32, tiles * kRadixDigits);
  }
};

// Largest n the radix form takes (global positions are 32-bit).
constexpr int64_t kRadixMaxN = (int64_t(1) << 32) - 1;

template <class CMP, class K, class V, bool HAS_V>
int radix_sort_launch(K* keys, V* vals, int64_t n, char* ws, hipStream_t stream) {
  static_assert(sizeof(K) == 4, "radix passes sort 4-byte keys");
  constexpr bool G = std::is_same<CMP, CmpGreater>::value;
  using C = RadixConfig<V>;
  const int64_t tiles = (n + C::TILE - 1) / C::TILE;
  const RadixWorkspace w(sizeof(K), HAS_V ? sizeof(V) : 0, n);
  K* tk = reinterpret_cast<K*>(ws + w.keys);
  V* tv = HAS_V ? reinterpret_cast<V*>(ws + w.vals) : nullptr;
  uint32_t* counts = reinterpret_cast<uint32_t*>(ws + w.counts);
  char* scan_ws = ws + w.scan;
  K* sk = keys;
  V* sv = vals;
  K* dk = tk;
  V* dv = tv;
  for (int p = 0; p < kRadixPasses; ++p) {
    const int shift = p * kRadixBits;
    radix_upsweep_kernel<K, G, C::BLOCK, C::ITEMS><<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(
        sk, n, shift, counts, tiles);
    BULKX_LAUNCHED();
    BULKX_TRY(exclusive_scan_u32(counts, counts, tiles * kRadixDigits, scan_ws, stream));
    radix_downsweep_kernel<K, V, HAS_V, G, C::BLOCK, C::ITEMS>
        <<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(sk, sv, n, shift, counts, tiles, dk, dv);
    BULKX_LAUNCHED();
    K* t1 = sk;
    sk = dk;
    dk = t1;
    V* t2 = sv;
    sv = dv;
    dv = t2;
  }
  // an even number of passes ends in the caller's arrays
  static_assert(kRadixPasses % 2 == 0, "the last pass writes the caller's arrays");
  return BULKX_SUCCESS;
}

}  // namespace bulkx
