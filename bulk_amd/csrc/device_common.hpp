// Device-side building blocks shared by every kernel of the library:
// operator functors, wavefront (64-lane) DPP scans and reductions, the
// decoupled look-back tile-status protocol, and vectorised tile loads.
//
// Target: gfx950 (CDNA4), wave64.  Nothing here is portable on purpose.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits>
#include <type_traits>

namespace bulkx {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Operators.  Integer addition wraps (computed in the unsigned type), which
// makes it associative and lets any evaluation order be bit-exact.
// min / max follow thrust::minimum / thrust::maximum: lhs < rhs ? lhs : rhs.
// ---------------------------------------------------------------------------
struct OpPlus {
  template <class T>
  __host__ __device__ __forceinline__ static T apply(T a, T b) {
    if constexpr (std::is_integral_v<T>) {
      using U = std::make_unsigned_t<T>;
      return T(U(a) + U(b));
    } else {
      return a + b;
    }
  }
  // -0.0 is the exact additive identity for IEEE floats (x + -0.0 == x for every x).
  template <class T>
  __host__ __device__ __forceinline__ static T identity() {
    if constexpr (std::is_floating_point_v<T>) return T(-0.0);
    else return T(0);
  }
};

struct OpMin {
  template <class T>
  __host__ __device__ __forceinline__ static T apply(T a, T b) { return a < b ? a : b; }
  template <class T>
  __host__ __device__ __forceinline__ static T identity() {
    if constexpr (std::is_floating_point_v<T>) return std::numeric_limits<T>::infinity();
    else return std::numeric_limits<T>::max();
  }
};

struct OpMax {
  template <class T>
  __host__ __device__ __forceinline__ static T apply(T a, T b) { return a < b ? b : a; }
  template <class T>
  __host__ __device__ __forceinline__ static T identity() {
    if constexpr (std::is_floating_point_v<T>) return -std::numeric_limits<T>::infinity();
    else return std::numeric_limits<T>::lowest();
  }
};

// ---------------------------------------------------------------------------
// Double-double accumulator: value = hi + lo with |lo| <= ulp(hi)/2.  Carries
// float64 sums across threads, tiles and the look-back chain so that the
// error of a float64 scan or reduction stays within a few ulps of sum|x|
// however long the chain of tiles is.  Addition is Knuth's TwoSum followed by
// renormalisation (Dekker, 1971; Hida, Li and Bailey's "sloppy" add).
// Non-finite sums short-circuit so inf stays inf and NaN stays NaN.
// ---------------------------------------------------------------------------
struct dd {
  double hi, lo;
};

__host__ __device__ __forceinline__ dd dd_add(dd a, dd b) {
#pragma clang fp contract(off)
  double s = a.hi + b.hi;
  if (!__builtin_isfinite(s)) return dd{s, 0.0};
  double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b.hi - bb);
  e += a.lo + b.lo;
  double hi = s + e;
  double lo = e - (hi - s);
  return dd{hi, lo};
}

struct OpPlusDD {
  __host__ __device__ __forceinline__ static dd apply(dd a, dd b) { return dd_add(a, b); }
  template <class T>
  __host__ __device__ __forceinline__ static dd identity() { return dd{-0.0, 0.0}; }
};

// Accumulator policy: which type and operator carry partial results between
// threads and tiles.  float sums accumulate in double, double sums in
// double-double; everything else (integers, min, max) stays in T and is exact.
//
// Two levels: `row` (type row_t, operator row_op) is used for the short
// wavefront-level scans (depth ~ log2 of a wave chunk), `type` / `op` for
// the carries from row to row, across waves, tiles and the look-back chain,
// whose depth grows with n.  lift() moves a row value to the carry level and
// finish(p, r) returns the element value p op r in T.
template <class T, class Op>
struct Acc {
  using type = T;
  using op = Op;
  using row_t = T;
  using row_op = Op;
  __host__ __device__ __forceinline__ static type from(T x) { return x; }
  __host__ __device__ __forceinline__ static T to(type a) { return a; }
  __host__ __device__ __forceinline__ static row_t row(T x) { return x; }
  __host__ __device__ __forceinline__ static type lift(row_t r) { return r; }
  __host__ __device__ __forceinline__ static T finish(type p, row_t r) { return Op::apply(p, r); }
};
template <>
struct Acc<float, OpPlus> {
  using type = double;
  using op = OpPlus;
  using row_t = double;
  using row_op = OpPlus;
  __host__ __device__ __forceinline__ static double from(float x) { return double(x); }
  __host__ __device__ __forceinline__ static float to(double a) { return float(a); }
  __host__ __device__ __forceinline__ static double row(float x) { return double(x); }
  __host__ __device__ __forceinline__ static double lift(double r) { return r; }
  __host__ __device__ __forceinline__ static float finish(double p, double r) { return float(p + r); }
};
template <>
struct Acc<double, OpPlus> {
  using type = dd;
  using op = OpPlusDD;
  using row_t = double;
  using row_op = OpPlus;
  __host__ __device__ __forceinline__ static dd from(double x) { return dd{x, 0.0}; }
  __host__ __device__ __forceinline__ static double to(dd a) {
    return __builtin_isfinite(a.hi) ? a.hi + a.lo : a.hi;
  }
  __host__ __device__ __forceinline__ static double row(double x) { return x; }
  __host__ __device__ __forceinline__ static dd lift(double r) { return dd{r, 0.0}; }
  // p.hi + (p.lo + r): one rounding of the low part and one of the result.
  __host__ __device__ __forceinline__ static double finish(dd p, double r) {
#pragma clang fp contract(off)
    return __builtin_isfinite(p.hi) ? p.hi + (p.lo + r) : p.hi + r;
  }
};

// The order of the comparators.  Integers: <.  Floats: the IEEE order for
// non-NaN values (-0.0 and +0.0 compare equal); a NaN, which < leaves
// unordered, orders by its bits -- before -inf when its sign bit is set,
// after +inf otherwise -- the order of the radix sort's key image, so the
// merge, the comparison sorts and the radix sort agree on inputs holding NaNs.
template <class U, class F>
__host__ __device__ __forceinline__ U float_order_image(F x) {
  constexpr U sign = U(1) << (8 * sizeof(U) - 1);
  U b = __builtin_bit_cast(U, x);
  if (b == sign) b = 0;  // -0.0 orders as +0.0
  return (b & sign) ? U(~b) : U(b | sign);
}
template <class T>
__host__ __device__ __forceinline__ bool order_less(T a, T b) {
  if constexpr (std::is_same_v<T, float>) {
    if (__builtin_expect(a != a || b != b, 0))
      return float_order_image<uint32_t>(a) < float_order_image<uint32_t>(b);
  } else if constexpr (std::is_same_v<T, double>) {
    if (__builtin_expect(a != a || b != b, 0))
      return float_order_image<uint64_t>(a) < float_order_image<uint64_t>(b);
  }
  return a < b;
}

struct CmpLess {
  template <class T>
  __host__ __device__ __forceinline__ static bool apply(T a, T b) { return order_less(a, b); }
};
struct CmpGreater {
  template <class T>
  __host__ __device__ __forceinline__ static bool apply(T a, T b) { return order_less(b, a); }
};

// ---------------------------------------------------------------------------
// Raw bit access.
// ---------------------------------------------------------------------------
template <class To, class From>
__device__ __forceinline__ To bitcast(From x) {
  static_assert(sizeof(To) == sizeof(From), "size mismatch");
  return __builtin_bit_cast(To, x);
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// ---------------------------------------------------------------------------
// DPP moves on any trivially copyable type whose size is a multiple of 4:
// each 32-bit chunk goes through update_dpp(old, src, ctrl, row_mask,
// bank_mask, bound_ctrl=false); lanes whose source is out of range, or whose
// row is masked off, keep `old`.
// ---------------------------------------------------------------------------
template <int N>
struct Chunks {
  uint32_t w[N];
};

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf, class T>
__device__ __forceinline__ T dpp_move(T old, T src) {
  static_assert(sizeof(T) % 4 == 0, "dpp_move needs a size that is a multiple of 4");
  constexpr int N = sizeof(T) / 4;
  auto o = bitcast<Chunks<N>>(old);
  auto s = bitcast<Chunks<N>>(src);
  Chunks<N> r;
#pragma unroll
  for (int i = 0; i < N; ++i)
    r.w[i] = uint32_t(__builtin_amdgcn_update_dpp(int(o.w[i]), int(s.w[i]), CTRL, ROW_MASK,
                                                  BANK_MASK, false));
  return bitcast<T>(r);
}

// DPP control codes (GFX9 encoding).
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kWaveShl1 = 0x130, kWaveShr1 = 0x138;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;

template <class T>
__device__ __forceinline__ T readlane(T x, int lane) {
  constexpr int N = sizeof(T) / 4;
  auto c = bitcast<Chunks<N>>(x);
#pragma unroll
  for (int i = 0; i < N; ++i) c.w[i] = uint32_t(__builtin_amdgcn_readlane(int(c.w[i]), lane));
  return bitcast<T>(c);
}

// A value every lane of the wavefront holds (e.g. a tile index broadcast
// through LDS), moved to a scalar register so that everything derived from it
// (job records, base addresses) is computed once per wave on the scalar unit.
__device__ __forceinline__ uint32_t wave_uniform(uint32_t x) {
  return uint32_t(__builtin_amdgcn_readfirstlane(int(x)));
}

// Inclusive scan across the 64 lanes of a wavefront: 4 row shifts inside
// each 16-lane row, then two row broadcasts join the rows.
template <class Op, class T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
  const T id = Op::template identity<T>();
  x = Op::apply(dpp_move<kRowShr1>(id, x), x);
  x = Op::apply(dpp_move<kRowShr2>(id, x), x);
  x = Op::apply(dpp_move<kRowShr4>(id, x), x);
  x = Op::apply(dpp_move<kRowShr8>(id, x), x);
  x = Op::apply(dpp_move<kRowBcast15, 0xa>(id, x), x);
  x = Op::apply(dpp_move<kRowBcast31, 0xc>(id, x), x);
  return x;
}

// Shift a wave-wide inclusive scan right by one lane (lane 0 gets identity),
// turning it into the exclusive scan.
template <class Op, class T>
__device__ __forceinline__ T wave_shift_right(T x) {
  return dpp_move<kWaveShr1>(Op::template identity<T>(), x);
}

// Reduction over the wavefront; every lane receives the total.
template <class Op, class T>
__device__ __forceinline__ T wave_reduce(T x) {
  return readlane(wave_inclusive_scan<Op>(x), kWave - 1);
}

// ---------------------------------------------------------------------------
// Block-level helpers (blockDim.x == BLOCK, BLOCK a multiple of 64).
// ---------------------------------------------------------------------------
template <int BLOCK, class Op, class T>
__device__ __forceinline__ T block_reduce(T x, T* s_wave /* BLOCK/64 */) {
  constexpr int NW = BLOCK / kWave;
  T w = wave_reduce<Op>(x);
  const int wid = threadIdx.x / kWave;
  if (lane_id() == 0) s_wave[wid] = w;
  __syncthreads();
  T r = s_wave[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r = Op::apply(r, s_wave[i]);
  __syncthreads();
  return r;
}

// Exclusive block scan of one value per thread.  Returns the thread's
// exclusive prefix; *total receives the block aggregate.
template <int BLOCK, class Op, class T>
__device__ __forceinline__ T block_exclusive_scan(T x, T* s_wave, T* total) {
  constexpr int NW = BLOCK / kWave;
  const int wid = threadIdx.x / kWave;
  T inc = wave_inclusive_scan<Op>(x);
  T exc = wave_shift_right<Op>(inc);
  if (lane_id() == kWave - 1) s_wave[wid] = inc;
  __syncthreads();
  T wp = Op::template identity<T>();
  T agg = Op::template identity<T>();
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    T s = s_wave[i];
    if (i < wid) wp = Op::apply(wp, s);
    agg = Op::apply(agg, s);
  }
  __syncthreads();
  *total = agg;
  return Op::apply(wp, exc);
}

// ---------------------------------------------------------------------------
// Agent-scope relaxed atomics (coherent across the 8 XCDs' L2 slices).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Decoupled look-back tile status.
//
// Each tile owns W = sizeof(T)/4 64-bit status words.  Word j holds bits
// [32j, 32j+32) of the value in its low half and the status kind in its
// high half.  Kinds: 0 = not yet published, 1 = tile aggregate,
// 2 = inclusive prefix.  A reader accepts a tile only when all of its words
// carry the same non-zero kind, so a value is never torn across words
// without any memory fence.  Words are cleared by the preparation kernel of
// each call.
// ---------------------------------------------------------------------------
constexpr uint32_t kStatusInvalid = 0, kStatusAggregate = 1, kStatusPrefix = 2;

// Bounded spin: a wave that waits this many polls for a predecessor gives up,
// flags the guard's error word and treats the predecessor as an identity
// prefix, so no kernel can hang forever on a protocol bug.  The host turns a
// set error word into BULKX_ERR_LOOKBACK (bulkx_stream_status,
// bulkx_batch_status), so a give-up is never reported as success.
constexpr uint32_t kSpinLimit = 1u << 22;

// Error word and spin limit shared by the look-back waits of a launch.
// spin_limit = 0 makes every look-back give up at once without reading its
// predecessors (a host-set knob for tests of the error path).
struct LookbackGuard {
  uint32_t err;
  uint32_t spin_limit;
  uint32_t taken;  // the error word as last read and cleared by the host
  uint32_t pad;
};

// Sleep between polls of a window whose statuses are not all published
// (units of 64 clocks).
#ifndef BULKX_LOOKBACK_SLEEP
#define BULKX_LOOKBACK_SLEEP 1
#endif

template <class T>
struct TileStatus {
  static constexpr int W = sizeof(T) / 4;
  uint64_t* words;  // W words per tile

  __device__ __forceinline__ void publish(int64_t tile, uint32_t kind, T value) const {
    auto c = bitcast<Chunks<W>>(value);
#pragma unroll
    for (int j = 0; j < W; ++j) st_agent(words + tile * W + j, (uint64_t(kind) << 32) | c.w[j]);
  }

  // Returns the kind (0 when not yet published or not yet consistent) and the value.
  __device__ __forceinline__ uint32_t read(int64_t tile, T* value) const {
    Chunks<W> c;
    uint32_t kind = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      uint64_t w = ld_agent(words + tile * W + j);
      uint32_t k = uint32_t(w >> 32);
      if (j == 0) kind = k;
      else if (k != kind) kind = kStatusInvalid;
      c.w[j] = uint32_t(w);
    }
    *value = bitcast<T>(c);
    return kind;
  }
};

// Wave-parallel look-back run by one full wavefront over windows of 64*Q
// predecessors: lane l inspects the Q consecutive tiles starting at
// (pos - 64*Q + 1 + l*Q), so lane order is tile order (oldest on the left)
// and non-commutative operators combine correctly.  The newest inclusive
// prefix in a window ends the walk; the aggregates after it are folded in.
// One window costs one memory round trip, so the inclusive-prefix front
// advances 64*Q tiles per round trip.  Returns the exclusive prefix of
// `tile` (tile >= 1).
template <class Op, int Q = 1, class T>
__device__ T lookback_exclusive_prefix(const TileStatus<T>& st, int64_t tile,
                                       LookbackGuard* guard) {
  constexpr int WIN = kWave * Q;
  const int lane = lane_id();
  const T id = Op::template identity<T>();
  const uint32_t limit = guard->spin_limit;
  if (limit == 0) {
    if (lane == 0) atomicOr(&guard->err, 1u);
    return id;
  }
  T acc = id;  // combination of every tile in (pos, tile)
  int64_t pos = tile - 1;  // newest tile of the current window
  while (true) {
    const int64_t first = pos - WIN + 1 + int64_t(lane) * Q;
    T val[Q];
    uint32_t kind[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      val[j] = id;
      kind[j] = kStatusPrefix;  // before tile 0: identity prefix
      if (first + j >= 0) kind[j] = st.read(first + j, &val[j]);
    }
    uint32_t spins = 0;
    while (true) {
      bool missing = false;
#pragma unroll
      for (int j = 0; j < Q; ++j) missing |= kind[j] == kStatusInvalid;
      if (!__any(missing)) break;
      __builtin_amdgcn_s_sleep(BULKX_LOOKBACK_SLEEP);
#pragma unroll
      for (int j = 0; j < Q; ++j)
        if (kind[j] == kStatusInvalid) kind[j] = st.read(first + j, &val[j]);
      if (++spins > limit) {
#pragma unroll
        for (int j = 0; j < Q; ++j) {
          if (kind[j] == kStatusInvalid) {
            atomicOr(&guard->err, 1u);
            kind[j] = kStatusPrefix;
            val[j] = id;
          }
        }
      }
    }
    // Newest prefix among this lane's Q tiles, and the fold from it onwards.
    int jp = -1;
#pragma unroll
    for (int j = 0; j < Q; ++j)
      if (kind[j] == kStatusPrefix) jp = j;
    T c = id;
#pragma unroll
    for (int j = 0; j < Q; ++j)
      if (j >= jp) c = Op::apply(c, val[j]);
    const uint64_t pmask = __ballot(jp >= 0);
    if (pmask) {
      const int p = 63 - __builtin_clzll(pmask);  // lane holding the newest prefix
      return Op::apply(wave_reduce<Op>(lane >= p ? c : id), acc);
    }
    acc = Op::apply(wave_reduce<Op>(c), acc);
    pos -= WIN;
  }
}

// Native 16-byte vector (the HIP uint4 is a struct; builtins want a vector).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Tile loads/stores: each thread owns ITEMS consecutive elements ("blocked"
// arrangement).  Full tiles with 16-byte aligned bases use dwordx4 accesses.
// ---------------------------------------------------------------------------
template <class T, int ITEMS>
__device__ __forceinline__ void load_blocked(const T* __restrict__ p, int64_t valid, T (&v)[ITEMS],
                                             T fill, bool full_aligned) {
  static_assert((ITEMS * sizeof(T)) % 16 == 0, "ITEMS*sizeof(T) must be a multiple of 16");
  if (full_aligned) {
    const u32x4* q = reinterpret_cast<const u32x4*>(p);
    constexpr int NV = ITEMS * sizeof(T) / 16;
    u32x4 tmp[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) tmp[i] = q[i];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      T* d = reinterpret_cast<T*>(&tmp[i]);
#pragma unroll
      for (int j = 0; j < int(16 / sizeof(T)); ++j) v[i * (16 / sizeof(T)) + j] = d[j];
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) v[i] = i < valid ? p[i] : fill;
  }
}

template <class T, int ITEMS>
__device__ __forceinline__ void store_blocked(T* __restrict__ p, int64_t valid, const T (&v)[ITEMS],
                                              bool full_aligned) {
  if (full_aligned) {
    u32x4* q = reinterpret_cast<u32x4*>(p);
    constexpr int NV = ITEMS * sizeof(T) / 16;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      u32x4 t;
      T* s = reinterpret_cast<T*>(&t);
#pragma unroll
      for (int j = 0; j < int(16 / sizeof(T)); ++j) s[j] = v[i * (16 / sizeof(T)) + j];
      q[i] = t;
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if (i < valid) p[i] = v[i];
  }
}

__device__ __forceinline__ bool aligned16(const void* p) {
  return (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

// Orders a wavefront's own LDS accesses (all LDS traffic of the wave has
// completed before any later access is issued).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// Striped register tiles.  A wavefront owns a contiguous chunk of 64*ITEMS
// elements viewed as NV rows of 64 16-byte vectors; lane l holds vector l of
// every row (x[j*VE + q] = chunk[(j*64 + l)*VE + q]).  Each load or store
// instruction of a wave therefore covers 1 KB of consecutive bytes, and no
// LDS is involved.  Scans over this layout combine a lane-local scan of each
// vector, a DPP scan across the row and a carry from row to row.
// ---------------------------------------------------------------------------
template <class T, int ITEMS, int VE_ = 16 / int(sizeof(T))>
struct Striped {
  static constexpr int VE = VE_;  // elements per lane per row
  static constexpr int NV = ITEMS / VE;
  static constexpr int N = kWave * ITEMS;
  static_assert(ITEMS % VE == 0, "ITEMS must fill whole rows");
  static_assert(VE * sizeof(T) == 4 || VE * sizeof(T) == 8 || VE * sizeof(T) == 16,
                "a lane's row slice must be 4, 8 or 16 bytes");
  // chunk offset of x[i] for lane l
  __device__ __forceinline__ static int offset(int i, int lane) {
    return ((i / VE) * kWave + lane) * VE + (i % VE);
  }
};

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
template <int BYTES>
struct RawVec;
template <>
struct RawVec<4> { using type = uint32_t; };
template <>
struct RawVec<8> { using type = u32x2; };
template <>
struct RawVec<16> { using type = u32x4; };

// NT = nontemporal loads (streams read exactly once).
template <class T, int ITEMS, int VE = 16 / int(sizeof(T)), bool NT = false>
__device__ __forceinline__ void striped_load(const T* chunk, int count, T (&x)[ITEMS], T fill) {
  using S = Striped<T, ITEMS, VE>;
  using RV = typename RawVec<VE * sizeof(T)>::type;
  const int lane = lane_id();
  if (count >= S::N && (reinterpret_cast<uintptr_t>(chunk) % (VE * sizeof(T))) == 0) {
    const RV* q = reinterpret_cast<const RV*>(chunk);
    RV t[S::NV];
#pragma unroll
    for (int j = 0; j < S::NV; ++j) {
      if constexpr (NT) t[j] = __builtin_nontemporal_load(q + j * kWave + lane);
      else t[j] = q[j * kWave + lane];
    }
#pragma unroll
    for (int j = 0; j < S::NV; ++j) {
      const T* e = reinterpret_cast<const T*>(&t[j]);
#pragma unroll
      for (int k = 0; k < S::VE; ++k) x[j * S::VE + k] = e[k];
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int o = S::offset(i, lane);
      x[i] = o < count ? chunk[o] : fill;
    }
  }
}

// NT = nontemporal stores (outputs not read again by this kernel).
template <class T, int ITEMS, int VE = 16 / int(sizeof(T)), bool NT = false>
__device__ __forceinline__ void striped_store(T* chunk, int count, const T (&x)[ITEMS]) {
  using S = Striped<T, ITEMS, VE>;
  using RV = typename RawVec<VE * sizeof(T)>::type;
  const int lane = lane_id();
  if (count >= S::N && (reinterpret_cast<uintptr_t>(chunk) % (VE * sizeof(T))) == 0) {
    RV* q = reinterpret_cast<RV*>(chunk);
#pragma unroll
    for (int j = 0; j < S::NV; ++j) {
      RV t;
      T* e = reinterpret_cast<T*>(&t);
#pragma unroll
      for (int k = 0; k < S::VE; ++k) e[k] = x[j * S::VE + k];
      if constexpr (NT) __builtin_nontemporal_store(t, q + j * kWave + lane);
      else q[j * kWave + lane] = t;
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int o = S::offset(i, lane);
      if (o < count) chunk[o] = x[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Wavefront-striped global access with an LDS transpose.
//
// A wavefront owns a contiguous chunk of 64*ITEMS elements; lane l works on
// the ITEMS consecutive elements [l*ITEMS, (l+1)*ITEMS) of it ("blocked"
// registers, what sequential per-thread scans and merges need).  Global
// memory is touched "striped": instruction j of lane l moves the 16-byte
// vector j*64 + l, so every instruction covers 1 KB of consecutive bytes.
// The per-wave LDS region holds one pad element after every 128 bytes, which
// makes the blocked reads and writes bank-conflict free for 4-byte types.
// Only the owning wavefront touches its region, so no workgroup barrier is
// needed.
// ---------------------------------------------------------------------------
template <class T, int ITEMS>
struct WaveTile {
  static constexpr int N = kWave * ITEMS;          // elements per wave chunk
  static constexpr int VE = 16 / int(sizeof(T));   // elements per 16-byte vector
  static constexpr int NV = ITEMS / VE;            // vectors per lane
  static constexpr int PAD = 128 / int(sizeof(T));  // one pad element per 128 bytes
  static constexpr int LDS_ELEMS = N + N / PAD;
  static_assert(ITEMS % VE == 0, "ITEMS must fill whole 16-byte vectors");
  __device__ __forceinline__ static int pos(int e) { return e + e / PAD; }
};

// Load chunk[0, count) (count <= N) into blocked registers; missing elements
// get `fill`.
template <class T, int ITEMS>
__device__ __forceinline__ void wave_load_blocked(const T* chunk, int count, T (&v)[ITEMS], T fill,
                                                  T* lds) {
  using W = WaveTile<T, ITEMS>;
  const int lane = lane_id();
  if (count >= W::N && aligned16(chunk)) {
    const u32x4* q = reinterpret_cast<const u32x4*>(chunk);
    u32x4 t[W::NV];
#pragma unroll
    for (int j = 0; j < W::NV; ++j) t[j] = q[j * kWave + lane];
#pragma unroll
    for (int j = 0; j < W::NV; ++j) {
      const T* e = reinterpret_cast<const T*>(&t[j]);
      const int base = W::pos((j * kWave + lane) * W::VE);
#pragma unroll
      for (int k = 0; k < W::VE; ++k) lds[base + k] = e[k];
    }
  } else {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const int e = j * kWave + lane;
      lds[W::pos(e)] = e < count ? chunk[e] : fill;
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) v[i] = lds[W::pos(lane * ITEMS + i)];
  wave_lds_sync();
}

// Store blocked registers to chunk[0, count) (count <= N).
template <class T, int ITEMS>
__device__ __forceinline__ void wave_store_blocked(T* chunk, int count, const T (&v)[ITEMS],
                                                   T* lds) {
  using W = WaveTile<T, ITEMS>;
  const int lane = lane_id();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) lds[W::pos(lane * ITEMS + i)] = v[i];
  wave_lds_sync();
  if (count >= W::N && aligned16(chunk)) {
    u32x4* q = reinterpret_cast<u32x4*>(chunk);
#pragma unroll
    for (int j = 0; j < W::NV; ++j) {
      u32x4 t;
      T* e = reinterpret_cast<T*>(&t);
      const int base = W::pos((j * kWave + lane) * W::VE);
#pragma unroll
      for (int k = 0; k < W::VE; ++k) e[k] = lds[base + k];
      q[j * kWave + lane] = t;
    }
  } else {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const int e = j * kWave + lane;
      if (e < count) chunk[e] = lds[W::pos(e)];
    }
  }
  wave_lds_sync();
}

}  // namespace bulkx
