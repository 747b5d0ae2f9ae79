// CPU oracle for the bulk algorithm layer.  TEST INFRASTRUCTURE ONLY: the
// tests, the smoke check and the CPU-baseline leg of the benchmark call it
// as the checker / the timed CPU restatement; the product never links it.
//
// It restates, sequentially and in plain C++, the reference's agent bodies
// and device-wide decompositions (each function cites the file:line it
// follows), and adds the exact / extended-precision references the float
// tolerances are measured against, plus the counter-based input generators
// shared by tests and fixtures.
//
// Pinning (how each restatement is checked, see tests/test_oracle.py):
//   scan / reduce (integer): closed forms of the reference's own checks
//     (ones + 13 -> 14 + i, scan.cu:158-184; sequence + 13 -> 1206807391,
//     reduce.cu:132-148), numpy cumsum/sum in wrapping arithmetic.
//   scan / reduce (float): long-double reference; numpy float64 cumsum on
//     small inputs; brute force.
//   merge / merge_by_key: numpy stable argsort of the concatenation
//     (== std::merge, first range wins ties).
//   stable_sort_by_key: numpy stable argsort; payload = index checks stability.
//   reduce_by_key (reference interval algorithm): the textbook run loop.
//   reduce_intervals / aligned_decomposition: closed form + brute force.
//   generators: golden fixtures committed under tests/golden.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Operators (thrust::plus / minimum / maximum; less / greater; equal_to).
// Integer addition wraps in the unsigned type.
// ---------------------------------------------------------------------------
template <class T>
T op_apply(int op, T a, T b) {
  switch (op) {
    case 0:
      if constexpr (std::is_integral<T>::value) {
        using U = typename std::make_unsigned<T>::type;
        return T(U(a) + U(b));
      } else {
        return a + b;
      }
    case 1: return a < b ? a : b;
    default: return a < b ? b : a;
  }
}

template <class T>
bool cmp_apply(int comp, T a, T b) { return comp == 0 ? (a < b) : (b < a); }

// ---------------------------------------------------------------------------
// aligned_decomposition(n, P, tile), decomposition.hpp:142-218.
// ---------------------------------------------------------------------------
struct Aligned {
  int64_t n, P, tile, q, r;
  Aligned(int64_t n_, int64_t P_, int64_t tile_) : n(n_), P(P_), tile(tile_) {
    int64_t tiles = (n + tile - 1) / tile;
    q = tiles / P;
    r = tiles % P;
  }
  void range(int64_t i, int64_t* first, int64_t* last) const {
    int64_t f = q * i + std::min(i, r);  // decomposition.hpp:197-198
    int64_t l = f + q + (i < r ? 1 : 0);  // decomposition.hpp:200
    *first = f * tile;
    *last = std::min(n, l * tile);  // decomposition.hpp:173-174
  }
};

// ---------------------------------------------------------------------------
// Sequential agent bodies.
// ---------------------------------------------------------------------------

// bulk::inclusive_scan on a bounded agent, scan.hpp:35-55:
//   init = op(init, first[i]); result[i] = init.
template <class T>
void seq_inclusive_scan(const T* x, int64_t n, T init, bool has_init, int op, T* out) {
  if (n == 0) return;
  int64_t i = 0;
  T acc;
  if (has_init) {
    acc = init;
  } else {  // no-init overload: the first element is the init (scan.hpp:489-526)
    acc = x[0];
    out[0] = acc;
    i = 1;
  }
  for (; i < n; ++i) {
    acc = op_apply(op, acc, x[i]);
    out[i] = acc;
  }
}

// bulk::exclusive_scan on a bounded agent, scan.hpp:58-78.
template <class T>
void seq_exclusive_scan(const T* x, int64_t n, T init, int op, T* out) {
  T acc = init;
  for (int64_t i = 0; i < n; ++i) {
    T xi = x[i];
    out[i] = acc;
    acc = op_apply(op, acc, xi);
  }
}

// bulk::reduce / bulk::accumulate on a bounded agent, reduce.hpp:33-58,
// accumulate.hpp:30-55: left fold from init.
template <class T>
T seq_reduce(const T* x, int64_t n, T init, int op) {
  T acc = init;
  for (int64_t i = 0; i < n; ++i) acc = op_apply(op, acc, x[i]);
  return acc;
}

// bulk::merge_path, merge.hpp:35-60: number of elements taken from range 1
// on diagonal `diag`.
template <class T>
int64_t merge_path(const T* a, int64_t n1, const T* b, int64_t n2, int64_t diag, int comp) {
  int64_t begin = std::max<int64_t>(0, diag - n2);
  int64_t end = std::min(diag, n1);
  while (begin < end) {
    int64_t mid = (begin + end) >> 1;
    if (cmp_apply(comp, b[diag - 1 - mid], a[mid])) end = mid;
    else begin = mid + 1;
  }
  return begin;
}

// Sequential merge / merge_by_key agent body, merge.hpp:63-170 / 173-305:
// take from range 1 unless range 2's head compares strictly less.
template <class K, class V>
void seq_merge_by_key(const K* k1, const V* v1, int64_t n1, const K* k2, const V* v2, int64_t n2,
                      K* ko, V* vo, int comp) {
  int64_t i1 = 0, i2 = 0, o = 0;
  while (i1 < n1 || i2 < n2) {
    bool p = (i2 >= n2) || ((i1 < n1) && !cmp_apply(comp, k2[i2], k1[i1]));  // merge.hpp:111
    if (p) {
      ko[o] = k1[i1];
      if (vo) vo[o] = v1[i1];
      ++i1;
    } else {
      ko[o] = k2[i2];
      if (vo) vo[o] = v2[i2];
      ++i2;
    }
    ++o;
  }
}

// Odd-even transposition sort by key of one agent's grain, sort.hpp:33-75:
// `bound` rounds of compare-swaps (j, j+1), strict comparator -> stable.
template <class K, class V>
void odd_even_sort_by_key(K* k, V* v, int n, int bound, int comp) {
  for (int i = 0; i < bound; ++i) {
    for (int j = 1 & i; j < bound - 1; j += 2) {
      if (j + 1 < n && cmp_apply(comp, k[j + 1], k[j])) {
        std::swap(k[j], k[j + 1]);
        if (v) std::swap(v[j], v[j + 1]);
      }
    }
  }
}

// stable_merge_sort_by_key, merge_sort_by_key.cu:150-209 with the block sort
// of stable_merge_sort.hpp:125-206: tiles of groupsize*grainsize (128*7),
// each agent's grain sorted by odd-even transposition, runs merged pairwise
// with doubling width inside the tile and then across tiles
// (log2_ri(num_groups) passes, ping-pong buffers).
template <class K, class V>
void ref_stable_sort_by_key(K* keys, V* vals, int64_t n, int comp) {
  const int grain = 7, group = 128;
  const int64_t tile = int64_t(grain) * group;
  for (int64_t g = 0; g * grain < n; ++g) {
    int64_t f = g * grain;
    int cnt = int(std::min<int64_t>(grain, n - f));
    odd_even_sort_by_key(keys + f, vals ? vals + f : nullptr, cnt, grain, comp);
  }
  std::vector<K> kt(n);
  std::vector<V> vt(vals ? n : 0);
  K* ks = keys;
  V* vs = vals;
  K* kd = kt.data();
  V* vd = vals ? vt.data() : nullptr;
  bool in_original = true;
  // tile = grain * 2^7, so doubling from the grain first covers the 7 in-tile
  // rounds (stable_merge_sort.hpp:75-118) and then the cross-tile passes
  // (merge_sort_by_key.cu:186-202).
  for (int64_t width = grain; width < n; width *= 2) {
    for (int64_t f = 0; f < n; f += 2 * width) {
      int64_t m = std::min(n, f + width), l = std::min(n, f + 2 * width);
      seq_merge_by_key(ks + f, vs ? vs + f : nullptr, m - f, ks + m, vs ? vs + m : nullptr, l - m,
                       kd + f, vd ? vd + f : nullptr, comp);
    }
    std::swap(ks, kd);
    std::swap(vs, vd);
    in_original = !in_original;
  }
  (void)tile;
  if (!in_original) {
    std::copy(ks, ks + n, keys);
    if (vals) std::copy(vs, vs + n, vals);
  }
}

// Textbook reduce_by_key (the pinning check of the interval algorithm).
template <class K, class V>
int64_t textbook_reduce_by_key(const K* k, const V* v, int64_t n, K* ko, V* vo, int op) {
  if (n == 0) return 0;
  int64_t m = 0;
  V acc = v[0];
  for (int64_t i = 1; i < n; ++i) {
    if (k[i] == k[i - 1]) {
      acc = op_apply(op, acc, v[i]);
    } else {
      ko[m] = k[i - 1];
      vo[m] = acc;
      ++m;
      acc = v[i];
    }
  }
  ko[m] = k[n - 1];
  vo[m] = acc;
  return m + 1;
}

// bulk::reduce_by_key for one interval, reduce_by_key.hpp:115-216, restated
// sequentially: flags against (init_key, init_value), segmented scan,
// tails written to slot flag-1, carry returned through init_*.
template <class K, class V>
int64_t interval_reduce_by_key(const K* k, const V* v, int64_t n, K* ko, V* vo, K* init_key,
                               V* init_value, int op) {
  int64_t m = 0;
  K ik = *init_key;
  V iv = *init_value;
  for (int64_t i = 0; i < n; ++i) {
    if (k[i] == ik) {
      iv = op_apply(op, iv, v[i]);
    } else {  // head: the running segment ends, its (last key, sum) is written
      ko[m] = ik;
      vo[m] = iv;
      ++m;
      iv = v[i];
    }
    ik = k[i];
  }
  *init_key = ik;
  *init_value = iv;
  return m;
}

// my_reduce_by_key large path, reduce_by_key.cu:215-272: intervals of an
// aligned_decomposition (tile 640, P = min(100*SMs, ceil(n/20000))), tail
// counts per interval (reduce_intervals, reduce_by_key.cu:240), their
// inclusive scan as output offsets (:243), per-interval reduce_by_key from
// the interval's first element with carry detection (:31-94), then the
// carries summed into the slot they belong to (:255-267).  The small path
// (n <= 20000, :197-213) is the same computation in one interval.
template <class K, class V>
int64_t ref_reduce_by_key(const K* k, const V* v, int64_t n, K* ko, V* vo, int op, int64_t sms) {
  if (n == 0) return 0;
  const int64_t interval = 20000, tile = 640;
  int64_t P = n <= interval ? 1 : std::min<int64_t>(100 * sms, (n + interval - 1) / interval);
  Aligned d(n, P, n <= interval ? n : tile);
  std::vector<int64_t> offsets(P);
  for (int64_t g = 0; g < P; ++g) {  // reduce_intervals over tail flags
    int64_t f, l;
    d.range(g, &f, &l);
    int64_t c = 0;
    for (int64_t i = f; i < l; ++i) c += (i == n - 1 || !(k[i] == k[i + 1])) ? 1 : 0;
    offsets[g] = c;
  }
  for (int64_t g = 1; g < P; ++g) offsets[g] += offsets[g - 1];
  std::vector<V> interval_values(P);
  std::vector<char> is_carry(P);
  for (int64_t g = 0; g < P; ++g) {
    int64_t f, l;
    d.range(g, &f, &l);
    if (f >= l) {  // empty interval (more partitions than tiles): no output, no carry
      is_carry[g] = 0;
      continue;
    }
    int64_t out = g == 0 ? 0 : offsets[g - 1];
    K ik = k[f];
    V iv = v[f];
    int64_t m = interval_reduce_by_key(k + f + 1, v + f + 1, l - f - 1, ko + out, vo + out, &ik,
                                       &iv, op);
    bool carry = !(l - 1 == n - 1 || !(k[l - 1] == k[l]));
    if (carry) {
      interval_values[g] = iv;
    } else {
      ko[out + m] = ik;
      vo[out + m] = iv;
    }
    is_carry[g] = carry;
  }
  // inclusive_scan_by_key over (offset, is_carry) then add each segment-tail
  // carry into values_result[offset] as op(carry, value) (reduce_by_key.cu:167-172).
  for (int64_t g = 0; g < P; ++g) {
    if (!is_carry[g]) continue;
    if (g > 0 && is_carry[g - 1] && offsets[g - 1] == offsets[g])
      interval_values[g] = op_apply(op, interval_values[g - 1], interval_values[g]);
    bool tail = (g == P - 1) || !(is_carry[g + 1] && offsets[g + 1] == offsets[g]);
    if (tail) vo[offsets[g]] = op_apply(op, interval_values[g], vo[offsets[g]]);
  }
  return offsets[P - 1];
}

// ---------------------------------------------------------------------------
// Counter-based generators: splitmix64 (Steele, Lea, Flood 2014) applied to
// seed-mixed counters, so element i depends only on (seed, stream, i).
// ---------------------------------------------------------------------------
inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t rng(uint64_t seed, uint64_t stream, uint64_t i) {
  return splitmix64(splitmix64(seed * 0x2545F4914F6CDD1Dull + stream) ^ (i * 0xD1B54A32D192ED03ull));
}
inline double u01_53(uint64_t r) { return double(r >> 11) * (1.0 / 9007199254740992.0); }
inline float u01_24(uint64_t r) { return float(r >> 40) * (1.0f / 16777216.0f); }

}  // namespace

// ===========================================================================
// extern "C" surface (ctypes)
// ===========================================================================
#define ORC_FOR_TYPES(X) X(i32, int32_t) X(u32, uint32_t) X(i64, int64_t) X(u64, uint64_t) X(f32, float) X(f64, double)

extern "C" {

#define ORC_SCAN(tag, T)                                                                       \
  void orc_inclusive_scan_##tag(const T* x, int64_t n, const T* init, int op, T* out) {       \
    seq_inclusive_scan<T>(x, n, init ? *init : T(0), init != nullptr, op, out);               \
  }                                                                                            \
  void orc_exclusive_scan_##tag(const T* x, int64_t n, T init, int op, T* out) {              \
    seq_exclusive_scan<T>(x, n, init, op, out);                                                \
  }                                                                                            \
  T orc_reduce_##tag(const T* x, int64_t n, T init, int op) { return seq_reduce<T>(x, n, init, op); } \
  void orc_reduce_intervals_##tag(const T* x, int64_t n, int64_t P, int64_t tile, int op, T* out) { \
    Aligned d(n, P, tile);                                                                     \
    for (int64_t g = 0; g < P; ++g) {                                                          \
      int64_t f, l;                                                                            \
      d.range(g, &f, &l);                                                                      \
      if (f >= l) { out[g] = T(0); continue; }                                                 \
      /* reduce_intervals.hpp:19-21: init = last element, fold the rest */                    \
      out[g] = seq_reduce<T>(x + f, l - 1 - f, x[l - 1], op);                                  \
    }                                                                                          \
  }
ORC_FOR_TYPES(ORC_SCAN)

#define ORC_MERGE(tag, K)                                                                          \
  void orc_merge_##tag(const K* a, int64_t n1, const K* b, int64_t n2, int comp, K* out) {        \
    seq_merge_by_key<K, int32_t>(a, nullptr, n1, b, nullptr, n2, out, nullptr, comp);              \
  }                                                                                                \
  int64_t orc_merge_path_##tag(const K* a, int64_t n1, const K* b, int64_t n2, int64_t diag,      \
                               int comp) {                                                         \
    return merge_path<K>(a, n1, b, n2, diag, comp);                                                \
  }
ORC_FOR_TYPES(ORC_MERGE)

#define ORC_KV(X) X(i32, int32_t, i32, int32_t) X(i32, int32_t, f32, float) X(i32, int32_t, i64, int64_t) \
  X(u32, uint32_t, u32, uint32_t) X(f32, float, i32, int32_t) X(f32, float, f32, float)                  \
  X(i64, int64_t, i64, int64_t) X(f64, double, f64, double) X(i32, int32_t, f64, double)                \
  X(u64, uint64_t, u64, uint64_t) X(u32, uint32_t, i32, int32_t) X(i64, int64_t, f32, float)

#define ORC_KV_FN(kt, K, vt, V)                                                                        \
  void orc_merge_by_key_##kt##_##vt(const K* k1, const V* v1, int64_t n1, const K* k2, const V* v2,   \
                                    int64_t n2, int comp, K* ko, V* vo) {                              \
    seq_merge_by_key<K, V>(k1, v1, n1, k2, v2, n2, ko, vo, comp);                                      \
  }                                                                                                    \
  void orc_stable_sort_by_key_##kt##_##vt(K* k, V* v, int64_t n, int comp) {                          \
    ref_stable_sort_by_key<K, V>(k, v, n, comp);                                                       \
  }                                                                                                    \
  int64_t orc_reduce_by_key_##kt##_##vt(const K* k, const V* v, int64_t n, K* ko, V* vo, int op,      \
                                        int64_t sms) {                                                 \
    return ref_reduce_by_key<K, V>(k, v, n, ko, vo, op, sms);                                          \
  }                                                                                                    \
  int64_t orc_textbook_reduce_by_key_##kt##_##vt(const K* k, const V* v, int64_t n, K* ko, V* vo,     \
                                                 int op) {                                             \
    return textbook_reduce_by_key<K, V>(k, v, n, ko, vo, op);                                          \
  }
ORC_KV(ORC_KV_FN)

// Sort keys only (bulk::stable_sort): payload-free form of the same algorithm.
#define ORC_SORT_KEYS(tag, K)                                              \
  void orc_stable_sort_##tag(K* k, int64_t n, int comp) {                  \
    ref_stable_sort_by_key<K, int32_t>(k, nullptr, n, comp);               \
  }
ORC_FOR_TYPES(ORC_SORT_KEYS)

// ---- aligned_decomposition ranges (decomposition.hpp:142-218) ----
void orc_aligned_ranges(int64_t n, int64_t P, int64_t tile, int64_t* first, int64_t* last) {
  Aligned d(n, P, tile);
  for (int64_t i = 0; i < P; ++i) d.range(i, first + i, last + i);
}

// ---- flags (tail_flags.hpp:52-61, bulk/detail/head_flags.hpp:63-75) ----
void orc_tail_flags_i32(const int32_t* k, int64_t n, uint8_t* f) {
  for (int64_t i = 0; i < n; ++i) f[i] = (i == n - 1 || !(k[i] == k[i + 1])) ? 1 : 0;
}
void orc_head_flags_i32(const int32_t* k, int64_t n, uint8_t* f) {
  for (int64_t i = 0; i < n; ++i) f[i] = (i == 0 || !(k[i] == k[i - 1])) ? 1 : 0;
}

// ---- extended-precision references for the float tolerances ----
// Prefix sums in long double (64-bit significand), returned as hi + lo
// doubles, together with the prefix sum of |x|.
// Prefix sums in long double with Neumaier's compensation (Kahan-Babuska:
// the rounding error of every addition is kept in a second accumulator,
// whichever operand is larger), so the prefix is exact to a few units of
// the long double format however long the array is (plain accumulation
// errs by up to n * 2^-64 * sum|x|, above the float64 bound at 2^28).
}  // extern "C"
template <class T>
static void prefix_sum_neumaier(const T* x, int64_t n, double init, double* hi, double* lo,
                                double* abs_prefix) {
  long double s = init, c = 0, a = fabsl((long double)init);
  for (int64_t i = 0; i < n; ++i) {
    const long double v = (long double)x[i];
    const long double t = s + v;
    if (fabsl(s) >= fabsl(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
    a += fabsl(v);
    const long double p = s + c;
    hi[i] = (double)p;
    lo[i] = (double)(p - (long double)hi[i]);
    abs_prefix[i] = (double)a;
  }
}
extern "C" {
void orc_prefix_sum_ld_f32(const float* x, int64_t n, double init, double* hi, double* lo,
                           double* abs_prefix) {
  prefix_sum_neumaier(x, n, init, hi, lo, abs_prefix);
}
void orc_prefix_sum_ld_f64(const double* x, int64_t n, double init, double* hi, double* lo,
                           double* abs_prefix) {
  prefix_sum_neumaier(x, n, init, hi, lo, abs_prefix);
}
// Sum and sum |x| in long double with Kahan compensation on top.
void orc_sum_ld_f64(const double* x, int64_t n, double* hi, double* lo, double* abs_sum) {
  long double s = 0, c = 0, a = 0;
  for (int64_t i = 0; i < n; ++i) {
    long double y = (long double)x[i] - c;
    long double t = s + y;
    c = (t - s) - y;
    s = t;
    a += fabsl((long double)x[i]);
  }
  *hi = (double)s;
  *lo = (double)(s - (long double)*hi);
  *abs_sum = (double)a;
}
void orc_sum_ld_f32(const float* x, int64_t n, double* hi, double* lo, double* abs_sum) {
  long double s = 0, a = 0;
  for (int64_t i = 0; i < n; ++i) {
    s += (long double)x[i];
    a += fabsl((long double)x[i]);
  }
  *hi = (double)s;
  *lo = (double)(s - (long double)*hi);
  *abs_sum = (double)a;
}

// ---- generators ----
void orc_gen_uniform_i32(uint64_t seed, int32_t lo, int32_t hi, int64_t n, int32_t* out) {
  uint64_t span = uint64_t(int64_t(hi) - int64_t(lo) + 1);
  for (int64_t i = 0; i < n; ++i) out[i] = int32_t(int64_t(lo) + int64_t(rng(seed, 1, i) % span));
}
void orc_gen_uniform_i64(uint64_t seed, int64_t lo, int64_t hi, int64_t n, int64_t* out) {
  uint64_t span = uint64_t(hi - lo) + 1;
  for (int64_t i = 0; i < n; ++i) out[i] = lo + int64_t(rng(seed, 2, i) % span);
}
void orc_gen_uniform_u32(uint64_t seed, int64_t n, uint32_t* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = uint32_t(rng(seed, 3, i) >> 32);
}
void orc_gen_uniform_f32(uint64_t seed, float lo, float hi, int64_t n, float* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * u01_24(rng(seed, 4, i));
}
void orc_gen_uniform_f64(uint64_t seed, double lo, double hi, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * u01_53(rng(seed, 5, i));
}
// Box-Muller in double: pair (2j, 2j+1) from two uniforms.
void orc_gen_normal_f64(uint64_t seed, int64_t n, double* out) {
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t j = 0; 2 * j < n; ++j) {
    double u1 = 1.0 - u01_53(rng(seed, 6, 2 * j));  // (0, 1]
    double u2 = u01_53(rng(seed, 6, 2 * j + 1));
    double r = sqrt(-2.0 * log(u1));
    out[2 * j] = r * cos(two_pi * u2);
    if (2 * j + 1 < n) out[2 * j + 1] = r * sin(two_pi * u2);
  }
}
void orc_gen_normal_f32(uint64_t seed, int64_t n, float* out) {
  std::vector<double> tmp(n);
  orc_gen_normal_f64(seed, n, tmp.data());
  for (int64_t i = 0; i < n; ++i) out[i] = float(tmp[i]);
}
// Contiguous runs, key = run index.  Geometric run lengths with mean `mean`
// (P(L = k) = p (1-p)^(k-1), p = 1/mean).  Returns the number of runs.
int64_t orc_gen_runs_geometric(uint64_t seed, double mean, int64_t n, int32_t* keys) {
  double p = 1.0 / mean;
  double lq = log1p(-p);
  int64_t i = 0, run = 0;
  while (i < n) {
    double u = 1.0 - u01_53(rng(seed, 7, uint64_t(run)));
    int64_t len = p >= 1.0 ? 1 : int64_t(floor(log(u) / lq)) + 1;
    if (len < 1) len = 1;
    for (int64_t j = 0; j < len && i < n; ++j) keys[i++] = int32_t(run);
    ++run;
  }
  return run;
}
// Zipf(s) run lengths on 1..cap by inverse CDF.  Returns the number of runs.
int64_t orc_gen_runs_zipf(uint64_t seed, double s, int64_t cap, int64_t n, int32_t* keys) {
  std::vector<double> cdf(cap);
  double acc = 0;
  for (int64_t k = 1; k <= cap; ++k) {
    acc += pow(double(k), -s);
    cdf[k - 1] = acc;
  }
  for (auto& c : cdf) c /= acc;
  int64_t i = 0, run = 0;
  while (i < n) {
    double u = u01_53(rng(seed, 8, uint64_t(run)));
    int64_t len = int64_t(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
    if (len > cap) len = cap;
    for (int64_t j = 0; j < len && i < n; ++j) keys[i++] = int32_t(run);
    ++run;
  }
  return run;
}
// Sorted float keys in [0,1) with a fraction of entries replaced by values of
// a shared pool (forcing ties across both merge inputs).
void orc_gen_merge_keys(uint64_t seed, uint64_t stream, int64_t n, const float* pool,
                        int64_t pool_n, double frac, float* out) {
  for (int64_t i = 0; i < n; ++i) {
    uint64_t r = rng(seed, 100 + stream, uint64_t(i));
    if (pool_n > 0 && u01_53(rng(seed, 200 + stream, uint64_t(i))) < frac)
      out[i] = pool[r % uint64_t(pool_n)];
    else
      out[i] = u01_24(r);
  }
  std::sort(out, out + n);
}
// The reference's hash(i) input fill (merge.cu:203-218, merge_sort_by_key.cu,
// reduce_by_key.cu:305-321), evaluated in int64 (thrust::tabulate passes the
// 64-bit signed difference_type) and truncated to int32.
void orc_ref_hash_i32(int64_t n, int mod10, int32_t* out) {
  for (int64_t i = 0; i < n; ++i) {
    int64_t x = i;
    x = (x + 0x7ed55d16) + (int64_t)((uint64_t)x << 12);
    x = (x ^ 0xc761c23c) ^ (x >> 19);
    x = (x + 0x165667b1) + (int64_t)((uint64_t)x << 5);
    x = (x + 0xd3a2646c) ^ (int64_t)((uint64_t)x << 9);
    x = (x + 0xfd7046c5) + (int64_t)((uint64_t)x << 3);
    x = (x ^ 0xb55a4f09) ^ (x >> 16);
    out[i] = mod10 ? int32_t(x % 10) : int32_t(x);
  }
}
// thrust::default_random_engine (minstd_rand, a = 48271, m = 2^31 - 1, seed 1).
void orc_minstd(int64_t count, uint32_t* out) {
  uint64_t s = 1;
  for (int64_t i = 0; i < count; ++i) {
    s = (s * 48271ull) % 2147483647ull;
    out[i] = uint32_t(s);
  }
}

// ---- CPU baseline: the agent bodies over a range, 1 or T threads ----
// scan: the reference's 3-phase decomposition (scan.cu:109-145) with SMs := T
// threads: upsweep partials per aligned partition, exclusive scan of the
// carries, downsweep scan seeded by the carry.
void orc_mt_inclusive_scan_i32(const int32_t* x, int64_t n, int32_t init, int32_t* out, int threads) {
  if (threads <= 1 || n < 20000) {
    seq_inclusive_scan<int32_t>(x, n, init, true, 0, out);
    return;
  }
  const int64_t tile = 128 * 9, P = std::min<int64_t>(20 * threads, (n + tile - 1) / tile);
  Aligned d(n, P, tile);
  std::vector<int32_t> carry(P);
  auto run = [&](auto&& body) {
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t)
      th.emplace_back([&, t] {
        for (int64_t g = t; g < P; g += threads) body(g);
      });
    for (auto& h : th) h.join();
  };
  run([&](int64_t g) {
    int64_t f, l;
    d.range(g, &f, &l);
    carry[g] = seq_reduce<int32_t>(x + f, l - f, 0, 0);
  });
  seq_exclusive_scan<int32_t>(carry.data(), P, init, 0, carry.data());
  run([&](int64_t g) {
    int64_t f, l;
    d.range(g, &f, &l);
    seq_inclusive_scan<int32_t>(x + f, l - f, carry[g], true, 0, out + f);
  });
}

double orc_now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

unsigned orc_hardware_concurrency() { return std::thread::hardware_concurrency(); }

}  // extern "C"
