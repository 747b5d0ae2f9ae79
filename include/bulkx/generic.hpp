// generic.hpp -- the device-wide drivers for operators, comparators,
// predicates and element types that cannot cross the C-ABI's enums (user
// functors, user structs): header-only, compiled in the caller's translation
// unit on top of the group algorithms of bulk.hpp, as the reference's
// drivers are built on its group algorithms (scan.cu:87-148, reduce.cu:58-97,
// merge.cu:127-162, merge_sort_by_key.cu:150-209, reduce_by_key.cu:176-272,
// reduce_intervals.hpp:31-47).  The C-ABI's hand-tuned kernels serve the
// enum cases; these serve the rest, with the same argument meaning.
//
// Operators need only be associative: every fold runs in input order.
// Work is split into P partitions of whole tiles (at most 4 per CU); a
// partition is one workgroup of 256 agents.
//   reduce:          fold each partition, then fold the partials with init.
//   scans:           fold each partition, exclusive-scan the partials from
//                    init in one group, then scan each partition from its carry.
//   merge:           one group per output tile; its merge-path split found by
//                    binary search, then the group merge of the two slices.
//   stable sort:     tiles of 1024 sorted in LDS by the bounded-group sort,
//                    then pairwise merge passes (ping-pong buffers).
//   reduce_by_key:   a segmented fold in three passes over (value, head
//                    flag, head count) triples: partition aggregates, their
//                    exclusive scan (which also gives the run count), then a
//                    scan of each partition that writes every run's last key
//                    and its value at the run's rank.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <iterator>
#include <type_traits>

#include "bulk.hpp"

namespace bulkx {
namespace generic {

constexpr int kGroup = 256;
constexpr int kGrain = 8;
constexpr long long kTile = (long long)kGroup * kGrain;  // elements per group and pass
constexpr int kSortGrain = 4;
constexpr long long kSortTile = (long long)kGroup * kSortGrain;  // block-sorted run
static_assert((2 * kSortTile) % kTile == 0, "a merge tile never spans two pairs of runs");

inline void hip_check(hipError_t e, const char* where) {
  if (e != hipSuccess) throw bulk::system_error(e, where);
}

// Stream-ordered device scratch of n T.
template <class T>
struct scratch {
  T* p = nullptr;
  hipStream_t s;
  scratch(long long n, hipStream_t stream) : s(stream) {
    hip_check(hipMallocAsync(reinterpret_cast<void**>(&p), sizeof(T) * size_t(n > 0 ? n : 1), s),
              "hipMallocAsync in a bulkx generic driver");
  }
  ~scratch() { (void)hipFreeAsync(p, s); }
  scratch(const scratch&) = delete;
  scratch& operator=(const scratch&) = delete;
};

// Partition p of n elements split into P partitions of whole tiles (the
// aligned decomposition, decomposition.hpp:142-218); never empty when
// P <= ceil(n / kTile).
__host__ __device__ inline void partition(long long n, long long P, long long p, long long* b,
                                          long long* e) {
  const long long tiles = (n + kTile - 1) / kTile, per = tiles / P, extra = tiles % P;
  const long long t0 = per * p + (p < extra ? p : extra);
  const long long t1 = t0 + per + (p < extra ? 1 : 0);
  *b = t0 * kTile;
  *e = t1 * kTile < n ? t1 * kTile : n;
}

inline long long partitions(long long n) {
  int dev = 0, cus = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
            "hipDeviceGetAttribute");
  const long long tiles = (n + kTile - 1) / kTile, want = 4LL * cus;
  return tiles < want ? tiles : want;
}

// n groups of kGroup agents with grain kGrain on stream s, no LDS heap.
inline auto groups(long long n, hipStream_t s) {
  return bulk::grid<kGroup, kGrain>(std::size_t(n), std::size_t(0), s);
}

// ------------------------------------------------------------ folds and scans
struct fold_partitions {
  template <class G, class It, class T, class Op>
  __device__ void operator()(G& g, It first, long long n, long long P, T* partials, Op op) {
    long long b, e;
    partition(n, P, g.index(), &b, &e);
    const T r = bulk::accumulate(g, first + (b + 1), first + e, T(first[b]), op);
    if (g.this_exec.index() == 0) partials[g.index()] = r;
  }
};

struct fold_partials {
  template <class G, class T, class Op>
  __device__ void operator()(G& g, const T* partials, long long P, T init, T* out, Op op) {
    const T r = bulk::accumulate(g, partials, partials + P, init, op);
    if (g.this_exec.index() == 0) *out = r;
  }
};

struct scan_partials {
  template <class G, class T, class Op>
  __device__ void operator()(G& g, const T* partials, long long P, T init, T* carries, Op op) {
    bulk::exclusive_scan(g, partials, partials + P, carries, init, op);
  }
};

template <bool INCLUSIVE>
struct scan_partitions {
  template <class G, class It1, class It2, class T, class Op>
  __device__ void operator()(G& g, It1 first, long long n, long long P, It2 result,
                             const T* carries, Op op) {
    long long b, e;
    partition(n, P, g.index(), &b, &e);
    if constexpr (INCLUSIVE)
      bulk::inclusive_scan(g, first + b, first + e, result + b, carries[g.index()], op);
    else
      bulk::exclusive_scan(g, first + b, first + e, result + b, carries[g.index()], op);
  }
};

// init op x[0] op ... op x[n-1] into *out (device), stream-ordered.
template <class It, class T, class Op>
void reduce_into(It first, long long n, T init, Op op, T* out, hipStream_t s) {
  if (n <= 0) {
    hip_check(hipMemcpyAsync(out, &init, sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");  // `init` is on this stack
    return;
  }
  const long long P = partitions(n);
  scratch<T> partials(P, s);
  bulk::async(groups(P, s), fold_partitions(), bulk::root.this_exec, first, n, P, partials.p, op);
  bulk::async(groups(1, s), fold_partials(), bulk::root.this_exec, (const T*)partials.p, P, init,
              out, op);
}

template <bool INCLUSIVE, class It1, class It2, class T, class Op>
void scan(It1 first, long long n, It2 result, T init, Op op, hipStream_t s) {
  if (n <= 0) return;
  const long long P = partitions(n);
  scratch<T> partials(P, s), carries(P, s);
  bulk::async(groups(P, s), fold_partitions(), bulk::root.this_exec, first, n, P, partials.p, op);
  bulk::async(groups(1, s), scan_partials(), bulk::root.this_exec, (const T*)partials.p, P, init,
              carries.p, op);
  bulk::async(groups(P, s), scan_partitions<INCLUSIVE>(), bulk::root.this_exec, first, n, P, result,
              (const T*)carries.p, op);
}

// -------------------------------------------------------------------- merge
struct merge_tiles {
  template <class G, class It1, class It2, class It3, class Compare>
  __device__ void operator()(G& g, It1 a, long long n1, It2 b, long long n2, It3 out,
                             Compare comp) {
    const long long n = n1 + n2, d0 = (long long)g.index() * kTile;
    const long long d1 = d0 + kTile < n ? d0 + kTile : n;
    const long long i0 = bulk::merge_path(a, n1, b, n2, d0, comp);
    const long long i1 = bulk::merge_path(a, n1, b, n2, d1, comp);
    bulk::merge(g, a + i0, a + i1, b + (d0 - i0), b + (d1 - i1), out + d0, comp);
  }
};

struct merge_by_key_tiles {
  template <class G, class K1, class K2, class V1, class V2, class K3, class V3, class Compare>
  __device__ void operator()(G& g, K1 a, long long n1, K2 b, long long n2, V1 va, V2 vb, K3 kout,
                             V3 vout, Compare comp) {
    const long long n = n1 + n2, d0 = (long long)g.index() * kTile;
    const long long d1 = d0 + kTile < n ? d0 + kTile : n;
    const long long i0 = bulk::merge_path(a, n1, b, n2, d0, comp);
    const long long i1 = bulk::merge_path(a, n1, b, n2, d1, comp);
    bulk::merge_by_key(g, a + i0, a + i1, b + (d0 - i0), b + (d1 - i1), va + i0, vb + (d0 - i0),
                       kout + d0, vout + d0, comp);
  }
};

template <class It1, class It2, class It3, class Compare>
void merge(It1 a, long long n1, It2 b, long long n2, It3 out, Compare comp, hipStream_t s) {
  const long long n = n1 + n2;
  if (n <= 0) return;
  bulk::async(groups((n + kTile - 1) / kTile, s), merge_tiles(), bulk::root.this_exec, a, n1, b,
              n2, out, comp);
}

template <class K1, class V1, class K2, class V2, class K3, class V3, class Compare>
void merge_by_key(K1 a, long long n1, V1 va, K2 b, long long n2, V2 vb, K3 kout, V3 vout,
                  Compare comp, hipStream_t s) {
  const long long n = n1 + n2;
  if (n <= 0) return;
  bulk::async(groups((n + kTile - 1) / kTile, s), merge_by_key_tiles(), bulk::root.this_exec, a,
              n1, b, n2, va, vb, kout, vout, comp);
}

// --------------------------------------------------------------- stable sort
struct sort_tiles {
  template <class G, class K, class V, class Compare>
  __device__ void operator()(G& g, K* keys, V* vals, long long n, Compare comp) {
    const long long b = (long long)g.index() * kSortTile;
    const long long cnt = n - b < kSortTile ? n - b : kSortTile;
    K* sk = static_cast<K*>(bulk::malloc(g, sizeof(K) * kSortTile));
    V* sv = static_cast<V*>(bulk::malloc(g, sizeof(V) * kSortTile));
    bulk::copy_n(g, keys + b, cnt, sk);
    bulk::copy_n(g, vals + b, cnt, sv);
    bulk::stable_sort_by_key(bulk::bound<kSortTile>(g), sk, sk + cnt, sv, comp);
    bulk::copy_n(g, sk, cnt, keys + b);
    bulk::copy_n(g, sv, cnt, vals + b);
    bulk::free(g, sv);
    bulk::free(g, sk);
  }
};

struct merge_pass {
  template <class G, class K, class V, class Compare>
  __device__ void operator()(G& g, const K* sk, const V* sv, long long n, long long w, K* dk, V* dv,
                             Compare comp) {
    const long long out0 = (long long)g.index() * kTile;
    const long long ps = (out0 / (2 * w)) * (2 * w);
    long long na = n - ps, nb = n - ps - w;
    na = na < 0 ? 0 : (na > w ? w : na);
    nb = nb < 0 ? 0 : (nb > w ? w : nb);
    const long long d0 = out0 - ps, d1 = d0 + kTile < na + nb ? d0 + kTile : na + nb;
    const K* a = sk + ps;
    const K* b = sk + ps + w;
    const long long i0 = bulk::merge_path(a, na, b, nb, d0, comp);
    const long long i1 = bulk::merge_path(a, na, b, nb, d1, comp);
    bulk::merge_by_key(g, a + i0, a + i1, b + (d0 - i0), b + (d1 - i1), sv + ps + i0,
                       sv + ps + w + (d0 - i0), dk + ps + d0, dv + ps + d0, comp);
  }
};

template <class K, class V, class Compare>
void stable_sort_by_key(K* keys, V* vals, long long n, Compare comp, hipStream_t s) {
  if (n <= 1) return;
  const long long tiles = (n + kSortTile - 1) / kSortTile;
  const std::size_t heap = (sizeof(K) + sizeof(V)) * std::size_t(kSortTile) + 256;
  bulk::async(bulk::grid<kGroup, kSortGrain>(std::size_t(tiles), heap, s), sort_tiles(),
              bulk::root.this_exec, keys, vals, n, comp);
  if (tiles == 1) return;
  scratch<K> tk(n, s);
  scratch<V> tv(n, s);
  K* sk = keys;
  V* sv = vals;
  K* dk = tk.p;
  V* dv = tv.p;
  for (long long w = kSortTile; w < n; w *= 2) {
    bulk::async(groups((n + kTile - 1) / kTile, s), merge_pass(), bulk::root.this_exec,
                (const K*)sk, (const V*)sv, n, w, dk, dv, comp);
    std::swap(sk, dk);
    std::swap(sv, dv);
  }
  if (sk != keys) {  // an odd number of passes ended in the scratch
    hip_check(hipMemcpyAsync(keys, sk, sizeof(K) * size_t(n), hipMemcpyDeviceToDevice, s),
              "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(vals, sv, sizeof(V) * size_t(n), hipMemcpyDeviceToDevice, s),
              "hipMemcpyAsync");
  }
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");  // scratch freed with the stream idle
}

// ------------------------------------------------------------- reduce_by_key
// A segmented partial: the fold of the values since the last run head
// (or of all of them when `head` is 0), whether a head was seen, and how
// many; `empty` marks the identity.  The combination is the reference's
// scan_head_flags operator (reduce_by_key.hpp:41-62) with a head count.
template <class V>
struct seg {
  V v;
  long long heads;
  int head;
  int empty;
};

template <class V, class Op>
struct seg_op {
  Op op;
  __host__ __device__ seg<V> operator()(const seg<V>& a, const seg<V>& b) const {
    if (a.empty) return b;
    if (b.empty) return a;
    seg<V> r;
    r.heads = a.heads + b.heads;
    r.head = a.head | b.head;
    r.empty = 0;
    r.v = b.head ? b.v : op(a.v, b.v);
    return r;
  }
};

// seg_iterator[i]: the segmented partial of element i alone.
template <class KIt, class VIt, class Pred>
struct seg_iterator {
  typedef typename std::iterator_traits<VIt>::value_type V;
  typedef seg<V> value_type;
  typedef long long difference_type;
  KIt keys;
  VIt vals;
  long long i;
  Pred pred;
  __host__ __device__ seg<V> operator[](long long j) const {
    const long long k = i + j;
    const int h = k == 0 || !pred(keys[k - 1], keys[k]);
    return seg<V>{V(vals[k]), (long long)h, h, 0};
  }
  __host__ __device__ seg<V> operator*() const { return (*this)[0]; }
  __host__ __device__ seg_iterator operator+(long long d) const {
    return seg_iterator{keys, vals, i + d, pred};
  }
  __host__ __device__ long long operator-(const seg_iterator& o) const { return i - o.i; }
  __host__ __device__ bool operator<(const seg_iterator& o) const { return i < o.i; }
};

// tail_writer[i] = s: when element i ends a run, its key and the run's value
// go to slot s.heads - 1 (the run's rank); otherwise nothing is written.
template <class KIt, class KOut, class VOut, class Pred>
struct tail_writer {
  KIt keys;
  long long n;
  KOut kout;
  VOut vout;
  long long i;
  Pred pred;
  struct proxy {
    const tail_writer* w;
    long long k;
    template <class S>
    __device__ proxy& operator=(const S& s) {
      if (k == w->n - 1 || !w->pred(w->keys[k], w->keys[k + 1])) {
        w->kout[s.heads - 1] = w->keys[k];
        w->vout[s.heads - 1] = s.v;
      }
      return *this;
    }
  };
  __device__ proxy operator[](long long j) const { return proxy{this, i + j}; }
  __host__ __device__ tail_writer operator+(long long d) const {
    return tail_writer{keys, n, kout, vout, i + d, pred};
  }
};

struct rbk_finish {
  template <class G, class S, class Op>
  __device__ void operator()(G& g, const S* aggs, long long P, S* carries, long long* count, Op op) {
    S empty;
    empty.empty = 1;
    empty.heads = 0;
    empty.head = 0;
    bulk::exclusive_scan(g, aggs, aggs + P, carries, empty, op);
    if (g.this_exec.index() == 0) *count = op(carries[P - 1], aggs[P - 1]).heads;
  }
};

struct rbk_write {
  template <class G, class SIt, class W, class S, class Op>
  __device__ void operator()(G& g, SIt segs, long long n, long long P, W writer, const S* carries,
                             Op op) {
    long long b, e;
    partition(n, P, g.index(), &b, &e);
    bulk::inclusive_scan(g, segs + b, segs + e, writer + b, carries[g.index()], op);
  }
};

// One output per maximal run of pred-equal keys: the run's last key and the
// op-fold of its values; the number of runs is written to *count (device).
template <class KIt, class VIt, class KOut, class VOut, class Pred, class Op>
void reduce_by_key(KIt keys, VIt vals, long long n, KOut kout, VOut vout, Pred pred, Op op,
                   long long* count, hipStream_t s) {
  typedef typename std::iterator_traits<VIt>::value_type V;
  typedef seg<V> S;
  if (n <= 0) {
    hip_check(hipMemsetAsync(count, 0, sizeof(long long), s), "hipMemsetAsync");
    return;
  }
  const long long P = partitions(n);
  const seg_op<V, Op> sop{op};
  const seg_iterator<KIt, VIt, Pred> segs{keys, vals, 0, pred};
  scratch<S> aggs(P, s), carries(P, s);
  bulk::async(groups(P, s), fold_partitions(), bulk::root.this_exec, segs, n, P, aggs.p, sop);
  bulk::async(groups(1, s), rbk_finish(), bulk::root.this_exec, (const S*)aggs.p, P, carries.p,
              count, sop);
  const tail_writer<KIt, KOut, VOut, Pred> writer{keys, n, kout, vout, 0, pred};
  bulk::async(groups(P, s), rbk_write(), bulk::root.this_exec, segs, n, P, writer,
              (const S*)carries.p, sop);
}

// ----------------------------------------------------------- reduce_intervals
// result[i] = the op-fold of partition i of any decomposition (operator[]
// giving a [first, second) range, size()); empty partitions leave their
// slot untouched.
struct intervals {
  template <class G, class It1, class D, class It2, class Op>
  __device__ void operator()(G& g, It1 first, D decomp, It2 result, Op op) {
    const auto r = decomp[g.index()];
    typedef typename std::iterator_traits<It2>::value_type T;
    if (r.second <= r.first) return;
    const T v = bulk::accumulate(g, first + (r.first + 1), first + r.second, T(first[r.first]), op);
    if (g.this_exec.index() == 0) result[g.index()] = v;
  }
};

template <class It1, class D, class It2, class Op>
void reduce_intervals(It1 first, const D& decomp, It2 result, Op op, hipStream_t s) {
  if (decomp.size() <= 0) return;
  bulk::async(groups((long long)decomp.size(), s), intervals(), bulk::root.this_exec, first, decomp,
              result, op);
}

}  // namespace generic
}  // namespace bulkx
