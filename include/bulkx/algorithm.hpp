// algorithm.hpp -- the reference's device-wide drivers as C++ templates over
// the C-ABI of bulkx.h, with the reference's names and argument order:
//
//   inclusive_scan(first, last, result, init, op)          scan.cu:87-148
//   exclusive_scan(first, last, result, init, op)          scan.hpp:553-591
//   my_reduce(first, last, init, op)                       reduce.cu:58-97
//   my_merge(first1, last1, first2, last2, result, comp)   merge.cu:127-162
//   merge_by_key(k1f, k1l, v1f, k2f, k2l, v2f, kout, vout, comp)
//                                                           merge.hpp:507-607
//   stable_merge_sort_by_key(keys_first, keys_last, values_first, comp)
//                                                           merge_sort_by_key.cu:150-209
//   my_reduce_by_key(kf, kl, vf, kr, vr, pred, op)         reduce_by_key.cu:176-272
//   reduce_intervals(first, decomposition, result, op)     reduce_intervals.hpp:31-47
//
// Iterators are contiguous device ranges: raw pointers, or (when Thrust is
// included first) thrust::device_ptr / device_vector iterators.  Operators
// map onto the library's enums: plus / minimum / maximum for scan and
// reduce, less / greater for merge and sort, equal_to for reduce_by_key
// (std:: and thrust:: functors, or the ones below).  Every call is ordered
// on `bulkx::default_stream()` (the null stream unless set); my_reduce and
// my_reduce_by_key read their scalar back with a blocking copy, as the
// reference does (reduce.cu:96, reduce_by_key.cu:269); async_reduce and
// async_reduce_by_key return the scalar through a scalar_future instead.
// Errors are thrown as bulkx::system_error, the counterpart of throw_on_error
// (bulk/detail/throw_on_error.hpp:33-49).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "../bulkx.h"
#include "generic.hpp"

// Thrust's functors and device pointers are recognised when Thrust is in
// use (any Thrust header included before this one).
#ifdef THRUST_VERSION
#include <thrust/functional.h>
#include <thrust/memory.h>
#endif

namespace bulkx {

class system_error : public std::runtime_error {
 public:
  system_error(int status, const std::string& where)
      : std::runtime_error(where + ": " + bulkx_status_string(status) +
                           (status == BULKX_ERR_HIP ? std::string(" (") + bulkx_last_error() + ")"
                                                    : std::string())),
        status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status, const char* where) {
  if (status != BULKX_SUCCESS) throw system_error(status, where);
}

inline hipStream_t& default_stream() {
  static thread_local hipStream_t s = nullptr;
  return s;
}

// Waits for the stream and throws if a look-back wait on this device gave
// up since the last check (the results of that launch are wrong): the
// counterpart of the reference's synchronize + throw_on_error
// (bulk/detail/synchronize.hpp:47-54, throw_on_error.hpp:33-49).
inline void synchronize(hipStream_t s = default_stream()) {
  check(bulkx_stream_status(s), "bulkx::synchronize");
}

// ------------------------------------------------------------- functors
template <class T = void>
struct plus {
  __host__ __device__ T operator()(const T& a, const T& b) const { return a + b; }
};
template <class T = void>
struct minimum {
  __host__ __device__ T operator()(const T& a, const T& b) const { return b < a ? b : a; }
};
template <class T = void>
struct maximum {
  __host__ __device__ T operator()(const T& a, const T& b) const { return a < b ? b : a; }
};
template <class T = void>
struct less {
  __host__ __device__ bool operator()(const T& a, const T& b) const { return a < b; }
};
template <class T = void>
struct greater {
  __host__ __device__ bool operator()(const T& a, const T& b) const { return b < a; }
};
template <class T = void>
struct equal_to {
  __host__ __device__ bool operator()(const T& a, const T& b) const { return a == b; }
};

namespace detail {

template <class T>
struct dtype_of;
template <> struct dtype_of<int32_t> { static constexpr int value = BULKX_INT32; };
template <> struct dtype_of<uint32_t> { static constexpr int value = BULKX_UINT32; };
template <> struct dtype_of<int64_t> { static constexpr int value = BULKX_INT64; };
template <> struct dtype_of<uint64_t> { static constexpr int value = BULKX_UINT64; };
template <> struct dtype_of<float> { static constexpr int value = BULKX_FLOAT32; };
template <> struct dtype_of<double> { static constexpr int value = BULKX_FLOAT64; };
template <> struct dtype_of<long long> { static constexpr int value = BULKX_INT64; };
template <> struct dtype_of<unsigned long long> { static constexpr int value = BULKX_UINT64; };

template <template <class> class F, class Op>
struct is_functor : std::false_type {};
template <template <class> class F, class T>
struct is_functor<F, F<T>> : std::true_type {};

template <class Op>
constexpr int op_code() {
  if constexpr (is_functor<plus, Op>::value || is_functor<std::plus, Op>::value) return BULKX_PLUS;
  else if constexpr (is_functor<minimum, Op>::value) return BULKX_MIN;
  else if constexpr (is_functor<maximum, Op>::value) return BULKX_MAX;
#ifdef THRUST_VERSION
  else if constexpr (is_functor<thrust::plus, Op>::value) return BULKX_PLUS;
  else if constexpr (is_functor<thrust::minimum, Op>::value) return BULKX_MIN;
  else if constexpr (is_functor<thrust::maximum, Op>::value) return BULKX_MAX;
#endif
  else return -1;
}

template <class Cmp>
constexpr int cmp_code() {
  if constexpr (is_functor<less, Cmp>::value || is_functor<std::less, Cmp>::value) return BULKX_LESS;
  else if constexpr (is_functor<greater, Cmp>::value || is_functor<std::greater, Cmp>::value)
    return BULKX_GREATER;
#ifdef THRUST_VERSION
  else if constexpr (is_functor<thrust::less, Cmp>::value) return BULKX_LESS;
  else if constexpr (is_functor<thrust::greater, Cmp>::value) return BULKX_GREATER;
#endif
  else return -1;
}

template <class Pred>
constexpr int pred_code() {
  if constexpr (is_functor<equal_to, Pred>::value || is_functor<std::equal_to, Pred>::value)
    return BULKX_EQUAL_TO;
#ifdef THRUST_VERSION
  else if constexpr (is_functor<thrust::equal_to, Pred>::value) return BULKX_EQUAL_TO;
#endif
  else return -1;
}

// Raw device pointer of a contiguous iterator.
template <class T>
T* raw(T* p) {
  return p;
}
#ifdef THRUST_VERSION
template <class It>
auto raw(It it) -> decltype(thrust::raw_pointer_cast(&*it)) {
  return thrust::raw_pointer_cast(&*it);
}
#endif

template <class It>
using value_t = std::remove_cv_t<std::remove_pointer_t<decltype(raw(std::declval<It>()))>>;

// Element types the C-ABI carries (everything else takes the generic path).
template <class T>
struct has_dtype : std::false_type {};
template <> struct has_dtype<int32_t> : std::true_type {};
template <> struct has_dtype<uint32_t> : std::true_type {};
template <> struct has_dtype<int64_t> : std::true_type {};
template <> struct has_dtype<uint64_t> : std::true_type {};
template <> struct has_dtype<float> : std::true_type {};
template <> struct has_dtype<double> : std::true_type {};
template <> struct has_dtype<long long> : std::true_type {};
template <> struct has_dtype<unsigned long long> : std::true_type {};

// Whether a call maps onto the C-ABI: an enum operator over enum types.
template <class Op, class... Ts>
constexpr bool abi_op() {
  return op_code<Op>() >= 0 && (has_dtype<Ts>::value && ...);
}
template <class Cmp, class... Ts>
constexpr bool abi_cmp() {
  return cmp_code<Cmp>() >= 0 && (has_dtype<Ts>::value && ...);
}

}  // namespace detail

// Each driver takes the C-ABI path when its operator (plus / minimum /
// maximum, less / greater, equal_to) and element types map onto the enums,
// and the header-only generic path of generic.hpp otherwise (any
// associative operator, any comparator or predicate, any element type).

template <class It1, class It2, class T, class Op>
It2 inclusive_scan(It1 first, It1 last, It2 result, T init, Op op) {
  using V = detail::value_t<It1>;
  using R = detail::value_t<It2>;
  const int64_t n = int64_t(last - first);
  if constexpr (detail::abi_op<Op, V>() && std::is_same<V, R>::value) {
    const V iv = V(init);
    check(bulkx_inclusive_scan(detail::dtype_of<V>::value, detail::op_code<Op>(), detail::raw(first),
                               detail::raw(result), n, &iv, nullptr, 0, default_stream()),
          "inclusive_scan");
  } else {
    generic::scan<true>(detail::raw(first), n, detail::raw(result), R(init), op, default_stream());
  }
  return result + n;
}

// exclusive_scan(first, last, result, init, op): result[0] = init,
// result[i] = init op x[0] op ... op x[i-1] (the device-wide form of
// bulk::exclusive_scan, scan.hpp:553-591).
template <class It1, class It2, class T, class Op>
It2 exclusive_scan(It1 first, It1 last, It2 result, T init, Op op) {
  using V = detail::value_t<It1>;
  using R = detail::value_t<It2>;
  const int64_t n = int64_t(last - first);
  if constexpr (detail::abi_op<Op, V>() && std::is_same<V, R>::value) {
    const V iv = V(init);
    check(bulkx_exclusive_scan(detail::dtype_of<V>::value, detail::op_code<Op>(), detail::raw(first),
                               detail::raw(result), n, &iv, nullptr, 0, default_stream()),
          "exclusive_scan");
  } else {
    generic::scan<false>(detail::raw(first), n, detail::raw(result), R(init), op, default_stream());
  }
  return result + n;
}

template <class It, class T, class Op>
T my_reduce(It first, It last, T init, Op op) {
  using V = detail::value_t<It>;
  using A = typename std::conditional<detail::abi_op<Op, V>(), V, T>::type;  // device result type
  const A iv = A(init);
  A* d = nullptr;
  check(hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(A), default_stream()) == hipSuccess
            ? BULKX_SUCCESS
            : BULKX_ERR_OUT_OF_MEMORY,
        "my_reduce");
  int st = BULKX_SUCCESS;
  if constexpr (detail::abi_op<Op, V>()) {
    st = bulkx_reduce(detail::dtype_of<V>::value, detail::op_code<Op>(), detail::raw(first),
                      int64_t(last - first), &iv, d, nullptr, 0, default_stream());
  } else {
    generic::reduce_into(detail::raw(first), int64_t(last - first), iv, op, d, default_stream());
  }
  A h = iv;
  if (st == BULKX_SUCCESS &&
      hipMemcpyAsync(&h, d, sizeof(A), hipMemcpyDeviceToHost, default_stream()) != hipSuccess)
    st = BULKX_ERR_HIP;
  if (hipStreamSynchronize(default_stream()) != hipSuccess && st == BULKX_SUCCESS) st = BULKX_ERR_HIP;
  (void)hipFreeAsync(d, default_stream());
  check(st, "my_reduce");
  return T(h);
}

template <class It1, class It2, class It3, class Compare>
It3 my_merge(It1 first1, It1 last1, It2 first2, It2 last2, It3 result, Compare comp) {
  using K = detail::value_t<It1>;
  const int64_t n1 = int64_t(last1 - first1), n2 = int64_t(last2 - first2);
  if constexpr (detail::abi_cmp<Compare, K>() && std::is_same<K, detail::value_t<It2>>::value &&
                std::is_same<K, detail::value_t<It3>>::value) {
    check(bulkx_merge(detail::dtype_of<K>::value, detail::cmp_code<Compare>(), detail::raw(first1),
                      n1, detail::raw(first2), n2, detail::raw(result), nullptr, 0,
                      default_stream()),
          "my_merge");
  } else {
    generic::merge(detail::raw(first1), n1, detail::raw(first2), n2, detail::raw(result), comp,
                   default_stream());
  }
  return result + (n1 + n2);
}

template <class KIt1, class VIt1, class KIt2, class VIt2, class KIt3, class VIt3, class Compare>
std::pair<KIt3, VIt3> merge_by_key(KIt1 keys_first1, KIt1 keys_last1, VIt1 values_first1,
                                   KIt2 keys_first2, KIt2 keys_last2, VIt2 values_first2,
                                   KIt3 keys_result, VIt3 values_result, Compare comp) {
  using K = detail::value_t<KIt1>;
  using V = detail::value_t<VIt1>;
  const int64_t n1 = int64_t(keys_last1 - keys_first1), n2 = int64_t(keys_last2 - keys_first2);
  if constexpr (detail::abi_cmp<Compare, K, V>() &&
                std::is_same<K, detail::value_t<KIt2>>::value &&
                std::is_same<V, detail::value_t<VIt2>>::value) {
    check(bulkx_merge_by_key(detail::dtype_of<K>::value, detail::dtype_of<V>::value,
                             detail::cmp_code<Compare>(), detail::raw(keys_first1),
                             detail::raw(values_first1), n1, detail::raw(keys_first2),
                             detail::raw(values_first2), n2, detail::raw(keys_result),
                             detail::raw(values_result), nullptr, 0, default_stream()),
          "merge_by_key");
  } else {
    generic::merge_by_key(detail::raw(keys_first1), n1, detail::raw(values_first1),
                          detail::raw(keys_first2), n2, detail::raw(values_first2),
                          detail::raw(keys_result), detail::raw(values_result), comp,
                          default_stream());
  }
  return {keys_result + (n1 + n2), values_result + (n1 + n2)};
}

template <class KIt, class VIt, class Compare>
void stable_merge_sort_by_key(KIt keys_first, KIt keys_last, VIt values_first, Compare comp) {
  using K = detail::value_t<KIt>;
  using V = detail::value_t<VIt>;
  const int64_t n = int64_t(keys_last - keys_first);
  if constexpr (detail::abi_cmp<Compare, K, V>()) {
    check(bulkx_stable_sort_by_key(detail::dtype_of<K>::value, detail::dtype_of<V>::value,
                                   detail::cmp_code<Compare>(), detail::raw(keys_first),
                                   detail::raw(values_first), n, nullptr, 0, default_stream()),
          "stable_merge_sort_by_key");
  } else {
    generic::stable_sort_by_key(detail::raw(keys_first), detail::raw(values_first), n, comp,
                                default_stream());
  }
}

template <class KIt1, class VIt1, class KIt2, class VIt2, class Pred, class Op>
std::pair<KIt2, VIt2> my_reduce_by_key(KIt1 keys_first, KIt1 keys_last, VIt1 values_first,
                                       KIt2 keys_result, VIt2 values_result, Pred pred, Op op) {
  using K = detail::value_t<KIt1>;
  using V = detail::value_t<VIt1>;
  int64_t* d = nullptr;
  check(hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(int64_t), default_stream()) == hipSuccess
            ? BULKX_SUCCESS
            : BULKX_ERR_OUT_OF_MEMORY,
        "my_reduce_by_key");
  int st = BULKX_SUCCESS;
  const int64_t n = int64_t(keys_last - keys_first);
  if constexpr (detail::pred_code<Pred>() >= 0 && detail::abi_op<Op, K, V>()) {
    st = bulkx_reduce_by_key(detail::dtype_of<K>::value, detail::dtype_of<V>::value,
                             detail::pred_code<Pred>(), detail::op_code<Op>(),
                             detail::raw(keys_first), detail::raw(values_first), n,
                             detail::raw(keys_result), detail::raw(values_result), d, nullptr, 0,
                             default_stream());
  } else {
    generic::reduce_by_key(detail::raw(keys_first), detail::raw(values_first), n,
                           detail::raw(keys_result), detail::raw(values_result), pred, op,
                           reinterpret_cast<long long*>(d), default_stream());
  }
  int64_t m = 0;
  if (st == BULKX_SUCCESS &&
      hipMemcpyAsync(&m, d, sizeof(m), hipMemcpyDeviceToHost, default_stream()) != hipSuccess)
    st = BULKX_ERR_HIP;
  if (hipStreamSynchronize(default_stream()) != hipSuccess && st == BULKX_SUCCESS) st = BULKX_ERR_HIP;
  // the call has blocked anyway: a look-back that gave up is reported here
  if (st == BULKX_SUCCESS) st = bulkx_stream_status(default_stream());
  (void)hipFreeAsync(d, default_stream());
  check(st, "my_reduce_by_key");
  return {keys_result + m, values_result + m};
}

// A scalar result that is still on its way to the host: the value is copied
// from device memory into pinned host memory behind the producing kernel on
// the stream, and an event marks its arrival.  get() waits for that event
// only; the stream is never synchronised, so the caller keeps issuing work
// while the value travels.  This is the non-blocking form of the scalar
// read that my_reduce and my_reduce_by_key make (reduce.cu:96,
// reduce_by_key.cu:269).  Move-only; the destructor waits for the copy
// before releasing the pinned slot and swallows errors, as the reference's
// future destructor does (bulk/future.hpp:53-75).
template <class T>
class scalar_future {
 public:
  scalar_future() = default;
  scalar_future(const scalar_future&) = delete;
  scalar_future& operator=(const scalar_future&) = delete;
  scalar_future(scalar_future&& o) noexcept { swap(o); }
  scalar_future& operator=(scalar_future&& o) noexcept {
    release();
    swap(o);
    return *this;
  }
  ~scalar_future() { release(); }

  bool valid() const { return m_host != nullptr; }
  // true once the value has reached the host (never blocks)
  bool ready() const {
    if (!valid()) return false;
    const hipError_t e = hipEventQuery(m_event);
    if (e == hipErrorNotReady) return false;
    check(e == hipSuccess ? BULKX_SUCCESS : BULKX_ERR_HIP, "scalar_future::ready");
    return true;
  }
  void wait() const {
    if (!valid()) throw system_error(BULKX_ERR_INVALID_ARGUMENT, "scalar_future::wait");
    check(hipEventSynchronize(m_event) == hipSuccess ? BULKX_SUCCESS : BULKX_ERR_HIP,
          "scalar_future::wait");
  }
  T get() const {
    wait();
    return *m_host;
  }

  // Enqueues the copy of *device_value on `stream` and records the event.
  // The device scalar stays the caller's.
  static scalar_future from_device(const T* device_value, hipStream_t stream) {
    scalar_future f;
    check(hipHostMalloc(reinterpret_cast<void**>(&f.m_host), sizeof(T), hipHostMallocDefault) ==
                  hipSuccess
              ? BULKX_SUCCESS
              : BULKX_ERR_OUT_OF_MEMORY,
          "scalar_future");
    check(hipEventCreateWithFlags(&f.m_event, hipEventDisableTiming) == hipSuccess
              ? BULKX_SUCCESS
              : BULKX_ERR_HIP,
          "scalar_future");
    check(hipMemcpyAsync(f.m_host, device_value, sizeof(T), hipMemcpyDeviceToHost, stream) ==
                      hipSuccess &&
                  hipEventRecord(f.m_event, stream) == hipSuccess
              ? BULKX_SUCCESS
              : BULKX_ERR_HIP,
          "scalar_future");
    return f;
  }

 private:
  void swap(scalar_future& o) noexcept {
    std::swap(m_host, o.m_host);
    std::swap(m_event, o.m_event);
  }
  void release() noexcept {
    if (m_event) {
      (void)hipEventSynchronize(m_event);
      (void)hipEventDestroy(m_event);
    }
    if (m_host) (void)hipHostFree(m_host);
    m_host = nullptr;
    m_event = nullptr;
  }
  T* m_host = nullptr;
  hipEvent_t m_event = nullptr;
};

namespace detail {
// A stream-ordered device scalar of type T, freed behind the work queued so far.
template <class T>
struct stream_scalar {
  T* p = nullptr;
  hipStream_t s;
  explicit stream_scalar(hipStream_t s_) : s(s_) {
    check(hipMallocAsync(reinterpret_cast<void**>(&p), sizeof(T), s) == hipSuccess
              ? BULKX_SUCCESS
              : BULKX_ERR_OUT_OF_MEMORY,
          "device scalar");
  }
  ~stream_scalar() { (void)hipFreeAsync(p, s); }
};
}  // namespace detail

// async_reduce(first, last, init, op[, stream]): my_reduce without the
// blocking read; the result arrives through the returned future.
template <class It, class T, class Op>
scalar_future<T> async_reduce(It first, It last, T init, Op op,
                              hipStream_t stream = default_stream()) {
  using V = detail::value_t<It>;
  detail::stream_scalar<T> d(stream);
  if constexpr (detail::abi_op<Op, V>() && std::is_same<V, T>::value) {
    const V iv = V(init);
    check(bulkx_reduce(detail::dtype_of<V>::value, detail::op_code<Op>(), detail::raw(first),
                       int64_t(last - first), &iv, d.p, nullptr, 0, stream),
          "async_reduce");
  } else {
    generic::reduce_into(detail::raw(first), int64_t(last - first), init, op, d.p, stream);
  }
  return scalar_future<T>::from_device(d.p, stream);
}

// The result of async_reduce_by_key: get() gives the ends of the written
// key and value ranges, as my_reduce_by_key returns them.
template <class KIt, class VIt>
class reduce_by_key_future {
 public:
  reduce_by_key_future(scalar_future<int64_t>&& count, KIt keys_result, VIt values_result)
      : m_count(std::move(count)), m_keys(keys_result), m_values(values_result) {}
  bool valid() const { return m_count.valid(); }
  bool ready() const { return m_count.ready(); }
  void wait() const { m_count.wait(); }
  int64_t num_runs() const { return m_count.get(); }
  std::pair<KIt, VIt> get() const {
    const int64_t m = m_count.get();
    return {m_keys + m, m_values + m};
  }

 private:
  scalar_future<int64_t> m_count;
  KIt m_keys;
  VIt m_values;
};

// async_reduce_by_key(...): my_reduce_by_key without the blocking read of
// the run count.
template <class KIt1, class VIt1, class KIt2, class VIt2, class Pred, class Op>
reduce_by_key_future<KIt2, VIt2> async_reduce_by_key(KIt1 keys_first, KIt1 keys_last,
                                                     VIt1 values_first, KIt2 keys_result,
                                                     VIt2 values_result, Pred pred, Op op,
                                                     hipStream_t stream = default_stream()) {
  using K = detail::value_t<KIt1>;
  using V = detail::value_t<VIt1>;
  detail::stream_scalar<int64_t> d(stream);
  if constexpr (detail::pred_code<Pred>() >= 0 && detail::abi_op<Op, K, V>()) {
    check(bulkx_reduce_by_key(detail::dtype_of<K>::value, detail::dtype_of<V>::value,
                              detail::pred_code<Pred>(), detail::op_code<Op>(),
                              detail::raw(keys_first), detail::raw(values_first),
                              int64_t(keys_last - keys_first), detail::raw(keys_result),
                              detail::raw(values_result), d.p, nullptr, 0, stream),
          "async_reduce_by_key");
  } else {
    generic::reduce_by_key(detail::raw(keys_first), detail::raw(values_first),
                           int64_t(keys_last - keys_first), detail::raw(keys_result),
                           detail::raw(values_result), pred, op,
                           reinterpret_cast<long long*>(d.p), stream);
  }
  return reduce_by_key_future<KIt2, VIt2>(scalar_future<int64_t>::from_device(d.p, stream),
                                          keys_result, values_result);
}

// Decompositions of [0, n) into partitions (decomposition.hpp:5-218): each
// has operator[](i) -> [first, second), size() (the number of partitions)
// and n().
//
// trivial_decomposition(n): one partition, [0, n).
template <class Size>
class trivial_decomposition {
 public:
  typedef Size size_type;
  typedef std::pair<Size, Size> range;
  trivial_decomposition() : m_n(0) {}
  explicit trivial_decomposition(Size n) : m_n(n) {}
  __host__ __device__ range operator[](Size) const { return range(0, m_n); }
  __host__ __device__ Size size() const { return 1; }
  __host__ __device__ Size n() const { return m_n; }

 private:
  Size m_n;
};
template <class Size>
trivial_decomposition<Size> make_trivial_decomposition(Size n) {
  return trivial_decomposition<Size>(n);
}

// blocked_decomposition(n, block): partitions of `block` elements, the last
// one short.
template <class Size>
class blocked_decomposition {
 public:
  typedef Size size_type;
  typedef std::pair<Size, Size> range;
  blocked_decomposition() : m_n(0), m_block(0), m_p(0) {}
  blocked_decomposition(Size n, Size block_size)
      : m_n(n), m_block(block_size), m_p(block_size > 0 ? (n + block_size - 1) / block_size : 0) {}
  __host__ __device__ range operator[](Size i) const {
    const Size b = i * m_block;
    return range(b, b + m_block < m_n ? b + m_block : m_n);
  }
  __host__ __device__ Size size() const { return m_p; }
  __host__ __device__ Size n() const { return m_n; }
  __host__ __device__ Size block_size() const { return m_block; }

 private:
  Size m_n, m_block, m_p;
};
template <class Size>
blocked_decomposition<Size> make_blocked_decomposition(Size n, Size block_size) {
  return blocked_decomposition<Size>(n, block_size);
}

// uniform_decomposition(n, P): blocks of n / P elements (so ceil(n / (n / P))
// partitions, the last one short), as the reference defines it.
template <class Size>
class uniform_decomposition : public blocked_decomposition<Size> {
 public:
  uniform_decomposition() = default;
  uniform_decomposition(Size n, Size num_partitions)
      : blocked_decomposition<Size>(n, num_partitions > 0 ? n / num_partitions : 0) {}
};
template <class Size>
uniform_decomposition<Size> make_uniform_decomposition(Size n, Size num_partitions) {
  return uniform_decomposition<Size>(n, num_partitions);
}

// aligned_decomposition(n, num_partitions, tile): ceil(n / tile) tiles split
// over the partitions, the first (tiles % P) partitions taking one more
// (decomposition.hpp:142-218).
template <class Size>
class aligned_decomposition {
 public:
  typedef Size size_type;
  typedef std::pair<Size, Size> range;
  aligned_decomposition(Size n, Size num_partitions, Size aligned_size)
      : m_n(n), m_p(num_partitions), m_tile(aligned_size) {}
  __host__ __device__ range operator[](Size i) const {
    const Size tiles = (m_n + m_tile - 1) / m_tile;
    const Size per = tiles / m_p, extra = tiles % m_p;
    Size b = per * i + (i < extra ? i : extra);
    Size e = b + per + (i < extra ? 1 : 0);
    b *= m_tile;
    e = e * m_tile < m_n ? e * m_tile : m_n;
    return range(b, e);
  }
  __host__ __device__ Size size() const { return m_p; }
  __host__ __device__ Size n() const { return m_n; }
  __host__ __device__ Size tile_size() const { return m_tile; }

 private:
  Size m_n, m_p, m_tile;
};

template <class Size>
aligned_decomposition<Size> make_aligned_decomposition(Size n, Size num_partitions, Size aligned_size) {
  return aligned_decomposition<Size>(n, num_partitions, aligned_size);
}

namespace detail {
template <class D>
struct is_aligned_decomposition : std::false_type {};
template <class Size>
struct is_aligned_decomposition<aligned_decomposition<Size>> : std::true_type {};
}  // namespace detail

// reduce_intervals(first, decomp, result, op): result[g] = the op-reduction
// of partition g of the decomposition (reduce_intervals.hpp:31-40).
// Aligned, blocked, uniform and trivial decompositions under an enum
// operator run the library kernel (empty partitions give the identity);
// any other decomposition or operator runs the generic path (empty
// partitions leave their slot untouched).
template <class It1, class Decomposition, class It2, class Op>
It2 reduce_intervals(It1 first, const Decomposition& decomp, It2 result, Op op) {
  using V = detail::value_t<It1>;
  using Size = typename Decomposition::size_type;
  constexpr bool aligned = detail::is_aligned_decomposition<Decomposition>::value;
  constexpr bool blocked = std::is_base_of<blocked_decomposition<Size>, Decomposition>::value;
  constexpr bool trivial = std::is_same<Decomposition, trivial_decomposition<Size>>::value;
  if constexpr ((aligned || blocked || trivial) && detail::abi_op<Op, V>() &&
                std::is_same<V, detail::value_t<It2>>::value) {
    int64_t tile;
    if constexpr (aligned) tile = int64_t(decomp.tile_size());
    else if constexpr (blocked) tile = int64_t(decomp.block_size());
    else tile = decomp.n() > 0 ? int64_t(decomp.n()) : 1;
    if (tile > 0) {
      check(bulkx_reduce_intervals(detail::dtype_of<V>::value, detail::op_code<Op>(),
                                   detail::raw(first), int64_t(decomp.n()), int64_t(decomp.size()),
                                   tile, detail::raw(result), default_stream()),
            "reduce_intervals");
      return result + decomp.size();
    }
  }
  generic::reduce_intervals(detail::raw(first), decomp, detail::raw(result), op, default_stream());
  return result + decomp.size();
}

// reduce_intervals(first, last, interval_size, result, op): the blocked
// decomposition of [first, last) (reduce_intervals.hpp:43-47).
template <class It1, class Size, class It2, class Op>
It2 reduce_intervals(It1 first, It1 last, Size interval_size, It2 result, Op op) {
  return reduce_intervals(first, make_blocked_decomposition<Size>(Size(last - first), interval_size),
                          result, op);
}

}  // namespace bulkx
