// bulk.hpp -- the bulk::async call surface on gfx950, header-only HIP C++.
//
// Mirrors the reference's execution-agent hierarchy and launch API
// (bulk/execution_policy.hpp:62-675, bulk/async.hpp:31-89,
// bulk/detail/async.inl:32-190, bulk/future.hpp:44-173, bulk/malloc.hpp)
// so that user functors written against it compile unchanged:
//
//   bulk::async(bulk::par(n), f, bulk::root.this_exec, args...)
//   bulk::async(bulk::grid<256, 8>(num_groups, heap_bytes), f, bulk::root.this_exec, ...)
//   bulk::async(bulk::con<512, 3>(heap_bytes), f, bulk::root, ...)
//   bulk::future<void> t = bulk::async(bulk::par(stream, 1), task); t.wait();
//
// Mapping onto the hardware:
//   * agent                -> one lane (work-item); agent<G> carries a grain of G items
//   * parallel_group<agent> of n agents -> a 1-D grid of n work-items (256 per workgroup)
//   * concurrent_group<agent, S>       -> one workgroup of S work-items (S <= 1024,
//                                         16 wavefronts of 64); wait() is a workgroup barrier
//   * parallel_group<concurrent_group> -> one workgroup per group (grid)
//   * bulk::malloc(g, bytes)           -> a first-fit heap in the workgroup's dynamic LDS
//                                         (heap_size bytes, or by default the LDS share
//                                         that keeps the kernel's occupancy; up to 160 KB
//                                         per CU on gfx950), falling back to device
//                                         malloc when exhausted
// Group algorithms (copy_n, reduce, accumulate, inclusive_scan,
// exclusive_scan, merge_path, merge, merge_by_key, inplace_merge,
// stable_sort_by_key, reduce_by_key, scatter_if, gather,
// adjacent_difference, for_each_n) run on a
// concurrent_group with wavefront shuffles and LDS; the same names on a
// bounded agent (bulk::bound<N>(self)) are the sequential bodies of one
// work-item.  head_flags / tail_flags are index-computed flag ranges and
// join_iterator views two ranges as one.
// Also: choose_sizes (bulk/choose_sizes.hpp), strided_iterator
// (bulk/iterator/strided_iterator.hpp) and uninitialized /
// uninitialized_array (bulk/uninitialized.hpp).
// Contracts: bulk/algorithm/*.hpp of the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include <thrust/pair.h>
#include <thrust/tuple.h>

namespace bulk {

static constexpr std::size_t use_default = ~std::size_t(0);
static constexpr std::size_t dynamic_group_size = 0;
static constexpr int invalid_index = -1;
// A group or grid size left to the launcher (use_default at construction):
// resolved at launch from the kernel's occupancy.
static constexpr int default_size = -1;

// ------------------------------------------------------------------ errors
class system_error : public std::runtime_error {
 public:
  system_error(hipError_t e, const std::string& where)
      : std::runtime_error(where + ": " + hipGetErrorString(e)), code_(e) {}
  hipError_t code() const { return code_; }

 private:
  hipError_t code_;
};

namespace detail {
inline void throw_on_error(hipError_t e, const char* where) {
  if (e != hipSuccess) throw system_error(e, where);
}
}  // namespace detail

// ------------------------------------------------------------ raw storage
// uninitialized<T>: storage for one T whose constructor has not run
// (bulk/uninitialized.hpp:30-160), so that types with constructors can
// live in __shared__ memory; construct(args...) / destroy() run them
// explicitly.
template <class T>
class uninitialized {
 public:
  __host__ __device__ T& get() { return *ptr(); }
  __host__ __device__ const T& get() const { return *ptr(); }
  __host__ __device__ operator T&() { return get(); }
  __host__ __device__ operator const T&() const { return get(); }
  __host__ __device__ uninitialized& operator=(const T& other) {
    get() = other;
    return *this;
  }
  template <class... Args>
  __host__ __device__ void construct(const Args&... args) {
    ::new (static_cast<void*>(m_storage)) T(args...);
  }
  __host__ __device__ void destroy() { get().~T(); }

 private:
  __host__ __device__ T* ptr() { return reinterpret_cast<T*>(m_storage); }
  __host__ __device__ const T* ptr() const { return reinterpret_cast<const T*>(m_storage); }
  alignas(T) unsigned char m_storage[sizeof(T)];
};

// uninitialized_array<T, N>: N unconstructed T with array access
// (bulk/uninitialized.hpp:163-296).
template <class T, std::size_t N>
class uninitialized_array {
 public:
  typedef T value_type;
  typedef T& reference;
  typedef const T& const_reference;
  typedef T* pointer;
  typedef const T* const_pointer;
  typedef pointer iterator;
  typedef const_pointer const_iterator;
  typedef std::size_t size_type;
  __host__ __device__ iterator begin() { return data(); }
  __host__ __device__ const_iterator begin() const { return data(); }
  __host__ __device__ iterator end() { return begin() + size(); }
  __host__ __device__ const_iterator end() const { return begin() + size(); }
  __host__ __device__ const_iterator cbegin() const { return begin(); }
  __host__ __device__ const_iterator cend() const { return end(); }
  __host__ __device__ size_type size() const { return N; }
  __host__ __device__ bool empty() const { return N == 0; }
  __host__ __device__ T* data() { return reinterpret_cast<T*>(m_storage); }
  __host__ __device__ const T* data() const { return reinterpret_cast<const T*>(m_storage); }
  __host__ __device__ reference operator[](size_type i) { return data()[i]; }
  __host__ __device__ const_reference operator[](size_type i) const { return data()[i]; }
  __host__ __device__ reference front() { return data()[0]; }
  __host__ __device__ const_reference front() const { return data()[0]; }
  __host__ __device__ reference back() { return data()[N - 1]; }
  __host__ __device__ const_reference back() const { return data()[N - 1]; }

 private:
  alignas(T) unsigned char m_storage[sizeof(T) * (N ? N : 1)];
};

// ------------------------------------------------------------- iterators
// strided_iterator: an iterator that steps `stride` elements of its base
// per increment (bulk/iterator/strided_iterator.hpp:28-105); the distance
// between two strided iterators rounds away from zero, as the reference's
// distance_to does.
template <class Iterator,
          class Size = typename std::iterator_traits<Iterator>::difference_type>
class strided_iterator {
 public:
  typedef typename std::iterator_traits<Iterator>::value_type value_type;
  typedef typename std::iterator_traits<Iterator>::reference reference;
  typedef typename std::iterator_traits<Iterator>::pointer pointer;
  typedef typename std::iterator_traits<Iterator>::difference_type difference_type;
  typedef std::random_access_iterator_tag iterator_category;
  typedef Size stride_type;

  __host__ __device__ strided_iterator() : m_base(), m_stride(1) {}
  __host__ __device__ strided_iterator(const Iterator& base, stride_type stride)
      : m_base(base), m_stride(stride) {}
  __host__ __device__ stride_type stride() const { return m_stride; }
  __host__ __device__ Iterator base() const { return m_base; }

  __host__ __device__ reference operator*() const { return *m_base; }
  __host__ __device__ reference operator[](difference_type n) const {
    return m_base[n * difference_type(m_stride)];
  }
  __host__ __device__ strided_iterator& operator++() {
    m_base += m_stride;
    return *this;
  }
  __host__ __device__ strided_iterator operator++(int) {
    strided_iterator t = *this;
    ++*this;
    return t;
  }
  __host__ __device__ strided_iterator& operator--() {
    m_base -= m_stride;
    return *this;
  }
  __host__ __device__ strided_iterator operator--(int) {
    strided_iterator t = *this;
    --*this;
    return t;
  }
  __host__ __device__ strided_iterator& operator+=(difference_type n) {
    m_base += n * difference_type(m_stride);
    return *this;
  }
  __host__ __device__ strided_iterator& operator-=(difference_type n) { return *this += -n; }
  __host__ __device__ strided_iterator operator+(difference_type n) const {
    strided_iterator t = *this;
    return t += n;
  }
  __host__ __device__ strided_iterator operator-(difference_type n) const {
    strided_iterator t = *this;
    return t -= n;
  }
  template <class OtherIterator, class OtherSize>
  __host__ __device__ difference_type operator-(
      const strided_iterator<OtherIterator, OtherSize>& other) const {
    const difference_type d = m_base - other.base();
    const difference_type s = difference_type(other.stride());
    return d >= 0 ? (d + (s - 1)) / s : (d - (s - 1)) / s;
  }
  __host__ __device__ bool operator==(const strided_iterator& o) const { return m_base == o.m_base; }
  __host__ __device__ bool operator!=(const strided_iterator& o) const { return m_base != o.m_base; }
  __host__ __device__ bool operator<(const strided_iterator& o) const { return m_base < o.m_base; }
  __host__ __device__ bool operator>(const strided_iterator& o) const { return m_base > o.m_base; }
  __host__ __device__ bool operator<=(const strided_iterator& o) const { return m_base <= o.m_base; }
  __host__ __device__ bool operator>=(const strided_iterator& o) const { return m_base >= o.m_base; }

 private:
  Iterator m_base;
  stride_type m_stride;
};

template <class Iterator, class Size>
__host__ __device__ strided_iterator<Iterator, Size> make_strided_iterator(Iterator it, Size stride) {
  return strided_iterator<Iterator, Size>(it, stride);
}

// ------------------------------------------------------------ agents/groups
template <std::size_t grainsize_ = 1>
class agent {
 public:
  typedef int size_type;
  static const size_type static_grainsize = size_type(grainsize_);
  __host__ __device__ agent(size_type i = invalid_index) : m_index(i) {}
  __host__ __device__ size_type index() const { return m_index; }
  __host__ __device__ size_type grainsize() const { return static_grainsize; }

 private:
  size_type m_index;
};

template <class ExecutionAgent = agent<>, std::size_t size_ = dynamic_group_size>
class parallel_group {
 public:
  typedef ExecutionAgent agent_type;
  typedef int size_type;
  static const size_type static_size = size_type(size_);
  agent_type this_exec;

  __host__ __device__ parallel_group(size_type size = size_type(size_), agent_type exec = agent_type(),
                                     size_type i = invalid_index)
      : this_exec(exec), m_size(size_ ? size_type(size_) : size), m_index(i) {}
  __host__ __device__ size_type index() const { return m_index; }
  __host__ __device__ size_type size() const { return m_size; }
  __host__ __device__ size_type global_index() const {
    return index() * size() + this_exec.index();
  }

 private:
  size_type m_size;
  size_type m_index;
};

namespace detail {
struct lds_arena {
  char* base;
  unsigned bytes;
};
}  // namespace detail

// A group of agents that may synchronize (one workgroup).
template <class ExecutionAgent = agent<>, std::size_t size_ = dynamic_group_size>
class concurrent_group : public parallel_group<ExecutionAgent, size_> {
  typedef parallel_group<ExecutionAgent, size_> super_t;

 public:
  typedef typename super_t::agent_type agent_type;
  typedef typename super_t::size_type size_type;

  // dynamically sized: concurrent_group<agent<G>> g(size, heap_size)
  template <std::size_t SS = size_, typename std::enable_if<SS == 0, int>::type = 0>
  __host__ __device__ concurrent_group(size_type size, size_type heap_size = size_type(0),
                                       agent_type exec = agent_type(),
                                       size_type i = invalid_index)
      : super_t(size, exec, i), m_heap_size(heap_size) {}
  // statically sized: concurrent_group<agent<G>, S> g(heap_size)
  template <std::size_t SS = size_, typename std::enable_if<SS != 0, int>::type = 0>
  __host__ __device__ explicit concurrent_group(std::size_t heap_size = 0,
                                                agent_type exec = agent_type(),
                                                size_type i = invalid_index)
      : super_t(size_type(SS), exec, i),
        m_heap_size(heap_size == use_default ? default_heap : size_type(heap_size)) {}

  // heap_size() == default_heap: the launcher sizes the LDS heap from the
  // kernel's occupancy (bulk/detail/cuda_launcher/cuda_launcher.hpp:105-143)
  static const size_type default_heap = -1;

  __device__ void wait() const { __syncthreads(); }
  __host__ __device__ size_type heap_size() const { return m_heap_size; }
  static size_type hardware_concurrency() {
    int dev = 0, cus = 0;
    detail::throw_on_error(hipGetDevice(&dev), "hipGetDevice");
    detail::throw_on_error(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                           "hipDeviceGetAttribute");
    return cus;
  }

  detail::lds_arena arena{nullptr, 0};  // set by the launcher on the device

 private:
  size_type m_heap_size;
};

// bounded agent: statically bounds the work of an agent
template <std::size_t bound_, class ExecutionAgent>
class bounded : public ExecutionAgent {
 public:
  typedef typename ExecutionAgent::size_type size_type;
  static const size_type static_bound = size_type(bound_);
  __host__ __device__ size_type bound() const { return static_bound; }
  __host__ __device__ ExecutionAgent& unbound() { return *this; }
  __host__ __device__ const ExecutionAgent& unbound() const { return *this; }
};

template <std::size_t bound_, class ExecutionAgent>
__host__ __device__ bounded<bound_, ExecutionAgent>& bound(ExecutionAgent& exec) {
  return static_cast<bounded<bound_, ExecutionAgent>&>(exec);
}

// --------------------------------------------------------------- futures
template <class T>
class future;

template <>
class future<void> {
 public:
  future() = default;
  future(hipStream_t s, bool owns) : m_stream(s), m_owns_stream(owns) {
    detail::throw_on_error(hipEventCreateWithFlags(&m_event, hipEventDisableTiming),
                           "hipEventCreateWithFlags in bulk::future");
    detail::throw_on_error(hipEventRecord(m_event, s), "hipEventRecord in bulk::future");
  }
  future(const future&) = delete;
  future& operator=(const future&) = delete;
  future(future&& o) noexcept { swap(o); }
  future& operator=(future&& o) noexcept {
    if (this != &o) {
      release();
      swap(o);
    }
    return *this;
  }
  ~future() { release(); }

  void wait() const {
    detail::throw_on_error(hipEventSynchronize(m_event), "hipEventSynchronize in future::wait");
  }
  bool valid() const { return m_event != nullptr; }
  hipEvent_t event() const { return m_event; }
  hipStream_t stream() const { return m_stream; }

 private:
  void swap(future& o) noexcept {
    std::swap(m_stream, o.m_stream);
    std::swap(m_event, o.m_event);
    std::swap(m_owns_stream, o.m_owns_stream);
  }
  void release() noexcept {  // errors are swallowed, as the reference's destructor does
    if (m_event) (void)hipEventDestroy(m_event);
    if (m_owns_stream && m_stream) (void)hipStreamDestroy(m_stream);
    m_event = nullptr;
    m_stream = nullptr;
    m_owns_stream = false;
  }
  hipStream_t m_stream = nullptr;
  hipEvent_t m_event = nullptr;
  bool m_owns_stream = false;
};

// ---------------------------------------------------------- launch policies
template <class ExecutionGroup>
class async_launch {
 public:
  async_launch(ExecutionGroup g, hipStream_t s, hipEvent_t before = nullptr)
      : m_stream_valid(true), m_group(g), m_stream(s), m_before(before) {}
  async_launch(ExecutionGroup g, hipEvent_t before)
      : m_stream_valid(false), m_group(g), m_stream(nullptr), m_before(before) {}
  ExecutionGroup exec() const { return m_group; }
  hipStream_t stream() const { return m_stream; }
  hipEvent_t before_event() const { return m_before; }
  bool is_stream_valid() const { return m_stream_valid; }

 private:
  bool m_stream_valid;
  ExecutionGroup m_group;
  hipStream_t m_stream;
  hipEvent_t m_before;
};

inline parallel_group<> par(std::size_t n) { return parallel_group<>(int(n)); }

template <class ExecutionAgent>
parallel_group<ExecutionAgent> par(ExecutionAgent exec, std::size_t n) {
  return parallel_group<ExecutionAgent>(int(n), exec);
}

inline async_launch<parallel_group<>> par(hipStream_t s, std::size_t n) {
  return async_launch<parallel_group<>>(par(n), s);
}

template <class ExecutionAgent>
async_launch<parallel_group<ExecutionAgent>> par(hipStream_t s, ExecutionAgent exec, std::size_t n) {
  return async_launch<parallel_group<ExecutionAgent>>(par(exec, n), s);
}

inline async_launch<parallel_group<>> par(future<void>& before, std::size_t n) {
  return async_launch<parallel_group<>>(par(n), before.event());
}

inline concurrent_group<> con(std::size_t size, std::size_t heap_size = use_default) {
  return concurrent_group<>(size == use_default ? default_size : int(size),
                            heap_size == use_default ? concurrent_group<>::default_heap
                                                     : int(heap_size));
}

template <class ExecutionAgent>
concurrent_group<ExecutionAgent> con(ExecutionAgent exec, std::size_t size,
                                     std::size_t heap_size = use_default) {
  return concurrent_group<ExecutionAgent>(
      size == use_default ? default_size : int(size),
      heap_size == use_default ? concurrent_group<ExecutionAgent>::default_heap : int(heap_size),
      exec);
}

template <std::size_t groupsize, std::size_t grainsize>
concurrent_group<agent<grainsize>, groupsize> con(std::size_t heap_size) {
  return concurrent_group<agent<grainsize>, groupsize>(heap_size);
}

inline parallel_group<concurrent_group<>> grid(std::size_t num_groups = use_default,
                                               std::size_t group_size = use_default,
                                               std::size_t heap_size = use_default) {
  return parallel_group<concurrent_group<>>(
      num_groups == use_default ? default_size : int(num_groups), con(group_size, heap_size));
}

inline async_launch<parallel_group<concurrent_group<>>> grid(std::size_t num_groups,
                                                             std::size_t group_size,
                                                             std::size_t heap_size, hipStream_t s) {
  return async_launch<parallel_group<concurrent_group<>>>(grid(num_groups, group_size, heap_size), s);
}

template <std::size_t groupsize, std::size_t grainsize>
parallel_group<concurrent_group<agent<grainsize>, groupsize>> grid(std::size_t num_groups,
                                                                   std::size_t heap_size = use_default) {
  return par(con<groupsize, grainsize>(heap_size), num_groups);
}

template <std::size_t groupsize, std::size_t grainsize>
async_launch<parallel_group<concurrent_group<agent<grainsize>, groupsize>>> grid(
    std::size_t num_groups, std::size_t heap_size, hipStream_t s) {
  return async_launch<parallel_group<concurrent_group<agent<grainsize>, groupsize>>>(
      grid<groupsize, grainsize>(num_groups, heap_size), s);
}

// ------------------------------------------------------------- placeholders
namespace detail {
template <unsigned D>
struct cursor;

template <unsigned D, class G>
struct at_depth {
  typedef typename at_depth<D - 1, typename G::agent_type>::type type;
};
template <class G>
struct at_depth<0, G> {
  typedef G type;
};

template <unsigned D, class G>
__device__ typename at_depth<D, G>::type& get_at(G& g) {
  if constexpr (D == 0) return g;
  else return get_at<D - 1>(g.this_exec);
}

template <unsigned D>
struct cursor {
  static const unsigned depth = D;
  cursor<D + 1> this_exec;
};
template <>
struct cursor<3> {
  static const unsigned depth = 3;
};

template <class T>
struct is_cursor : std::false_type {};
template <unsigned D>
struct is_cursor<cursor<D>> : std::true_type {};

// An argument, or the hierarchy level a cursor names.
template <class G, class A>
__device__ decltype(auto) substitute(G& root, A& a) {
  if constexpr (is_cursor<A>::value) return (get_at<A::depth>(root));
  else return (a);
}
}  // namespace detail

// bulk::root stands for the root of the launched hierarchy;
// bulk::root.this_exec for its agents (or groups), and so on.
static constexpr detail::cursor<0> root{};

// ------------------------------------------------------------------ malloc
// The group's on-chip heap: the workgroup's dynamic LDS (the launcher sizes
// it, heap_size bytes or, by default, the LDS share that keeps the kernel's
// occupancy) managed as a first-fit list of blocks, the allocator of
// bulk/malloc.hpp:133-199 written afresh for LDS.  Arena layout: a 16-byte
// header {break offset, lock}, then blocks of a 16-byte header {data bytes,
// free flag, offset of the previous block} followed by their data.  Blocks
// can be freed in any order; a freed block merges with free neighbours, and
// a free block at the end moves the break back.  When the arena has no room,
// allocation falls back to the device heap (::malloc); if that fails too the
// result is null, as with the reference's shmalloc (malloc.hpp:519-531).
namespace detail {

struct lds_block {
  unsigned size;  // data bytes, a multiple of 16
  unsigned free;
  unsigned prev;  // offset of the previous block's header; 0 for the first block
  unsigned pad;
};
static_assert(sizeof(lds_block) == 16, "block headers keep data 16-byte aligned");

constexpr unsigned kArenaHeader = 16;

__device__ inline lds_block* lds_block_at(char* base, unsigned off) {
  return reinterpret_cast<lds_block*>(base + off);
}
__device__ inline unsigned& lds_break(char* base) { return *reinterpret_cast<unsigned*>(base); }
__device__ inline unsigned& lds_lock(char* base) { return reinterpret_cast<unsigned*>(base)[1]; }

// First fit over the block list, else a new block at the break; null when
// the arena has no room.  Not thread safe: one agent at a time.
__device__ inline void* lds_allocate(const lds_arena& a, std::size_t bytes) {
  if (!a.base || a.bytes <= kArenaHeader) return nullptr;
  char* base = a.base;
  const std::size_t need = (bytes + 15) & ~std::size_t(15);
  if (need > a.bytes) return nullptr;
  const unsigned brk = lds_break(base);
  unsigned last = 0;
  for (unsigned off = kArenaHeader; off < brk;) {
    lds_block* b = lds_block_at(base, off);
    if (b->free && b->size >= need) {
      if (b->size - need >= 2 * sizeof(lds_block)) {  // split off the rest as a free block
        const unsigned rest = off + unsigned(sizeof(lds_block) + need);
        lds_block* r = lds_block_at(base, rest);
        r->size = b->size - unsigned(need) - unsigned(sizeof(lds_block));
        r->free = 1;
        r->prev = off;
        const unsigned after = rest + unsigned(sizeof(lds_block)) + r->size;
        if (after < brk) lds_block_at(base, after)->prev = rest;
        b->size = unsigned(need);
      }
      b->free = 0;
      return base + off + sizeof(lds_block);
    }
    last = off;
    off += unsigned(sizeof(lds_block)) + b->size;
  }
  if (std::size_t(brk) + sizeof(lds_block) + need > a.bytes) return nullptr;
  lds_block* b = lds_block_at(base, brk);
  b->size = unsigned(need);
  b->free = 0;
  b->prev = last;
  lds_break(base) = brk + unsigned(sizeof(lds_block) + need);
  return base + brk + sizeof(lds_block);
}

__device__ inline bool lds_owns(const lds_arena& a, const void* p) {
  const char* c = static_cast<const char*>(p);
  return a.base && c >= a.base + kArenaHeader + sizeof(lds_block) && c < a.base + a.bytes;
}

__device__ inline void lds_deallocate(const lds_arena& a, void* p) {
  char* base = a.base;
  unsigned off = unsigned(static_cast<char*>(p) - base) - unsigned(sizeof(lds_block));
  lds_block* b = lds_block_at(base, off);
  b->free = 1;
  unsigned brk = lds_break(base);
  // merge with a free successor
  unsigned next = off + unsigned(sizeof(lds_block)) + b->size;
  if (next < brk && lds_block_at(base, next)->free) {
    b->size += unsigned(sizeof(lds_block)) + lds_block_at(base, next)->size;
    next = off + unsigned(sizeof(lds_block)) + b->size;
    if (next < brk) lds_block_at(base, next)->prev = off;
  }
  // merge into a free predecessor
  if (off > kArenaHeader && lds_block_at(base, b->prev)->free) {
    const unsigned p0 = b->prev;
    lds_block* pb = lds_block_at(base, p0);
    pb->size += unsigned(sizeof(lds_block)) + b->size;
    off = p0;
    b = pb;
    if (next < brk) lds_block_at(base, next)->prev = off;
  }
  // a free block at the end gives its space back to the break
  if (next >= brk) lds_break(base) = off;
}

}  // namespace detail

// Collective allocation (bulk/malloc.hpp:589-604): every agent of the group
// calls it with the same size and receives the same pointer.
template <class Agent, std::size_t S>
__device__ void* malloc(concurrent_group<Agent, S>& g, std::size_t bytes) {
  __shared__ void* s_ptr;
  g.wait();
  if (g.this_exec.index() == 0) {
    void* p = detail::lds_allocate(g.arena, bytes);
    s_ptr = p ? p : ::malloc(bytes);  // device heap when the arena has no room; may be null
  }
  g.wait();
  return s_ptr;
}

// Collective free (bulk/malloc.hpp:607-615); any order, null ignored.
template <class Agent, std::size_t S>
__device__ void free(concurrent_group<Agent, S>& g, void* ptr) {
  g.wait();
  if (g.this_exec.index() == 0 && ptr) {
    if (detail::lds_owns(g.arena, ptr)) detail::lds_deallocate(g.arena, ptr);
    else ::free(ptr);
  }
  g.wait();
}

// Allocation by one agent (bulk/malloc.hpp:519-569: shmalloc / shfree): the
// arena is guarded by an LDS lock, so agents of a group may call it
// independently.  The lock loop keeps the critical section inside the loop
// body, so lanes of one wavefront that contend take turns instead of
// waiting on each other.
template <class Agent, std::size_t S>
__device__ void* shmalloc(concurrent_group<Agent, S>& g, std::size_t bytes) {
  void* p = nullptr;
  if (g.arena.base) {
    unsigned* lock = &detail::lds_lock(g.arena.base);
    bool done = false;
    while (!done) {
      if (atomicCAS(lock, 0u, 1u) == 0u) {
        p = detail::lds_allocate(g.arena, bytes);
        __threadfence_block();
        atomicExch(lock, 0u);
        done = true;
      }
    }
  }
  return p ? p : ::malloc(bytes);
}

template <class Agent, std::size_t S>
__device__ void shfree(concurrent_group<Agent, S>& g, void* ptr) {
  if (!ptr) return;
  if (!detail::lds_owns(g.arena, ptr)) {
    ::free(ptr);
    return;
  }
  unsigned* lock = &detail::lds_lock(g.arena.base);
  bool done = false;
  while (!done) {
    if (atomicCAS(lock, 0u, 1u) == 0u) {
      detail::lds_deallocate(g.arena, ptr);
      __threadfence_block();
      atomicExch(lock, 0u);
      done = true;
    }
  }
}

// -------------------------------------------------------------- the launcher
namespace detail {

template <class G>
struct launch_traits;

// par(n): n agents, 256 per workgroup
template <class A, std::size_t S>
struct launch_traits<parallel_group<A, S>> {
  static dim3 grid(const parallel_group<A, S>& g) { return dim3(unsigned((g.size() + 255) / 256)); }
  static dim3 block(const parallel_group<A, S>&) { return dim3(256); }
  static std::size_t lds(const parallel_group<A, S>&) { return 0; }
};
// con(n): one workgroup.  heap(): the requested heap, or default_heap.
template <class A, std::size_t S>
struct launch_traits<concurrent_group<A, S>> {
  static dim3 grid(const concurrent_group<A, S>&) { return dim3(1); }
  static dim3 block(const concurrent_group<A, S>& g) { return dim3(unsigned(g.size())); }
  static int heap(const concurrent_group<A, S>& g) { return g.heap_size(); }
};
// grid: one workgroup per concurrent group
template <class A, std::size_t S, std::size_t PS>
struct launch_traits<parallel_group<concurrent_group<A, S>, PS>> {
  typedef parallel_group<concurrent_group<A, S>, PS> G;
  static dim3 grid(const G& g) { return dim3(unsigned(g.size())); }
  static dim3 block(const G& g) { return dim3(unsigned(g.this_exec.size())); }
  static int heap(const G& g) { return g.this_exec.heap_size(); }
};

// Dynamic LDS of a launch whose group asked for `requested` heap bytes
// (choose_heap_size, bulk/detail/cuda_launcher/cuda_launcher.hpp:105-143):
// 0 -> no heap; default_heap -> the LDS share of one workgroup at the
// kernel's occupancy with no dynamic LDS (the heap then costs no
// occupancy); an explicit request -> the larger of the request (plus the
// arena's bookkeeping) and that share.  Capped by what one workgroup may
// allocate.
template <class K>
inline std::size_t heap_bytes(K kernel, int requested, int block_threads) {
  if (requested == 0) return 0;
  int dev = 0;
  throw_on_error(hipGetDevice(&dev), "hipGetDevice in bulk::async");
  int lds_cu = 0, lds_block_max = 0;
  throw_on_error(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                       dev),
                 "hipDeviceGetAttribute in bulk::async");
  throw_on_error(hipDeviceGetAttribute(&lds_block_max, hipDeviceAttributeSharedMemPerBlockOptin,
                                       dev),
                 "hipDeviceGetAttribute in bulk::async");
  if (lds_block_max <= 0) lds_block_max = lds_cu;
  hipFuncAttributes attr;
  throw_on_error(hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(kernel)),
                 "hipFuncGetAttributes in bulk::async");
  int blocks = 0;
  throw_on_error(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                     &blocks, reinterpret_cast<const void*>(kernel), block_threads, 0),
                 "hipOccupancyMaxActiveBlocksPerMultiprocessor in bulk::async");
  if (blocks < 1) blocks = 1;
  const long long cap = (long long)lds_block_max - (long long)attr.sharedSizeBytes;
  long long share = (long long)lds_cu / blocks - (long long)attr.sharedSizeBytes;
  share = share < cap ? share : cap;
  share = share > 0 ? share & ~15LL : 0;
  long long want = share;
  if (requested > 0) {
    const long long req = (long long)((requested + 15) & ~15) + 64;  // + block headers
    if (req > want) want = req;
  }
  if (want > cap) want = cap & ~15LL;
  return want > (long long)detail::kArenaHeader ? std::size_t(want) : 0;
}

template <class A, std::size_t S>
__device__ void bind_arena(concurrent_group<A, S>& cg, std::size_t bytes) {
  extern __shared__ __align__(16) char bulk_dynamic_lds[];
  cg.arena.base = bytes ? bulk_dynamic_lds : nullptr;
  cg.arena.bytes = unsigned(bytes);
  if (bytes && threadIdx.x == 0) {
    detail::lds_break(bulk_dynamic_lds) = detail::kArenaHeader;
    detail::lds_lock(bulk_dynamic_lds) = 0;
  }
  __syncthreads();
}

template <class A, std::size_t S, class F, class... Args>
__global__ void par_kernel(parallel_group<A, S> proto, F f, Args... args) {
  const int i = int(blockIdx.x) * 256 + int(threadIdx.x);
  if (i >= proto.size()) return;
  parallel_group<A, S> g(proto.size(), A(i), 0);
  f(substitute(g, args)...);
}

// The group object of one workgroup: agent = work-item, index = workgroup.
template <class A, std::size_t S>
__device__ concurrent_group<A, S> make_group(const concurrent_group<A, S>& proto, int index) {
  if constexpr (S == 0)
    return concurrent_group<A, S>(proto.size(), proto.heap_size(), A(int(threadIdx.x)), index);
  else
    return concurrent_group<A, S>(std::size_t(proto.heap_size()), A(int(threadIdx.x)), index);
}

template <class A, std::size_t S, class F, class... Args>
__global__ void con_kernel(concurrent_group<A, S> proto, std::size_t lds, F f, Args... args) {
  concurrent_group<A, S> g = make_group(proto, 0);
  bind_arena(g, lds);
  f(substitute(g, args)...);
}

template <class A, std::size_t S, std::size_t PS, class F, class... Args>
__global__ void grid_kernel(parallel_group<concurrent_group<A, S>, PS> proto, std::size_t lds, F f,
                            Args... args) {
  concurrent_group<A, S> cg = make_group(proto.this_exec, int(blockIdx.x));
  bind_arena(cg, lds);
  parallel_group<concurrent_group<A, S>, PS> g(proto.size(), cg, 0);
  f(substitute(g, args)...);
}

template <class K>
inline void allow_lds(K kernel, std::size_t bytes) {
  if (bytes > 64 * 1024)
    throw_on_error(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)),
                   "hipFuncSetAttribute in bulk::async");
}

template <class A, std::size_t S, class F, class... Args>
void launch(const parallel_group<A, S>& g, hipStream_t s, F f, Args... args) {
  if (g.size() <= 0) return;
  par_kernel<A, S, F, Args...><<<launch_traits<parallel_group<A, S>>::grid(g), dim3(256), 0, s>>>(
      g, f, args...);
}

// Group size the launcher picks for `kernel` when the group's size was
// left to it: the block size of maximum potential occupancy
// (choose_group_size, bulk/detail/cuda_launcher/cuda_launcher.hpp:147-159).
template <class K>
inline int default_group_size(K kernel) {
  int min_grid = 0, block = 0;
  throw_on_error(hipOccupancyMaxPotentialBlockSize(&min_grid, &block,
                                                   reinterpret_cast<const void*>(kernel), 0, 1024),
                 "hipOccupancyMaxPotentialBlockSize in bulk::async");
  return block > 0 ? block : 256;
}

// Number of groups when left to the launcher: enough groups of `group_size`
// to fill every CU's thread slots (choose_num_groups / choose_subscription,
// cuda_launcher.hpp:163-185).
inline int default_num_groups(int group_size) {
  int dev = 0, cus = 0, per_cu = 0;
  throw_on_error(hipGetDevice(&dev), "hipGetDevice in bulk::async");
  throw_on_error(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                 "hipDeviceGetAttribute in bulk::async");
  throw_on_error(
      hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxThreadsPerMultiProcessor, dev),
      "hipDeviceGetAttribute in bulk::async");
  const int subscription = group_size > 0 ? per_cu / group_size : 0;
  return subscription > 0 ? subscription * cus : cus;
}

// The group with its launcher-chosen sizes filled in.
template <class A, std::size_t S, class K>
concurrent_group<A, S> resolve(const concurrent_group<A, S>& g, K kernel) {
  if constexpr (S == 0) {
    if (g.size() == default_size)
      return concurrent_group<A, S>(default_group_size(kernel), g.heap_size(), g.this_exec,
                                    g.index());
  }
  return g;
}
template <class A, std::size_t S, std::size_t PS, class K>
parallel_group<concurrent_group<A, S>, PS> resolve(const parallel_group<concurrent_group<A, S>, PS>& g,
                                                   K kernel) {
  const concurrent_group<A, S> cg = resolve(g.this_exec, kernel);
  const int n = g.size() == default_size ? default_num_groups(cg.size()) : g.size();
  return parallel_group<concurrent_group<A, S>, PS>(n, cg, g.index());
}

template <class A, std::size_t S, class F, class... Args>
void launch(const concurrent_group<A, S>& g0, hipStream_t s, F f, Args... args) {
  const concurrent_group<A, S> g = resolve(g0, con_kernel<A, S, F, Args...>);
  if (g.size() <= 0 || g.size() > 1024)
    throw system_error(hipErrorInvalidValue, "bulk::async: a concurrent_group holds 1..1024 agents");
  const std::size_t lds = heap_bytes(con_kernel<A, S, F, Args...>,
                                     launch_traits<concurrent_group<A, S>>::heap(g), g.size());
  allow_lds(con_kernel<A, S, F, Args...>, lds);
  con_kernel<A, S, F, Args...><<<dim3(1), dim3(unsigned(g.size())), lds, s>>>(g, lds, f, args...);
}

template <class A, std::size_t S, std::size_t PS, class F, class... Args>
void launch(const parallel_group<concurrent_group<A, S>, PS>& g0, hipStream_t s, F f, Args... args) {
  typedef parallel_group<concurrent_group<A, S>, PS> G;
  const G g = resolve(g0, grid_kernel<A, S, PS, F, Args...>);
  if (g.size() <= 0) return;
  if (g.this_exec.size() <= 0 || g.this_exec.size() > 1024)
    throw system_error(hipErrorInvalidValue, "bulk::async: a concurrent_group holds 1..1024 agents");
  const std::size_t lds = heap_bytes(grid_kernel<A, S, PS, F, Args...>, launch_traits<G>::heap(g),
                                     g.this_exec.size());
  allow_lds(grid_kernel<A, S, PS, F, Args...>, lds);
  grid_kernel<A, S, PS, F, Args...><<<launch_traits<G>::grid(g), launch_traits<G>::block(g), lds, s>>>(
      g, lds, f, args...);
}

template <class G, class F, class... Args>
future<void> async_in_stream(const G& g, hipStream_t s, hipEvent_t before, bool owns, F f,
                             Args... args) {
  if (before) throw_on_error(hipStreamWaitEvent(s, before, 0), "hipStreamWaitEvent in bulk::async");
  launch(g, s, f, args...);
  throw_on_error(hipGetLastError(), "kernel launch in bulk::async");
  return future<void>(s, owns);
}

template <class G>
struct is_async_launch : std::false_type {};
template <class G>
struct is_async_launch<async_launch<G>> : std::true_type {};

}  // namespace detail

// bulk::async(policy, f, args...): launches f over the policy's hierarchy.
// A plain group launches on the null stream; par(stream, ...) on that
// stream; par(future, ...) on a new stream (owned by the returned future)
// that first waits for the future's event -- as bulk/detail/async.inl does.
template <class G, class F, class... Args>
future<void> async(G g, F f, Args... args) {
  if constexpr (detail::is_async_launch<G>::value) {
    if (g.is_stream_valid())
      return detail::async_in_stream(g.exec(), g.stream(), g.before_event(), false, f, args...);
    hipStream_t s = nullptr;
    detail::throw_on_error(hipStreamCreateWithFlags(&s, hipStreamNonBlocking),
                           "hipStreamCreate in bulk::async");
    return detail::async_in_stream(g.exec(), s, g.before_event(), true, f, args...);
  } else {
    return detail::async_in_stream(g, nullptr, nullptr, false, f, args...);
  }
}

// choose_sizes(g, f, args...): the (number of groups, group size) a launch
// of f(args...) over the grid g would use, sizes g leaves to the launcher
// (use_default) filled in from the kernel's occupancy
// (bulk/choose_sizes.hpp:29-75, bulk/detail/choose_sizes.inl:32-117).
template <class A, std::size_t S, std::size_t PS, class F, class... Args>
thrust::pair<int, int> choose_sizes(parallel_group<concurrent_group<A, S>, PS> g, F f,
                                    Args... args) {
  (void)f;
  const auto r = detail::resolve(g, detail::grid_kernel<A, S, PS, F, Args...>);
  return thrust::make_pair(r.size(), r.this_exec.size());
}

// ------------------------------------------------------------ group algorithms
namespace detail {

template <class T>
__device__ T shfl_up(T x, int delta) {
  static_assert(sizeof(T) % 4 == 0, "4-byte multiple");
  constexpr int W = sizeof(T) / 4;
  union {
    T t;
    int w[W];
  } u;
  u.t = x;
#pragma unroll
  for (int i = 0; i < W; ++i) u.w[i] = __shfl_up(u.w[i], delta, 64);
  return u.t;
}

template <class T>
__device__ T shfl_idx(T x, int lane) {
  constexpr int W = sizeof(T) / 4;
  union {
    T t;
    int w[W];
  } u;
  u.t = x;
#pragma unroll
  for (int i = 0; i < W; ++i) u.w[i] = __shfl(u.w[i], lane, 64);
  return u.t;
}

// Inclusive scan across the group's agents (agent order).  Returns the
// agent's inclusive value; *total receives the group total.
template <class G, class T, class Op>
__device__ T group_inclusive(G& g, T x, Op op, T* total) {
  __shared__ T s_wave[16];
  const int tid = g.this_exec.index();
  const int lane = tid & 63, wid = tid >> 6;
  const int nw = (g.size() + 63) / 64;
  for (int d = 1; d < 64; d <<= 1) {
    T y = shfl_up(x, d);
    if (lane >= d) x = op(y, x);
  }
  if (lane == 63 || tid == g.size() - 1) s_wave[wid] = x;
  g.wait();
  T carry = x;  // placeholder, set below
  bool have = false;
  for (int w = 0; w < nw; ++w) {
    const T v = s_wave[w];
    if (w < wid) {
      carry = have ? op(carry, v) : v;
      have = true;
    }
    if (w == nw - 1 && total) {
      T t = s_wave[0];
      for (int u = 1; u < nw; ++u) t = op(t, s_wave[u]);
      *total = t;
    }
  }
  g.wait();
  return have ? op(carry, x) : x;
}

}  // namespace detail

// copy_n: result[i] = first[i] for i < n, striped over the group's agents.
template <class Agent, std::size_t S, class It1, class Size, class It2>
__device__ It2 copy_n(concurrent_group<Agent, S>& g, It1 first, Size n, It2 result) {
  for (Size i = g.this_exec.index(); i < n; i += g.size()) result[i] = first[i];
  g.wait();
  return result + n;
}

// reduce: init op x[0] op ... op x[n-1] (op assumed commutative, as the
// reference's strided group reduce, reduce.hpp:110-219).  Every agent
// receives the result.
template <class Agent, std::size_t S, class It, class T, class Op>
__device__ T reduce(concurrent_group<Agent, S>& g, It first, It last, T init, Op op) {
  const auto n = last - first;
  const int tid = g.this_exec.index();
  bool have = false;
  T acc = init;
  for (decltype(n + 0) i = tid; i < n; i += g.size()) {
    const T v = first[i];
    acc = have ? op(acc, v) : v;
    have = true;
  }
  // agents without elements contribute nothing: fold in order with a flag
  struct P {
    T v;
    int has;
  };
  P p{acc, have ? 1 : 0};
  auto pop = [op](P a, P b) { return b.has ? (a.has ? P{op(a.v, b.v), 1} : b) : a; };
  P total;
  (void)detail::group_inclusive(g, p, pop, &total);
  return total.has ? op(init, total.v) : init;
}

// inclusive_scan with init: result[i] = init op x[0] op ... op x[i].
// Tiles of size*grain elements; each agent scans a blocked grain.
template <class Agent, std::size_t S, class It1, class It2, class T, class Op>
__device__ It2 inclusive_scan(concurrent_group<Agent, S>& g, It1 first, It1 last, It2 result,
                              T init, Op op) {
  const auto n = last - first;
  constexpr int G = Agent::static_grainsize;
  const int tid = g.this_exec.index();
  const int tile = g.size() * G;
  T carry = init;
  for (decltype(n + 0) base = 0; base < n; base += tile) {
    T v[G];
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const auto i = base + decltype(n + 0)(tid) * G + j;
      if (i < n) {
        v[j] = first[i];
        v[j] = j ? op(v[j - 1], v[j]) : v[j];
        ++cnt;
      }
    }
    struct P {
      T v;
      int has;
    };
    auto pop = [op](P a, P b) { return b.has ? (a.has ? P{op(a.v, b.v), 1} : b) : a; };
    P total;
    P inc = detail::group_inclusive(g, P{cnt ? v[cnt - 1] : carry, cnt ? 1 : 0}, pop, &total);
    // exclusive prefix of this agent = inclusive of the previous agent
    P prev = detail::shfl_up(inc, 1);
    __shared__ P s_last[16];
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 63) s_last[wid] = inc;
    g.wait();
    if (lane == 0) prev = wid > 0 ? s_last[wid - 1] : P{carry, 0};
    const T pre = prev.has ? op(carry, prev.v) : carry;
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const auto i = base + decltype(n + 0)(tid) * G + j;
      if (i < n) result[i] = op(pre, v[j]);
    }
    if (total.has) carry = op(carry, total.v);
    g.wait();
  }
  return result + n;
}

// exclusive_scan: result[0] = init, result[i] = init op x[0] op ... op x[i-1].
template <class Agent, std::size_t S, class It1, class It2, class T, class Op>
__device__ It2 exclusive_scan(concurrent_group<Agent, S>& g, It1 first, It1 last, It2 result,
                              T init, Op op) {
  const auto n = last - first;
  constexpr int G = Agent::static_grainsize;
  const int tid = g.this_exec.index();
  const int tile = g.size() * G;
  T carry = init;
  for (decltype(n + 0) base = 0; base < n; base += tile) {
    T v[G];
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const auto i = base + decltype(n + 0)(tid) * G + j;
      if (i < n) {
        v[j] = first[i];
        ++cnt;
      }
    }
    T agg = v[0];
    for (int j = 1; j < cnt; ++j) agg = op(agg, v[j]);
    struct P {
      T v;
      int has;
    };
    auto pop = [op](P a, P b) { return b.has ? (a.has ? P{op(a.v, b.v), 1} : b) : a; };
    P total;
    P inc = detail::group_inclusive(g, P{cnt ? agg : carry, cnt ? 1 : 0}, pop, &total);
    P prev = detail::shfl_up(inc, 1);
    __shared__ P s_last[16];
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 63) s_last[wid] = inc;
    g.wait();
    if (lane == 0) prev = wid > 0 ? s_last[wid - 1] : P{carry, 0};
    T run = prev.has ? op(carry, prev.v) : carry;
    for (int j = 0; j < cnt; ++j) {
      const auto i = base + decltype(n + 0)(tid) * G + j;
      result[i] = run;
      run = op(run, v[j]);
    }
    if (total.has) carry = op(carry, total.v);
    g.wait();
  }
  return result + n;
}

// inclusive_scan without init (scan.hpp:489-526): the first input is the
// init, written through unchanged; in place (result == first) allowed.
template <class Agent, std::size_t S, class It1, class It2, class Op>
__device__ It2 inclusive_scan(concurrent_group<Agent, S>& g, It1 first, It1 last, It2 result,
                              Op op) {
  if (first < last) {
    const typename std::iterator_traits<It1>::value_type init = *first;
    g.wait();  // every agent has read *first before result[0] may overwrite it
    if (g.this_exec.index() == 0) *result = init;
    inclusive_scan(g, first + 1, last, result + 1, init, op);
  }
  return result + (last - first);
}

// The bounded-group forms (scan.hpp:222-313, 391-452, 529-550): a group
// whose input is statically bounded by one tile (bound <= size * grain)
// scans it in one pass.  The reference's bounded exclusive_scan forwards to
// its inclusive form (scan.hpp:548); this one computes the exclusive scan
// its name and the unbounded form (scan.hpp:553-591) promise.
template <std::size_t B, std::size_t S, std::size_t G, class It1, class It2, class T, class Op>
__device__ typename std::enable_if<(B <= S * G), It2>::type inclusive_scan(
    bounded<B, concurrent_group<agent<G>, S>>& g, It1 first, It1 last, It2 result, T init, Op op) {
  return inclusive_scan(static_cast<concurrent_group<agent<G>, S>&>(g), first, last, result, init,
                        op);
}

template <std::size_t B, std::size_t S, std::size_t G, class It1, class It2, class Op>
__device__ typename std::enable_if<(B <= S * G), It2>::type inclusive_scan(
    bounded<B, concurrent_group<agent<G>, S>>& g, It1 first, It1 last, It2 result, Op op) {
  return inclusive_scan(static_cast<concurrent_group<agent<G>, S>&>(g), first, last, result, op);
}

template <std::size_t B, std::size_t S, std::size_t G, class It1, class It2, class T, class Op>
__device__ typename std::enable_if<(B <= S * G), It2>::type exclusive_scan(
    bounded<B, concurrent_group<agent<G>, S>>& g, It1 first, It1 last, It2 result, T init, Op op) {
  return exclusive_scan(static_cast<concurrent_group<agent<G>, S>&>(g), first, last, result, init,
                        op);
}

// merge_path(a, n1, b, n2, diag, comp): how many of the first diag outputs of
// the stable merge come from a (ties take a first), bulk/algorithm/merge.hpp:35-60.
template <class It1, class Size, class It2, class Compare>
__host__ __device__ Size merge_path(It1 a, Size n1, It2 b, Size n2, Size diag, Compare comp) {
  Size lo = diag > n2 ? diag - n2 : Size(0);
  Size hi = diag < n1 ? diag : n1;
  while (lo < hi) {
    const Size m = lo + (hi - lo) / 2;
    if (comp(b[diag - 1 - m], a[m])) hi = m;
    else lo = m + 1;
  }
  return lo;
}

// Group merge: each agent finds its diagonal and merges a grain sequentially.
template <class Agent, std::size_t S, class It1, class It2, class It3, class Compare>
__device__ It3 merge(concurrent_group<Agent, S>& g, It1 first1, It1 last1, It2 first2, It2 last2,
                     It3 result, Compare comp) {
  typedef long long size_type;
  const size_type n1 = last1 - first1, n2 = last2 - first2, n = n1 + n2;
  constexpr int G = Agent::static_grainsize;
  for (size_type d = size_type(g.this_exec.index()) * G; d < n; d += size_type(g.size()) * G) {
    size_type i = merge_path(first1, n1, first2, n2, d, comp);
    size_type j = d - i;
    for (int k = 0; k < G && d + k < n; ++k) {
      const bool take_a = j >= n2 || (i < n1 && !comp(first2[j], first1[i]));
      result[d + k] = take_a ? first1[i++] : first2[j++];
    }
  }
  g.wait();
  return result + n;
}

// ------------------------------------------------------ bounded-agent bodies
// The sequential bodies of one work-item whose work is statically bounded
// (bulk::bound<N>(self)); each touches at most N elements.

// inclusive_scan: init = op(init, x[i]); result[i] = init (scan.hpp:35-56).
template <std::size_t B, std::size_t G, class It1, class It2, class T, class Op>
__device__ It2 inclusive_scan(const bounded<B, agent<G>>& exec, It1 first, It1 last, It2 result,
                              T init, Op op) {
  const auto n = last - first;
  for (int i = 0; i < exec.bound(); ++i) {
    if (i < n) {
      init = op(init, first[i]);
      result[i] = init;
    }
  }
  return result + n;
}

// exclusive_scan: result[i] = init; init = op(init, x[i]) (scan.hpp:59-78).
template <std::size_t B, std::size_t G, class It1, class It2, class T, class Op>
__device__ It2 exclusive_scan(const bounded<B, agent<G>>& exec, It1 first, It1 last, It2 result,
                              T init, Op op) {
  const auto n = last - first;
  for (int i = 0; i < exec.bound(); ++i) {
    if (i < n) {
      result[i] = init;
      init = op(init, first[i]);
    }
  }
  return result + n;
}

// reduce / accumulate: left fold from init (reduce.hpp:33-58, accumulate.hpp:30-55).
template <std::size_t B, std::size_t G, class It, class T, class Op>
__device__ T accumulate(const bounded<B, agent<G>>& exec, It first, It last, T init, Op op) {
  const auto n = last - first;
  for (int i = 0; i < exec.bound(); ++i)
    if (i < n) init = op(init, first[i]);
  return init;
}

template <std::size_t B, std::size_t G, class It, class T, class Op>
__device__ T reduce(const bounded<B, agent<G>>& exec, It first, It last, T init, Op op) {
  return bulk::accumulate(exec, first, last, init, op);
}

// copy_n of at most the bound's elements.
template <std::size_t B, std::size_t G, class It1, class Size, class It2>
__device__ It2 copy_n(const bounded<B, agent<G>>& exec, It1 first, Size n, It2 result) {
  for (int i = 0; i < exec.bound(); ++i)
    if (Size(i) < n) result[i] = first[i];
  return result + n;
}

// merge: the first min(bound, n1 + n2) outputs of the stable merge; a is
// taken unless b's head is strictly less (merge.hpp:63-170).
template <std::size_t B, std::size_t G, class It1, class It2, class It3, class Compare>
__device__ It3 merge(const bounded<B, agent<G>>& exec, It1 first1, It1 last1, It2 first2,
                     It2 last2, It3 result, Compare comp) {
  const auto n1 = last1 - first1;
  const auto n2 = last2 - first2;
  decltype(n1 + n2) i = 0, j = 0, k = 0;
  for (; k < exec.bound() && i + j < n1 + n2; ++k) {
    const bool take_a = j >= n2 || (i < n1 && !comp(first2[j], first1[i]));
    result[k] = take_a ? first1[i++] : first2[j++];
  }
  return result + k;
}

// merge_by_key: as merge, each key's value travelling with it (merge.hpp:173-305).
template <std::size_t B, std::size_t G, class It1, class It2, class It3, class It4, class It5,
          class It6, class Compare>
__device__ thrust::pair<It5, It6> merge_by_key(const bounded<B, agent<G>>& exec, It1 keys_first1,
                                               It1 keys_last1, It2 keys_first2, It2 keys_last2,
                                               It3 values_first1, It4 values_first2,
                                               It5 keys_result, It6 values_result, Compare comp) {
  const auto n1 = keys_last1 - keys_first1;
  const auto n2 = keys_last2 - keys_first2;
  decltype(n1 + n2) i = 0, j = 0, k = 0;
  for (; k < exec.bound() && i + j < n1 + n2; ++k) {
    const bool take_a = j >= n2 || (i < n1 && !comp(keys_first2[j], keys_first1[i]));
    if (take_a) {
      keys_result[k] = keys_first1[i];
      values_result[k] = values_first1[i++];
    } else {
      keys_result[k] = keys_first2[j];
      values_result[k] = values_first2[j++];
    }
  }
  return thrust::make_pair(keys_result + k, values_result + k);
}

// stable_sort_by_key / stable_sort: odd-even transposition over at most the
// bound's elements; a pair swaps only when the later key is strictly less,
// so equal keys keep their order (sort.hpp:33-149).
template <std::size_t B, std::size_t G, class It1, class It2, class Compare>
__device__ void stable_sort_by_key(const bounded<B, agent<G>>& exec, It1 keys_first, It1 keys_last,
                                   It2 values_first, Compare comp) {
  const int n = int(keys_last - keys_first);
  for (int r = 0; r < exec.bound() && r < n; ++r) {
    for (int j = r & 1; j + 1 < n; j += 2) {
      if (comp(keys_first[j + 1], keys_first[j])) {
        auto k = keys_first[j];
        keys_first[j] = keys_first[j + 1];
        keys_first[j + 1] = k;
        auto v = values_first[j];
        values_first[j] = values_first[j + 1];
        values_first[j + 1] = v;
      }
    }
  }
}

template <std::size_t B, std::size_t G, class It, class Compare>
__device__ void stable_sort(const bounded<B, agent<G>>& exec, It first, It last, Compare comp) {
  const int n = int(last - first);
  for (int r = 0; r < exec.bound() && r < n; ++r) {
    for (int j = r & 1; j + 1 < n; j += 2) {
      if (comp(first[j + 1], first[j])) {
        auto k = first[j];
        first[j] = first[j + 1];
        first[j + 1] = k;
      }
    }
  }
}

// scatter_if: result[map[i]] = first[i] where stencil[i] (scatter.hpp:27-33).
template <std::size_t B, std::size_t G, class It1, class It2, class It3, class It4>
__device__ void scatter_if(const bounded<B, agent<G>>& exec, It1 first, It1 last, It2 map,
                           It3 stencil, It4 result) {
  const auto n = last - first;
  for (int i = 0; i < exec.bound(); ++i)
    if (i < n && stencil[i]) result[map[i]] = first[i];
}

// gather: result[i] = input[map[i]] (gather.hpp:31-64).
template <std::size_t B, std::size_t G, class It1, class It2, class It3>
__device__ It3 gather(const bounded<B, agent<G>>& exec, It1 map_first, It1 map_last,
                      It2 input_first, It3 result) {
  const auto n = map_last - map_first;
  for (int i = 0; i < exec.bound(); ++i)
    if (i < n) result[i] = input_first[map_first[i]];
  return result + n;
}

// --------------------------------------------------------- flag ranges
namespace detail {

// A random-access iterator whose element d is f(i + d): the flag ranges
// below compute each flag from the keys when it is read.
template <class F, class Value, class Index>
class index_iterator {
 public:
  typedef std::random_access_iterator_tag iterator_category;
  typedef Value value_type;
  typedef Index difference_type;
  typedef const Value* pointer;
  typedef Value reference;

  __host__ __device__ index_iterator(F f = F(), Index i = 0) : m_f(f), m_i(i) {}
  __host__ __device__ Value operator*() const { return m_f(m_i); }
  __host__ __device__ Value operator[](Index d) const { return m_f(m_i + d); }
  __host__ __device__ index_iterator operator+(Index d) const { return index_iterator(m_f, m_i + d); }
  __host__ __device__ index_iterator operator-(Index d) const { return index_iterator(m_f, m_i - d); }
  __host__ __device__ Index operator-(const index_iterator& o) const { return m_i - o.m_i; }
  __host__ __device__ index_iterator& operator+=(Index d) { m_i += d; return *this; }
  __host__ __device__ index_iterator& operator-=(Index d) { m_i -= d; return *this; }
  __host__ __device__ index_iterator& operator++() { ++m_i; return *this; }
  __host__ __device__ index_iterator& operator--() { --m_i; return *this; }
  __host__ __device__ bool operator==(const index_iterator& o) const { return m_i == o.m_i; }
  __host__ __device__ bool operator!=(const index_iterator& o) const { return m_i != o.m_i; }
  __host__ __device__ bool operator<(const index_iterator& o) const { return m_i < o.m_i; }

 private:
  F m_f;
  Index m_i;
};

struct equal_to_ {
  template <class A, class B>
  __host__ __device__ bool operator()(const A& a, const B& b) const { return a == b; }
};

}  // namespace detail

// head_flags: flag i = (i == 0 || !pred(x[i], x[i-1])) (head_flags.hpp:118-196).
template <class It, class Pred = detail::equal_to_, class ValueType = bool,
          class IndexType = typename std::iterator_traits<It>::difference_type>
class head_flags {
  struct flag {
    It first;
    Pred pred;
    __host__ __device__ ValueType operator()(IndexType i) const {
      return ValueType(i == 0 || !pred(first[i], first[i - 1]));
    }
  };

 public:
  typedef detail::index_iterator<flag, ValueType, IndexType> iterator;
  __host__ __device__ head_flags(It first, It last, Pred pred = Pred())
      : m_begin(flag{first, pred}, 0), m_end(flag{first, pred}, IndexType(last - first)) {}
  __host__ __device__ iterator begin() const { return m_begin; }
  __host__ __device__ iterator end() const { return m_end; }
  __host__ __device__ ValueType operator[](IndexType i) const { return m_begin[i]; }

 private:
  iterator m_begin, m_end;
};

// head_flags_with_init: flag 0 compares with init instead of being set:
// flag 0 = !pred(init, x[0]) (head_flags.hpp:26-110, bulk/detail/head_flags.hpp:34-122).
template <class It, class Pred = detail::equal_to_, class ValueType = bool,
          class IndexType = typename std::iterator_traits<It>::difference_type>
class head_flags_with_init {
  typedef typename std::iterator_traits<It>::value_type init_type;
  struct flag {
    It first;
    init_type init;
    Pred pred;
    __host__ __device__ ValueType operator()(IndexType i) const {
      return ValueType(i == 0 ? !pred(init, first[0]) : !pred(first[i], first[i - 1]));
    }
  };

 public:
  typedef detail::index_iterator<flag, ValueType, IndexType> iterator;
  __host__ __device__ head_flags_with_init(It first, It last, init_type init, Pred pred = Pred())
      : m_begin(flag{first, init, pred}, 0),
        m_end(flag{first, init, pred}, IndexType(last - first)) {}
  __host__ __device__ iterator begin() const { return m_begin; }
  __host__ __device__ iterator end() const { return m_end; }
  __host__ __device__ ValueType operator[](IndexType i) const { return m_begin[i]; }

 private:
  iterator m_begin, m_end;
};

// tail_flags: flag i = (i == n-1 || !pred(x[i], x[i+1])) (tail_flags.hpp:26-109).
template <class It, class Pred = detail::equal_to_, class ValueType = bool,
          class IndexType = typename std::iterator_traits<It>::difference_type>
class tail_flags {
  struct flag {
    It first;
    IndexType n;
    Pred pred;
    __host__ __device__ ValueType operator()(IndexType i) const {
      return ValueType(i == n - 1 || !pred(first[i], first[i + 1]));
    }
  };

 public:
  typedef detail::index_iterator<flag, ValueType, IndexType> iterator;
  __host__ __device__ tail_flags(It first, It last, Pred pred = Pred())
      : m_begin(flag{first, IndexType(last - first), pred}, 0),
        m_end(flag{first, IndexType(last - first), pred}, IndexType(last - first)) {}
  __host__ __device__ iterator begin() const { return m_begin; }
  __host__ __device__ iterator end() const { return m_end; }
  __host__ __device__ ValueType operator[](IndexType i) const { return m_begin[i]; }

 private:
  iterator m_begin, m_end;
};

template <class It>
__host__ __device__ head_flags<It> make_head_flags(It first, It last) {
  return head_flags<It>(first, last);
}
template <class It, class Pred>
__host__ __device__ head_flags<It, Pred> make_head_flags(It first, It last, Pred pred) {
  return head_flags<It, Pred>(first, last, pred);
}
template <class It>
__host__ __device__ tail_flags<It> make_tail_flags(It first, It last) {
  return tail_flags<It>(first, last);
}
template <class It, class Pred>
__host__ __device__ tail_flags<It, Pred> make_tail_flags(It first, It last, Pred pred) {
  return tail_flags<It, Pred>(first, last, pred);
}

// ------------------------------------------------- more group algorithms

// accumulate: init op x[0] op ... op x[n-1] in order, so the op need only be
// associative (accumulate.hpp:198-217; the reference reduces when the op is
// commutative, which gives the same value for exact arithmetic).  Each agent
// folds a blocked grain of each tile; the grains combine in agent order.
// Every agent receives the result.
template <class Agent, std::size_t S, class It, class T, class Op>
__device__ T accumulate(concurrent_group<Agent, S>& g, It first, It last, T init, Op op) {
  const auto n = last - first;
  constexpr int G = Agent::static_grainsize;
  const int tid = g.this_exec.index();
  const auto tile = decltype(n + 0)(g.size()) * G;
  struct P {
    T v;
    int has;
  };
  auto pop = [op](P a, P b) { return b.has ? (a.has ? P{op(a.v, b.v), 1} : b) : a; };
  for (decltype(n + 0) base = 0; base < n; base += tile) {
    P p{init, 0};
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const auto i = base + decltype(n + 0)(tid) * G + j;
      if (i < n) p = pop(p, P{T(first[i]), 1});
    }
    P total;
    (void)detail::group_inclusive(g, p, pop, &total);
    if (total.has) init = op(init, total.v);
  }
  return init;
}

// merge_by_key: the stable merge of two key ranges, each value travelling
// with its key (merge.hpp:507-607); each agent finds the merge path of its
// grain and merges it sequentially.  Outputs must not overlap the inputs.
template <class Agent, std::size_t S, class It1, class It2, class It3, class It4, class It5,
          class It6, class Compare>
__device__ thrust::pair<It5, It6> merge_by_key(concurrent_group<Agent, S>& g, It1 keys_first1,
                                               It1 keys_last1, It2 keys_first2, It2 keys_last2,
                                               It3 values_first1, It4 values_first2,
                                               It5 keys_result, It6 values_result, Compare comp) {
  typedef long long size_type;
  const size_type n1 = keys_last1 - keys_first1, n2 = keys_last2 - keys_first2, n = n1 + n2;
  constexpr int G = Agent::static_grainsize;
  for (size_type d = size_type(g.this_exec.index()) * G; d < n; d += size_type(g.size()) * G) {
    size_type i = merge_path(keys_first1, n1, keys_first2, n2, d, comp);
    size_type j = d - i;
    for (int k = 0; k < G && d + k < n; ++k) {
      const bool take_a = j >= n2 || (i < n1 && !comp(keys_first2[j], keys_first1[i]));
      if (take_a) {
        keys_result[d + k] = keys_first1[i];
        values_result[d + k] = values_first1[i++];
      } else {
        keys_result[d + k] = keys_first2[j];
        values_result[d + k] = values_first2[j++];
      }
    }
  }
  g.wait();
  return thrust::make_pair(keys_result + n, values_result + n);
}

namespace detail {

// One merge step of a bounded group: agent t produces outputs [t*G, t*G+G)
// of the stable merge of [a, a+n1) and [a+n1, a+n1+n2) (keys, and values when
// VALUES), holds them in registers until every agent has read its inputs,
// then writes them back in place.
template <bool VALUES, int G, class G_, class It1, class It2, class Compare>
__device__ void inplace_merge_step(G_& g, It1 keys, It2 vals, long long pair_start, long long n1,
                                   long long n2, long long p, long long count, Compare comp) {
  typedef typename std::iterator_traits<It1>::value_type K;
  typedef typename std::conditional<VALUES, typename std::iterator_traits<It2>::value_type,
                                    int>::type V;
  K k_out[G];
  V v_out[G];
  It1 a = keys + pair_start, b = a + n1;
  long long i = 0, j = 0;
  if (count > 0) {
    const long long d = p - pair_start;
    i = merge_path(a, n1, b, n2, d, comp);
    j = d - i;
#pragma unroll
    for (int k = 0; k < G; ++k) {
      if (k < count) {
        const bool take_a = j >= n2 || (i < n1 && !comp(b[j], a[i]));
        const long long src = take_a ? pair_start + i : pair_start + n1 + j;
        k_out[k] = keys[src];
        if constexpr (VALUES) v_out[k] = vals[src];
        if (take_a) ++i;
        else ++j;
      }
    }
  }
  g.wait();
#pragma unroll
  for (int k = 0; k < G; ++k) {
    if (k < count) {
      keys[p + k] = k_out[k];
      if constexpr (VALUES) vals[p + k] = v_out[k];
    }
  }
  g.wait();
}

}  // namespace detail

// inplace_merge on a bounded group: [first, middle) and [middle, last), each
// sorted, become one sorted range; n <= bound <= groupsize * grainsize
// (merge.hpp:308-353).
template <std::size_t B, std::size_t S, std::size_t G, class It, class Compare>
__device__ typename std::enable_if<(B <= S * G)>::type inplace_merge(
    bounded<B, concurrent_group<agent<G>, S>>& g, It first, It middle, It last, Compare comp) {
  const long long n1 = middle - first, n2 = last - middle, n = n1 + n2;
  const long long p = (long long)(g.this_exec.index()) * G;
  const long long count = p < n ? (n - p < (long long)G ? n - p : (long long)G) : 0;
  detail::inplace_merge_step<false, int(G)>(g, first, (int*)nullptr, 0, n1, n2, p, count, comp);
}

// stable_sort_by_key on a bounded group: each agent sorts its grain by odd-even
// transposition, then rounds of in-place merges of adjacent sorted runs
// double the run length until one run remains (sort.hpp:152-166,
// stable_merge_sort.hpp:125-206).  n <= bound <= groupsize * grainsize.
template <std::size_t B, std::size_t S, std::size_t G, class It1, class It2, class Compare>
__device__ typename std::enable_if<(B <= S * G)>::type stable_sort_by_key(
    bounded<B, concurrent_group<agent<G>, S>>& g, It1 keys_first, It1 keys_last, It2 values_first,
    Compare comp) {
  typedef typename std::iterator_traits<It1>::value_type K;
  typedef typename std::iterator_traits<It2>::value_type V;
  const long long n = keys_last - keys_first;
  const long long p = (long long)(g.this_exec.index()) * G;
  const int cnt = p < n ? int(n - p < (long long)G ? n - p : (long long)G) : 0;
  {
    K k[G];
    V v[G];
#pragma unroll
    for (int j = 0; j < int(G); ++j) {
      if (j < cnt) {
        k[j] = keys_first[p + j];
        v[j] = values_first[p + j];
      }
    }
#pragma unroll
    for (int r = 0; r < int(G); ++r) {
#pragma unroll
      for (int j = r & 1; j + 1 < int(G); j += 2) {
        if (j + 1 < cnt && comp(k[j + 1], k[j])) {
          K tk = k[j];
          k[j] = k[j + 1];
          k[j + 1] = tk;
          V tv = v[j];
          v[j] = v[j + 1];
          v[j + 1] = tv;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < int(G); ++j) {
      if (j < cnt) {
        keys_first[p + j] = k[j];
        values_first[p + j] = v[j];
      }
    }
  }
  g.wait();
  for (long long w = G; w < n; w *= 2) {
    // 2w is a multiple of the grain, so an agent's outputs lie in one pair of runs
    const long long pair_start = p < n ? (p / (2 * w)) * (2 * w) : 0;
    const long long rest = n - pair_start;
    const long long n1 = rest < w ? rest : w;
    const long long n2 = rest - n1 < w ? rest - n1 : w;
    detail::inplace_merge_step<true, int(G)>(g, keys_first, values_first, pair_start, n1, n2, p,
                                             cnt, comp);
  }
}

// reduce_by_key on a group (reduce_by_key.hpp:115-216): runs of keys equal
// under pred are reduced with op, continuing the run (init_key, init_value)
// carried in.  Per tile of groupsize * grainsize elements, element i is a
// head when !pred(x[i], x[i-1]) (element 0 compares with the carried key as
// !pred(init_key, x[0])); every run that ends inside the tile writes its last
// key and its sum to the next output slot, and the carried run is written
// first when the tile's first element starts a new run.  The tile's last run
// is carried to the next tile; the returned tuple holds the output ends and
// the final carry (key, value), which is not written.  Outputs may alias the
// inputs' start (each output slot is at or before the element that fills it).
template <std::size_t S, std::size_t G, class It1, class It2, class It3, class It4, class T1,
          class T2, class Pred, class Op>
__device__ thrust::tuple<It3, It4, typename std::iterator_traits<It1>::value_type,
                         typename std::iterator_traits<It4>::value_type>
reduce_by_key(concurrent_group<agent<G>, S>& g, It1 keys_first, It1 keys_last, It2 values_first,
              It3 keys_result, It4 values_result, T1 init_key, T2 init_value, Pred pred, Op op) {
  typedef typename std::iterator_traits<It1>::value_type K;
  typedef typename std::iterator_traits<It2>::value_type V;
  typedef typename std::iterator_traits<It4>::value_type VOut;
  // (heads so far, fold of the values since the last head); `has` marks a
  // non-empty part; the combine is the scan_head_flags operator
  // (reduce_by_key.hpp:41-62)
  struct P {
    long long cnt;
    V v;
    int has;
  };
  auto pop = [op](P a, P b) {
    if (!b.has) return a;
    if (!a.has) return b;
    return P{a.cnt + b.cnt, b.cnt ? b.v : V(op(a.v, b.v)), 1};
  };
  const int tid = g.this_exec.index();
  const long long tile = (long long)(g.size()) * G;
  K carry_key = K(init_key);
  V carry_val = V(init_value);
  for (; keys_first < keys_last; keys_first += tile, values_first += tile) {
    const long long rem = keys_last - keys_first;
    const long long n = rem < tile ? rem : tile;
    const long long base = (long long)(tid) * G;
    K k[G];
    V v[G];
    bool head[G];
    bool tail[G];
    P mine{0, V(), 0};
#pragma unroll
    for (int j = 0; j < int(G); ++j) {
      const long long i = base + j;
      head[j] = false;
      tail[j] = false;
      if (i < n) {
        k[j] = keys_first[i];
        v[j] = values_first[i];
        if (i == 0) head[j] = !pred(carry_key, k[j]);
        else head[j] = !pred(k[j], j > 0 ? k[j > 0 ? j - 1 : 0] : K(keys_first[i - 1]));
        mine = pop(mine, P{head[j] ? 1 : 0, v[j], 1});
      }
    }
    // element i ends a run written out when i < n - 1 and i + 1 is a head
#pragma unroll
    for (int j = 0; j < int(G); ++j) {
      const long long i = base + j;
      if (i + 1 < n)
        tail[j] = j + 1 < int(G) ? head[j + 1 < int(G) ? j + 1 : j]
                                 : !pred(K(keys_first[i + 1]), k[j]);
    }
    const K last_key = keys_first[n - 1];
    P total;
    const P inc = detail::group_inclusive(g, mine, pop, &total);  // waits: all loads are done
    // exclusive prefix of this agent, seeded with the carried run (1 slot, carry_val)
    P excl = detail::shfl_up(inc, 1);
    __shared__ P s_last[16];
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 63) s_last[wid] = inc;
    g.wait();
    if (lane == 0) excl = wid > 0 ? s_last[wid - 1] : P{0, V(), 0};
    P r = pop(P{1, carry_val, 1}, excl);
#pragma unroll
    for (int j = 0; j < int(G); ++j) {
      const long long i = base + j;
      if (i < n) {
        r = pop(r, P{head[j] ? 1 : 0, v[j], 1});
        if (tail[j]) {
          keys_result[r.cnt - 1] = k[j];
          values_result[r.cnt - 1] = VOut(r.v);
        }
      }
    }
    if (tid == 0 && head[0]) {
      keys_result[0] = carry_key;
      values_result[0] = VOut(carry_val);
    }
    const P fin = pop(P{1, carry_val, 1}, total);
    keys_result += fin.cnt - 1;
    values_result += fin.cnt - 1;
    carry_key = last_key;
    carry_val = fin.v;
    g.wait();  // s_last is rewritten by the next tile
  }
  return thrust::make_tuple(keys_result, values_result, carry_key, VOut(carry_val));
}

// scatter_if: result[map[i]] = first[i] where stencil[i], striped over the
// group (scatter.hpp:96-178).
template <class Agent, std::size_t S, class It1, class It2, class It3, class It4>
__device__ void scatter_if(concurrent_group<Agent, S>& g, It1 first, It1 last, It2 map,
                           It3 stencil, It4 result) {
  const auto n = last - first;
  for (decltype(n + 0) i = g.this_exec.index(); i < n; i += g.size())
    if (stencil[i]) result[map[i]] = first[i];
  g.wait();
}

// gather: result[i] = input[map[i]], striped over the group (gather.hpp:67-78).
template <class Agent, std::size_t S, class It1, class It2, class It3>
__device__ It3 gather(concurrent_group<Agent, S>& g, It1 map_first, It1 map_last,
                      It2 input_first, It3 result) {
  const auto n = map_last - map_first;
  for (decltype(n + 0) i = g.this_exec.index(); i < n; i += g.size())
    result[i] = input_first[map_first[i]];
  g.wait();
  return result + n;
}

// adjacent_difference with init: result[i] = op(x[i], x[i-1]), x[-1] = init
// (adjacent_difference.hpp:27-47, 50-107).  Each agent reads its grain and
// the element before it, the group waits, then writes, so result may equal
// first; the last element of each tile is carried to the next.  (The
// reference's group loop hands every agent the tile's last element as the
// predecessor of its next grain, adjacent_difference.hpp:80, which is right
// for agent 0 only; this form gives every element its own predecessor.)
template <class Agent, std::size_t S, class It1, class It2, class T, class Op>
__device__ It2 adjacent_difference(concurrent_group<Agent, S>& g, It1 first, It1 last, It2 result,
                                   T init, Op op) {
  const long long n = last - first;
  constexpr int G = Agent::static_grainsize;
  const long long tile = (long long)(g.size()) * G;
  const int tid = g.this_exec.index();
  T carry = init;  // element before the tile
  for (long long base = 0; base < n; base += tile) {
    const long long lo = base + (long long)(tid)*G;
    const long long rest = n - lo;
    const int cnt = rest <= 0 ? 0 : (rest < G ? int(rest) : G);
    T x[G];
    T before = carry;
    if (cnt > 0 && tid > 0) before = T(first[lo - 1]);
#pragma unroll
    for (int j = 0; j < G; ++j)
      if (j < cnt) x[j] = T(first[lo + j]);
    const long long tend = base + tile < n ? base + tile : n;
    const T next_carry = T(first[tend - 1]);
    g.wait();
#pragma unroll
    for (int j = 0; j < G; ++j)
      if (j < cnt) result[lo + j] = op(x[j], j > 0 ? x[j > 0 ? j - 1 : 0] : before);
    carry = next_carry;
    g.wait();
  }
  return result + n;
}

// adjacent_difference without init: result[0] = x[0], then as above
// (adjacent_difference.hpp:110-137).
template <class Agent, std::size_t S, class It1, class It2, class Op>
__device__ It2 adjacent_difference(concurrent_group<Agent, S>& g, It1 first, It1 last, It2 result,
                                   Op op) {
  if (!(first < last)) return result;
  typedef typename std::iterator_traits<It1>::value_type T;
  const T init = first[0];
  g.wait();  // first may equal result
  if (g.this_exec.index() == 0) result[0] = init;
  return bulk::adjacent_difference(g, first + 1, last, result + 1, init, op);
}

// adjacent_difference of one agent, sequential (adjacent_difference.hpp:27-47).
template <std::size_t G, class It1, class It2, class T, class Op>
__device__ It2 adjacent_difference(agent<G>& exec, It1 first, It1 last, It2 result, T init,
                                   Op op) {
  for (; first != last; ++first, ++result) {
    const T x = *first;
    *result = op(x, init);
    init = x;
  }
  return result;
}

// for_each_n: f(x[i]) for i < n, striped over the group, then a wait
// (for_each.hpp:28-45); on a bounded agent, sequential (for_each.hpp:48-70).
template <class Agent, std::size_t S, class It, class Size, class F>
__device__ It for_each_n(concurrent_group<Agent, S>& g, It first, Size n, F f) {
  for (Size i = Size(g.this_exec.index()); i < n; i += Size(g.size())) f(first[i]);
  g.wait();
  return first + n;
}

template <std::size_t B, std::size_t G, class It, class Size, class F>
__device__ It for_each_n(bounded<B, agent<G>>& exec, It first, Size n, F f) {
  for (int i = 0; i < exec.bound(); ++i)
    if (Size(i) < n) f(first[i]);
  return first + n;
}

// join_iterator: two ranges viewed as one, element i being first1[i] for
// i < n1 and first2[i - n1] after (join_iterator.hpp:47-100); read-only.
namespace detail {
template <class It1, class It2, class Difference>
struct join_at {
  It1 first1;
  Difference n1;
  It2 first2;
  __host__ __device__ typename std::iterator_traits<It1>::value_type operator()(Difference i) const {
    return i < n1 ? first1[i] : typename std::iterator_traits<It1>::value_type(first2[i - n1]);
  }
};
}  // namespace detail

template <class It1, class It2,
          class Difference = typename std::iterator_traits<It1>::difference_type>
using join_iterator = detail::index_iterator<detail::join_at<It1, It2, Difference>,
                                             typename std::iterator_traits<It1>::value_type,
                                             Difference>;

template <class It1, class Size, class It2>
__host__ __device__ join_iterator<It1, It2, Size> make_join_iterator(It1 first1, Size n1, It2 first2) {
  return join_iterator<It1, It2, Size>(detail::join_at<It1, It2, Size>{first1, n1, first2}, Size(0));
}

}  // namespace bulk
