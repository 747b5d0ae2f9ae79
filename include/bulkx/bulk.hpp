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
//   * bulk::malloc(g, bytes)           -> a bump arena in the workgroup's dynamic LDS
//                                         (heap_size bytes, up to 160 KB per CU on gfx950),
//                                         falling back to device malloc when exhausted
// Group algorithms (copy_n, reduce, inclusive_scan, exclusive_scan,
// merge_path, merge) run on a concurrent_group with wavefront shuffles and
// LDS; see bulk/algorithm.hpp of the reference for their contracts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

namespace bulk {

static constexpr std::size_t use_default = ~std::size_t(0);
static constexpr std::size_t dynamic_group_size = 0;
static constexpr int invalid_index = -1;

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
        m_heap_size(size_type(heap_size == use_default ? 0 : heap_size)) {}

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
  return concurrent_group<>(int(size == use_default ? 256 : size),
                            int(heap_size == use_default ? 0 : heap_size));
}

template <class ExecutionAgent>
concurrent_group<ExecutionAgent> con(ExecutionAgent exec, std::size_t size,
                                     std::size_t heap_size = use_default) {
  return concurrent_group<ExecutionAgent>(int(size == use_default ? 256 : size),
                                          int(heap_size == use_default ? 0 : heap_size), exec);
}

template <std::size_t groupsize, std::size_t grainsize>
concurrent_group<agent<grainsize>, groupsize> con(std::size_t heap_size) {
  return concurrent_group<agent<grainsize>, groupsize>(heap_size);
}

inline parallel_group<concurrent_group<>> grid(std::size_t num_groups = use_default,
                                               std::size_t group_size = use_default,
                                               std::size_t heap_size = use_default) {
  if (num_groups == use_default) num_groups = std::size_t(concurrent_group<>::hardware_concurrency());
  return par(con(group_size, heap_size), num_groups);
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
// Collective allocation in the group's LDS arena: every agent of the group
// calls it with the same size and receives the same pointer.
// The arena is a stack: its first 16 bytes hold the offset of the free
// space; free(p) of an arena block releases p and every block allocated
// after it (the reference's kernels free in reverse order of allocation).
template <class Agent, std::size_t S>
__device__ void* malloc(concurrent_group<Agent, S>& g, std::size_t bytes) {
  __shared__ void* s_ptr;
  g.wait();
  if (g.this_exec.index() == 0) {
    unsigned* top = reinterpret_cast<unsigned*>(g.arena.base);
    const std::size_t need = (bytes + 15) & ~std::size_t(15);
    if (top && std::size_t(*top) + need <= g.arena.bytes) {
      s_ptr = g.arena.base + *top;
      *top += unsigned(need);
    } else {
      s_ptr = ::malloc(bytes);  // device heap when the LDS arena is exhausted
    }
  }
  g.wait();
  return s_ptr;
}

template <class Agent, std::size_t S>
__device__ void free(concurrent_group<Agent, S>& g, void* ptr) {
  g.wait();
  if (g.this_exec.index() == 0 && ptr) {
    char* p = static_cast<char*>(ptr);
    if (g.arena.base && p >= g.arena.base + 16 && p < g.arena.base + g.arena.bytes)
      *reinterpret_cast<unsigned*>(g.arena.base) = unsigned(p - g.arena.base);
    else
      ::free(ptr);
  }
  g.wait();
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
// con(n): one workgroup
template <class A, std::size_t S>
struct launch_traits<concurrent_group<A, S>> {
  static dim3 grid(const concurrent_group<A, S>&) { return dim3(1); }
  static dim3 block(const concurrent_group<A, S>& g) { return dim3(unsigned(g.size())); }
  static std::size_t lds(const concurrent_group<A, S>& g) { return arena_bytes(g.heap_size()); }
  static std::size_t arena_bytes(int heap) { return heap > 0 ? std::size_t(heap) + 16 : 0; }
};
// grid: one workgroup per concurrent group
template <class A, std::size_t S, std::size_t PS>
struct launch_traits<parallel_group<concurrent_group<A, S>, PS>> {
  typedef parallel_group<concurrent_group<A, S>, PS> G;
  static dim3 grid(const G& g) { return dim3(unsigned(g.size())); }
  static dim3 block(const G& g) { return dim3(unsigned(g.this_exec.size())); }
  static std::size_t lds(const G& g) {
    return launch_traits<concurrent_group<A, S>>::arena_bytes(g.this_exec.heap_size());
  }
};

template <class A, std::size_t S>
__device__ void bind_arena(concurrent_group<A, S>& cg, std::size_t bytes) {
  extern __shared__ __align__(16) char bulk_dynamic_lds[];
  cg.arena.base = bytes ? bulk_dynamic_lds : nullptr;
  cg.arena.bytes = unsigned(bytes);
  if (bytes && threadIdx.x == 0) *reinterpret_cast<unsigned*>(bulk_dynamic_lds) = 16;
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

template <class A, std::size_t S, class F, class... Args>
void launch(const concurrent_group<A, S>& g, hipStream_t s, F f, Args... args) {
  if (g.size() <= 0 || g.size() > 1024)
    throw system_error(hipErrorInvalidValue, "bulk::async: a concurrent_group holds 1..1024 agents");
  const std::size_t lds = launch_traits<concurrent_group<A, S>>::lds(g);
  allow_lds(con_kernel<A, S, F, Args...>, lds);
  con_kernel<A, S, F, Args...><<<dim3(1), dim3(unsigned(g.size())), lds, s>>>(g, lds, f, args...);
}

template <class A, std::size_t S, std::size_t PS, class F, class... Args>
void launch(const parallel_group<concurrent_group<A, S>, PS>& g, hipStream_t s, F f, Args... args) {
  if (g.size() <= 0) return;
  if (g.this_exec.size() <= 0 || g.this_exec.size() > 1024)
    throw system_error(hipErrorInvalidValue, "bulk::async: a concurrent_group holds 1..1024 agents");
  typedef parallel_group<concurrent_group<A, S>, PS> G;
  const std::size_t lds = launch_traits<G>::lds(g);
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

}  // namespace bulk
