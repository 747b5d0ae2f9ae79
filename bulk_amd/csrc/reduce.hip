// Device-wide reduce and reduce_intervals.
//
// reduce replaces my_reduce (reduce.cu:58-97): the reference runs a
// partial-sum launch over an aligned_decomposition of 10 groups per SM
// (reduce_partitions, reduce.cu:13-55, built on bulk::reduce,
// reduce.hpp:110-219), a second one-group launch over the partials, and a
// blocking copy of the scalar back to the host.  Here the first launch is a
// grid-stride stream of dwordx4 loads with several accumulators per thread,
// a DPP wavefront reduction and an LDS combine; the second launch folds the
// per-workgroup partials and the init and writes the result to device
// memory, so the call never blocks the host.
//
// reduce_intervals replaces reduce_intervals (reduce_intervals.hpp:31-40):
// one workgroup per partition of an aligned_decomposition
// (decomposition.hpp:142-218).
#include "device_common.hpp"
#include "host_common.hpp"
#include "tuning.hpp"

namespace bulkx {

// 256 lanes x 2 vectors in flight, 3 workgroups per CU: a sweep of
// (lanes, vectors, workgroups per CU) on MI355X at 2^28 float32 found that
// fewer bytes in flight per CU (about 24 KB) stream faster from HBM
// (7.2 TB/s) than deeper queues (8 per CU x 4 vectors: 5.5 TB/s).
constexpr int kReduceBlock = 256;
constexpr int kReduceUnroll = 2;
constexpr int kReduceBlocksPerCU = 3;

template <class T, class Op, int BLOCK, int UNROLL>
__global__ __launch_bounds__(BLOCK) void reduce_partials_kernel(
    const T* __restrict__ in, int64_t n, typename Acc<T, Op>::type* __restrict__ partials) {
  using AccP = Acc<T, Op>;
  using A = typename AccP::type;
  using AOp = typename AccP::op;
  constexpr int V = 16 / sizeof(T);
  __shared__ A s_wave[BLOCK / kWave];

  // Elements before the first 16-byte boundary and after the last full vector.
  const uintptr_t addr = reinterpret_cast<uintptr_t>(in);
  int64_t head = int64_t(((16 - (addr & 15)) & 15) / sizeof(T));
  if (head > n) head = n;
  const int64_t nvec = (n - head) / V;
  const int64_t tail = head + nvec * V;
  const u32x4* vin = reinterpret_cast<const u32x4*>(in + head);

  A acc[UNROLL];
#pragma unroll
  for (int u = 0; u < UNROLL; ++u) acc[u] = AOp::template identity<A>();

  const int64_t stride = int64_t(gridDim.x) * BLOCK;
  int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  for (; i + (UNROLL - 1) * stride < nvec; i += UNROLL * stride) {
    u32x4 t[UNROLL];
    // Nontemporal loads: the stream is read once, so it should not displace
    // anything in L2 / MALL (read-only streams reach ~87% of HBM peak this way).
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) t[u] = __builtin_nontemporal_load(vin + i + u * stride);
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const T* e = reinterpret_cast<const T*>(&t[u]);
      // Fold the vector's V elements pairwise before they meet the accumulator.
      T s = e[0];
#pragma unroll
      for (int j = 1; j < V; ++j) s = Op::apply(s, e[j]);
      acc[u] = AOp::apply(acc[u], AccP::from(s));
    }
  }
  for (; i < nvec; i += stride) {
    u32x4 t = __builtin_nontemporal_load(vin + i);
    const T* e = reinterpret_cast<const T*>(&t);
    T s = e[0];
#pragma unroll
    for (int j = 1; j < V; ++j) s = Op::apply(s, e[j]);
    acc[0] = AOp::apply(acc[0], AccP::from(s));
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x < head) acc[0] = AOp::apply(acc[0], AccP::from(in[threadIdx.x]));
    const int64_t t = tail + threadIdx.x;
    if (t < n) acc[0] = AOp::apply(acc[0], AccP::from(in[t]));
  }
#pragma unroll
  for (int u = 1; u < UNROLL; ++u) acc[0] = AOp::apply(acc[0], acc[u]);
  A r = block_reduce<BLOCK, AOp>(acc[0], s_wave);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

template <class T, class Op, int BLOCK>
__global__ __launch_bounds__(BLOCK) void reduce_final_kernel(
    const typename Acc<T, Op>::type* __restrict__ partials, int num, T init, bool has_init,
    T* __restrict__ out) {
  using AccP = Acc<T, Op>;
  using A = typename AccP::type;
  using AOp = typename AccP::op;
  __shared__ A s_wave[BLOCK / kWave];
  A acc = AOp::template identity<A>();
  for (int i = threadIdx.x; i < num; i += BLOCK) acc = AOp::apply(acc, partials[i]);
  A r = block_reduce<BLOCK, AOp>(acc, s_wave);
  if (threadIdx.x == 0) {
    if (has_init) r = AOp::apply(AccP::from(init), r);
    *out = AccP::to(r);
  }
}

// Workgroups per CU of each tuning variant: 0 = 3 (default), 1 = 2, 2 = 6.
static int reduce_blocks_per_cu(int variant) {
  return variant == 1 ? 2 : variant == 2 ? 6 : kReduceBlocksPerCU;
}
static constexpr int kReduceMaxBlocksPerCU = 6;

static int reduce_grid(int64_t n, int elem_size, int variant) {
  DeviceProps p;
  int cus = current_device_props(&p) == BULKX_SUCCESS ? p.num_cus : 256;
  const int64_t vec = 16 / elem_size;
  const int64_t per_block = int64_t(kReduceBlock) * vec * kReduceUnroll;
  int64_t want = (n + per_block - 1) / per_block;
  int64_t cap = int64_t(cus) * reduce_blocks_per_cu(variant);
  if (want < 1) want = 1;
  return int(want < cap ? want : cap);
}

size_t reduce_workspace_bytes(int dtype, int64_t n) {
  size_t es = dtype_size(dtype);
  if (es == 0) return 0;
  // partials in the accumulator type (at most 16 bytes) for the largest grid
  DeviceProps p;
  int cus = current_device_props(&p) == BULKX_SUCCESS ? p.num_cus : 256;
  (void)n;
  return size_t(cus) * kReduceMaxBlocksPerCU * 16;
}

template <class T, class Op>
static int reduce_launch(const T* in, int64_t n, const T* init, T* out, char* ws,
                         hipStream_t stream, int variant) {
  using A = typename Acc<T, Op>::type;
  const T init_v = init ? *init : Op::template identity<T>();
  const int grid = n > 0 ? reduce_grid(n, sizeof(T), variant) : 0;
  A* partials = reinterpret_cast<A*>(ws);
  if (grid > 0) {
    reduce_partials_kernel<T, Op, kReduceBlock, kReduceUnroll>
        <<<grid, kReduceBlock, 0, stream>>>(in, n, partials);
    BULKX_LAUNCHED();
  }
  reduce_final_kernel<T, Op, 1024><<<1, 1024, 0, stream>>>(partials, grid, init_v, init != nullptr,
                                                          out);
  BULKX_LAUNCHED();
  return BULKX_SUCCESS;
}

// ---------------------------------------------------------------------------
// reduce_intervals over aligned_decomposition(n, P, tile):
// partition g covers tiles [g*q + min(g, r), ...) with q = tiles / P,
// r = tiles % P (decomposition.hpp:163-203), clamped to n.
// ---------------------------------------------------------------------------
template <class T, class Op, int BLOCK>
__global__ __launch_bounds__(BLOCK) void reduce_intervals_kernel(const T* __restrict__ in,
                                                                 int64_t n, int64_t num_parts,
                                                                 int64_t tile, T* __restrict__ out) {
  using AccP = Acc<T, Op>;
  using A = typename AccP::type;
  using AOp = typename AccP::op;
  __shared__ A s_wave[BLOCK / kWave];
  const int64_t tiles = (n + tile - 1) / tile;
  const int64_t q = tiles / num_parts, r = tiles % num_parts;
  for (int64_t g = blockIdx.x; g < num_parts; g += gridDim.x) {
    int64_t t0 = q * g + (g < r ? g : r);
    int64_t t1 = t0 + q + (g < r ? 1 : 0);
    int64_t first = t0 * tile;
    int64_t last = t1 * tile < n ? t1 * tile : n;
    A acc = AOp::template identity<A>();
    for (int64_t i = first + threadIdx.x; i < last; i += BLOCK) acc = AOp::apply(acc, AccP::from(in[i]));
    A s = block_reduce<BLOCK, AOp>(acc, s_wave);
    if (threadIdx.x == 0) out[g] = AccP::to(s);
  }
}

}  // namespace bulkx

using namespace bulkx;

extern "C" int bulkx_reduce(int dtype, int op, const void* in, int64_t n, const void* init,
                            void* out, void* workspace, size_t workspace_bytes, void* stream_) {
  if (dtype_size(dtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  if (!valid_op(op)) return BULKX_ERR_UNSUPPORTED_OP;
  if (n < 0 || out == nullptr || (n > 0 && in == nullptr)) return BULKX_ERR_INVALID_ARGUMENT;
  if (n > 0 && (reinterpret_cast<uintptr_t>(in) % dtype_size(dtype)) != 0)
    return BULKX_ERR_INVALID_ARGUMENT;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  Scratch scratch;
  BULKX_TRY(scratch.acquire(workspace, workspace_bytes, reduce_workspace_bytes(dtype, n), stream));
  int st = BULKX_SUCCESS;
  const int v = tuned_variant(BULKX_ALGO_REDUCE, dtype, n);
  BULKX_DISPATCH_DTYPE(dtype, T, BULKX_DISPATCH_OP(op, OP, {
    st = reduce_launch<T, OP>(static_cast<const T*>(in), n, static_cast<const T*>(init),
                              static_cast<T*>(out), scratch.ptr(), stream, v);
  }));
  int rs = scratch.release();
  return st != BULKX_SUCCESS ? st : rs;
}

extern "C" int bulkx_reduce_intervals(int dtype, int op, const void* in, int64_t n,
                                      int64_t num_partitions, int64_t tile_size, void* out,
                                      void* stream_) {
  if (dtype_size(dtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  if (!valid_op(op)) return BULKX_ERR_UNSUPPORTED_OP;
  if (n < 0 || num_partitions < 0 || tile_size <= 0) return BULKX_ERR_INVALID_ARGUMENT;
  if (num_partitions == 0) return BULKX_SUCCESS;
  if (out == nullptr || (n > 0 && in == nullptr)) return BULKX_ERR_INVALID_ARGUMENT;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  const int64_t grid = num_partitions < 65536 ? num_partitions : 65536;
  BULKX_DISPATCH_DTYPE(dtype, T, BULKX_DISPATCH_OP(op, OP, {
    reduce_intervals_kernel<T, OP, 256><<<dim3(uint32_t(grid)), 256, 0, stream>>>(
        static_cast<const T*>(in), n, num_partitions, tile_size, static_cast<T*>(out));
  }));
  BULKX_LAUNCHED();
  return BULKX_SUCCESS;
}
