// Device-wide inclusive / exclusive scan: one single-pass kernel with
// decoupled look-back.
//
// Replaces ::inclusive_scan (scan.cu:87-148), which runs three launches
// (upsweep accumulate_tiles scan.cu:57-84, a one-group exclusive scan of the
// carries scan.cu:133-135, and the downsweep inclusive_downsweep
// scan.cu:34-54 built on bulk::inclusive_scan, scan.hpp:455-486) and so moves
// 3n elements.  Here every tile is read once and written once (2n), and the
// carry between tiles travels through the tile-status words.
//
// Per tile: each thread owns ITEMS consecutive elements, scans them
// sequentially (the bounded-agent body, scan.hpp:35-55), the per-thread
// totals are scanned across the wavefront with DPP and across the
// workgroup through LDS, and wave 0 looks back over the predecessors.
#include "device_common.hpp"
#include "host_common.hpp"
#include "scan_config.hpp"

namespace bulkx {

// Preparation: clears the tile counter, the error word and the status words.
__global__ void scan_prepare_kernel(uint32_t* header, uint64_t* words, int64_t num_words) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0) {
    header[0] = 0;
    header[1] = 0;
  }
  for (; i < num_words; i += int64_t(gridDim.x) * blockDim.x) words[i] = 0;
}

template <class T, class Op, int BLOCK, int ITEMS, bool EXCLUSIVE>
__global__ __launch_bounds__(BLOCK) void scan_kernel(const T* in, T* out, int64_t n, T init,
                                                     bool has_init, uint32_t* header,
                                                     uint64_t* status_words, bool single_tile) {
  using AccP = Acc<T, Op>;
  using A = typename AccP::type;
  using AOp = typename AccP::op;
  using SL = Striped<T, ITEMS>;
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int NW = BLOCK / kWave;
  __shared__ A s_wave[NW];
  __shared__ A s_prefix;
  __shared__ uint32_t s_tile;
  const TileStatus<A> st{status_words};

  int64_t tile;
  if (single_tile) {
    tile = 0;
  } else {
    if (threadIdx.x == 0) s_tile = atomicAdd(header, 1u);
    __syncthreads();
    tile = s_tile;
  }
  const int wid = threadIdx.x / kWave;
  const int64_t chunk_base = tile * TILE + int64_t(wid) * SL::N;
  const int64_t rem = n - chunk_base;
  const int count = rem <= 0 ? 0 : (rem >= SL::N ? SL::N : int(rem));
  const T id = Op::template identity<T>();
  const A aid = AOp::template identity<A>();

  T x[ITEMS];
  striped_load<T, ITEMS>(in + (count > 0 ? chunk_base : 0), count, x, id);

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
  if (lane_id() == 0) s_wave[wid] = carry;
  __syncthreads();
  A wave_prefix = aid, tile_agg = aid;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const A t = s_wave[w];
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
      A excl = lookback_exclusive_prefix<AOp>(st, tile, header + 1);
      if (threadIdx.x == 0) {
        st.publish(tile, kStatusPrefix, AOp::apply(excl, tile_agg));
        s_prefix = excl;
      }
    }
    __syncthreads();
    tile_prefix = s_prefix;
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
  if (count > 0) striped_store<T, ITEMS>(out + chunk_base, count, x);
}

template <class T>
static int64_t scan_num_tiles(int64_t n) {
  constexpr int TILE = ScanConfig<T>::BLOCK * ScanConfig<T>::ITEMS;
  return (n + TILE - 1) / TILE;
}

static constexpr size_t kHeaderBytes = 256;

// Status words per tile: one per 32 bits of the accumulator (int32: 1,
// int64: 2, float32 -> double: 2, float64 -> double-double: 4).  The size is
// taken for the widest accumulator of the dtype so one workspace serves every op.
static int status_words_per_tile(int dtype) {
  switch (dtype) {
    case BULKX_INT32: case BULKX_UINT32: return 1;
    case BULKX_INT64: case BULKX_UINT64: case BULKX_FLOAT32: return 2;
    case BULKX_FLOAT64: return 4;
    default: return 0;
  }
}

size_t scan_workspace_bytes(int dtype, int64_t n) {
  size_t es = dtype_size(dtype);
  if (es == 0 || n <= 0) return 0;
  int64_t tiles = es == 4 ? scan_num_tiles<float>(n) : scan_num_tiles<double>(n);
  if (tiles <= 1) return 0;
  return kHeaderBytes + size_t(tiles) * status_words_per_tile(dtype) * sizeof(uint64_t);
}

template <class T, class Op, bool EXCLUSIVE>
static int scan_launch(const T* in, T* out, int64_t n, const T* init, char* ws,
                       hipStream_t stream) {
  using C = ScanConfig<T>;
  const int64_t tiles = scan_num_tiles<T>(n);
  using A = typename Acc<T, Op>::type;
  const T init_v = init ? *init : Op::template identity<T>();
  if (tiles <= 1) {
    scan_kernel<T, Op, C::BLOCK, C::ITEMS, EXCLUSIVE><<<1, C::BLOCK, 0, stream>>>(
        in, out, n, init_v, init != nullptr, nullptr, nullptr, true);
    BULKX_LAUNCHED();
    return BULKX_SUCCESS;
  }
  uint32_t* header = reinterpret_cast<uint32_t*>(ws);
  uint64_t* words = reinterpret_cast<uint64_t*>(ws + kHeaderBytes);
  const int64_t num_words = tiles * TileStatus<A>::W;
  const int prep_blocks = int((num_words + 255) / 256 < 1024 ? (num_words + 255) / 256 : 1024);
  scan_prepare_kernel<<<prep_blocks, 256, 0, stream>>>(header, words, num_words);
  BULKX_LAUNCHED();
  scan_kernel<T, Op, C::BLOCK, C::ITEMS, EXCLUSIVE><<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(
      in, out, n, init_v, init != nullptr, header, words, false);
  BULKX_LAUNCHED();
  return BULKX_SUCCESS;
}

template <bool EXCLUSIVE>
static int scan_entry(int dtype, int op, const void* in, void* out, int64_t n, const void* init,
                      void* workspace, size_t workspace_bytes, void* stream_) {
  if (dtype_size(dtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  if (!valid_op(op)) return BULKX_ERR_UNSUPPORTED_OP;
  if (n < 0) return BULKX_ERR_INVALID_ARGUMENT;
  if (EXCLUSIVE && init == nullptr) return BULKX_ERR_INVALID_ARGUMENT;
  if (n == 0) return BULKX_SUCCESS;
  if (in == nullptr || out == nullptr) return BULKX_ERR_INVALID_ARGUMENT;
  if (n > (int64_t(1) << 40)) return BULKX_ERR_INVALID_ARGUMENT;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  Scratch scratch;
  BULKX_TRY(scratch.acquire(workspace, workspace_bytes, scan_workspace_bytes(dtype, n), stream));
  int st = BULKX_SUCCESS;
  BULKX_DISPATCH_DTYPE(dtype, T, BULKX_DISPATCH_OP(op, OP, {
    st = scan_launch<T, OP, EXCLUSIVE>(static_cast<const T*>(in), static_cast<T*>(out), n,
                                       static_cast<const T*>(init), scratch.ptr(), stream);
  }));
  int rs = scratch.release();
  return st != BULKX_SUCCESS ? st : rs;
}

}  // namespace bulkx

extern "C" int bulkx_inclusive_scan(int dtype, int op, const void* in, void* out, int64_t n,
                                    const void* init, void* workspace, size_t workspace_bytes,
                                    void* stream) {
  return bulkx::scan_entry<false>(dtype, op, in, out, n, init, workspace, workspace_bytes, stream);
}

extern "C" int bulkx_exclusive_scan(int dtype, int op, const void* in, void* out, int64_t n,
                                    const void* init, void* workspace, size_t workspace_bytes,
                                    void* stream) {
  return bulkx::scan_entry<true>(dtype, op, in, out, n, init, workspace, workspace_bytes, stream);
}
