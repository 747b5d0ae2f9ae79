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
#include <mutex>

#include "device_common.hpp"
#include "host_common.hpp"
#include "scan_config.hpp"
#include "scan_impl.hpp"
#include "tuning.hpp"

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
                                                     uint64_t* status_words, LookbackGuard* guard,
                                                     bool single_tile) {
  __shared__ ScanSmem<T, Op, BLOCK> sm;
  __shared__ uint32_t s_tile;
  int64_t tile = 0;
  if (!single_tile) {
    if (threadIdx.x == 0) s_tile = atomicAdd(header, 1u);
    __syncthreads();
    tile = wave_uniform(s_tile);
  }
  scan_tile<T, Op, BLOCK, ITEMS, EXCLUSIVE>(sm, in, out, n, init, has_init, tile, status_words,
                                            guard, single_tile, nullptr);
}

template <class T, int V>
static int64_t scan_num_tiles(int64_t n) {
  constexpr int TILE = ScanConfig<T, V>::TILE;
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
  // sized for the variant with the smallest tile, so any variant fits
  int64_t tiles = es == 4 ? scan_num_tiles<float, 1>(n) : scan_num_tiles<double, 1>(n);
  static_assert(ScanConfig<float, 1>::TILE <= ScanConfig<float, 3>::TILE &&
                    ScanConfig<double, 1>::TILE <= ScanConfig<double, 3>::TILE,
                "variant 1 has the smallest tile");
  if (tiles <= 1) return 0;
  return kHeaderBytes + size_t(tiles) * status_words_per_tile(dtype) * sizeof(uint64_t);
}

template <class T, class Op, bool EXCLUSIVE, int V>
static int scan_launch(const T* in, T* out, int64_t n, const T* init, char* ws,
                       hipStream_t stream) {
  using C = ScanConfig<T, V>;
  const int64_t tiles = scan_num_tiles<T, V>(n);
  using A = typename Acc<T, Op>::type;
  const T init_v = init ? *init : Op::template identity<T>();
  if (tiles <= 1) {
    scan_kernel<T, Op, C::BLOCK, C::ITEMS, EXCLUSIVE><<<1, C::BLOCK, 0, stream>>>(
        in, out, n, init_v, init != nullptr, nullptr, nullptr, nullptr, true);
    BULKX_LAUNCHED();
    return BULKX_SUCCESS;
  }
  LookbackGuard* guard = nullptr;
  BULKX_TRY(lookback_guard(&guard));
  uint32_t* header = reinterpret_cast<uint32_t*>(ws);
  uint64_t* words = reinterpret_cast<uint64_t*>(ws + kHeaderBytes);
  const int64_t num_words = tiles * TileStatus<A>::W;
  const int prep_blocks = int((num_words + 255) / 256 < 1024 ? (num_words + 255) / 256 : 1024);
  scan_prepare_kernel<<<prep_blocks, 256, 0, stream>>>(header, words, num_words);
  BULKX_LAUNCHED();
  scan_kernel<T, Op, C::BLOCK, C::ITEMS, EXCLUSIVE><<<dim3(uint32_t(tiles)), C::BLOCK, 0, stream>>>(
      in, out, n, init_v, init != nullptr, header, words, guard, false);
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
  Scratch ws;
  BULKX_TRY(ws.acquire(workspace, workspace_bytes, scan_workspace_bytes(dtype, n), stream));
  int st = BULKX_SUCCESS;
  const int v = tuned_variant(EXCLUSIVE ? BULKX_ALGO_EXCLUSIVE_SCAN : BULKX_ALGO_INCLUSIVE_SCAN,
                              dtype, n);
  BULKX_DISPATCH_DTYPE(dtype, T, BULKX_DISPATCH_OP(op, OP, {
    // the other shapes are instantiated for plus only (min / max use the default)
    bool done = false;
    if constexpr (std::is_same<OP, OpPlus>::value) {
      const T* i = static_cast<const T*>(in);
      T* o = static_cast<T*>(out);
      const T* ini = static_cast<const T*>(init);
      done = true;
      switch (v) {
        case 1: st = scan_launch<T, OP, EXCLUSIVE, 1>(i, o, n, ini, ws.ptr(), stream); break;
        case 2: st = scan_launch<T, OP, EXCLUSIVE, 2>(i, o, n, ini, ws.ptr(), stream); break;
        case 3: st = scan_launch<T, OP, EXCLUSIVE, 3>(i, o, n, ini, ws.ptr(), stream); break;
        case 4:
          if constexpr (ScanConfig<T, 4>::kWide)
            st = scan_launch<T, OP, EXCLUSIVE, 4>(i, o, n, ini, ws.ptr(), stream);
          else
            done = false;
          break;
        default: done = false; break;
      }
    }
    if (!done)
      st = scan_launch<T, OP, EXCLUSIVE, 0>(static_cast<const T*>(in), static_cast<T*>(out), n,
                                            static_cast<const T*>(init), ws.ptr(), stream);
  }));
  int rs = ws.release();
  return st != BULKX_SUCCESS ? st : rs;
}

// Exclusive plus-scan of uint32 values (in == out allowed) on `stream`, with
// a workspace of scan_workspace_bytes(BULKX_UINT32, n) bytes: the radix
// sort's digit counts become global positions with it.
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, char* ws,
                       hipStream_t stream) {
  if (n <= 0) return BULKX_SUCCESS;
  const uint32_t zero = 0;
  return scan_launch<uint32_t, OpPlus, true, 0>(in, out, n, &zero, ws, stream);
}

// Scratch of the allocation-free scan: the look-back header and status
// words live in this array, one copy per device (each device that loads the
// library's code object holds its own until the process ends).  4 MiB of
// status words plus the 256-byte header: enough for 2^32 4-byte integers,
// 2^31 floats, 2^30 8-byte integers or 2^29 doubles.
constexpr size_t kNoAllocScratchBytes = (size_t(4) << 20) + kHeaderBytes;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

// The device address of `scratch` on the current device, looked up once per
// device with that device current.  In host code the array's name is only a
// host-side stand-in; kernels get this address.
static int noalloc_scratch(char** out) {
  int dev = 0;
  BULKX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices) return BULKX_ERR_INTERNAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    BULKX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return BULKX_SUCCESS;
}

}  // namespace bulkx

extern "C" int bulkx_inclusive_scan(int dtype, int op, const void* in, void* out, int64_t n,
                                    const void* init, void* workspace, size_t workspace_bytes,
                                    void* stream) {
  return bulkx::scan_entry<false>(dtype, op, in, out, n, init, workspace, workspace_bytes, stream);
}

extern "C" int bulkx_inclusive_scan_noalloc(int dtype, int op, const void* in, void* out,
                                            int64_t n, const void* init, void* stream) {
  if (bulkx::dtype_size(dtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  if (n < 0) return BULKX_ERR_INVALID_ARGUMENT;
  if (bulkx::scan_workspace_bytes(dtype, n) > bulkx::kNoAllocScratchBytes)
    return BULKX_ERR_WORKSPACE_TOO_SMALL;
  char* ws = nullptr;
  BULKX_TRY(bulkx::noalloc_scratch(&ws));
  // the scan's preparation kernel clears the header and the status words
  // of every call, so the previous call's contents never leak in
  return bulkx::scan_entry<false>(dtype, op, in, out, n, init, ws, bulkx::kNoAllocScratchBytes,
                                  stream);
}

extern "C" int bulkx_exclusive_scan(int dtype, int op, const void* in, void* out, int64_t n,
                                    const void* init, void* workspace, size_t workspace_bytes,
                                    void* stream) {
  return bulkx::scan_entry<true>(dtype, op, in, out, n, init, workspace, workspace_bytes, stream);
}
