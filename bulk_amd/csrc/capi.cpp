// C-ABI utilities: status strings, device properties, scratch management,
// workspace sizing and the choose_sizes replacement.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "tuning.hpp"

namespace bulkx {

static thread_local std::string g_last_error;

void set_last_error(const char* where, hipError_t e) {
  g_last_error = std::string(where) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
}

int current_device_props(DeviceProps* out) {
  static std::mutex mu;
  static std::vector<DeviceProps> cache;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
      set_last_error("hipGetDeviceCount", e);
      (void)hipGetLastError();
      return BULKX_ERR_NO_DEVICE;
    }
    set_last_error("hipGetDevice", e);
    return BULKX_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (dev < int(cache.size()) && cache[dev].device == dev) {
    *out = cache[dev];
    return BULKX_SUCCESS;
  }
  hipDeviceProp_t prop;
  BULKX_HIP(hipGetDeviceProperties(&prop, dev));
  DeviceProps p;
  p.device = dev;
  p.num_cus = prop.multiProcessorCount;
  p.lds_per_cu = int(prop.maxSharedMemoryPerMultiProcessor);
  p.wave = prop.warpSize;
  p.max_threads = prop.maxThreadsPerBlock;
  p.hbm_bytes = (long long)prop.totalGlobalMem;
  if (int(cache.size()) <= dev) cache.resize(dev + 1);
  cache[dev] = p;
  *out = p;
  return BULKX_SUCCESS;
}

Scratch::~Scratch() { release(); }

int Scratch::acquire(void* ws, size_t ws_bytes, size_t required, hipStream_t stream) {
  stream_ = stream;
  if (required == 0) {
    ptr_ = static_cast<char*>(ws);
    return BULKX_SUCCESS;
  }
  if (ws != nullptr) {
    if (ws_bytes < required) return BULKX_ERR_WORKSPACE_TOO_SMALL;
    ptr_ = static_cast<char*>(ws);
    return BULKX_SUCCESS;
  }
  void* p = nullptr;
  BULKX_HIP(hipMallocAsync(&p, required, stream));
  ptr_ = static_cast<char*>(p);
  owned_ = true;
  return BULKX_SUCCESS;
}

int Scratch::release() {
  if (owned_ && ptr_) {
    owned_ = false;
    char* p = ptr_;
    ptr_ = nullptr;
    BULKX_HIP(hipFreeAsync(p, stream_));
  }
  return BULKX_SUCCESS;
}

}  // namespace bulkx

using namespace bulkx;

extern "C" const char* bulkx_status_string(int status) {
  switch (status) {
    case BULKX_SUCCESS: return "BULKX_SUCCESS";
    case BULKX_ERR_INVALID_ARGUMENT: return "BULKX_ERR_INVALID_ARGUMENT";
    case BULKX_ERR_UNSUPPORTED_DTYPE: return "BULKX_ERR_UNSUPPORTED_DTYPE";
    case BULKX_ERR_UNSUPPORTED_OP: return "BULKX_ERR_UNSUPPORTED_OP";
    case BULKX_ERR_WORKSPACE_TOO_SMALL: return "BULKX_ERR_WORKSPACE_TOO_SMALL";
    case BULKX_ERR_NO_DEVICE: return "BULKX_ERR_NO_DEVICE";
    case BULKX_ERR_HIP: return "BULKX_ERR_HIP";
    case BULKX_ERR_OUT_OF_MEMORY: return "BULKX_ERR_OUT_OF_MEMORY";
    case BULKX_ERR_BUDGET: return "BULKX_ERR_BUDGET";
    case BULKX_ERR_NOT_READY: return "BULKX_ERR_NOT_READY";
    case BULKX_ERR_INTERNAL: return "BULKX_ERR_INTERNAL";
    case BULKX_ERR_LOOKBACK: return "BULKX_ERR_LOOKBACK";
    default: return "BULKX_ERR_UNKNOWN";
  }
}

extern "C" const char* bulkx_last_error(void) { return g_last_error.c_str(); }

extern "C" const char* bulkx_version(void) { return "0.2.0"; }

extern "C" size_t bulkx_dtype_size(int dtype) { return dtype_size(dtype); }

extern "C" int bulkx_device_info(int* num_cus, int* lds_bytes_per_cu, int* wavefront_size,
                                 int* max_threads_per_group, long long* hbm_bytes) {
  DeviceProps p;
  BULKX_TRY(current_device_props(&p));
  if (num_cus) *num_cus = p.num_cus;
  if (lds_bytes_per_cu) *lds_bytes_per_cu = p.lds_per_cu;
  if (wavefront_size) *wavefront_size = p.wave;
  if (max_threads_per_group) *max_threads_per_group = p.max_threads;
  if (hbm_bytes) *hbm_bytes = p.hbm_bytes;
  return BULKX_SUCCESS;
}

extern "C" int bulkx_workspace_size(int algo, int dtype, int value_dtype, int64_t n1, int64_t n2,
                                    size_t* bytes) {
  if (bytes == nullptr || n1 < 0 || n2 < 0) return BULKX_ERR_INVALID_ARGUMENT;
  if (dtype_size(dtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  switch (algo) {
    case BULKX_ALGO_INCLUSIVE_SCAN:
    case BULKX_ALGO_EXCLUSIVE_SCAN: *bytes = scan_workspace_bytes(dtype, n1); break;
    case BULKX_ALGO_REDUCE: *bytes = reduce_workspace_bytes(dtype, n1); break;
    case BULKX_ALGO_REDUCE_INTERVALS: *bytes = 0; break;
    case BULKX_ALGO_MERGE: *bytes = merge_workspace_bytes(dtype, -1, n1, n2); break;
    case BULKX_ALGO_MERGE_BY_KEY:
      if (dtype_size(value_dtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
      *bytes = merge_workspace_bytes(dtype, value_dtype, n1, n2);
      break;
    case BULKX_ALGO_SORT_BY_KEY:
      if (value_dtype >= 0 && dtype_size(value_dtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
      *bytes = sort_workspace_bytes(dtype, value_dtype, n1);
      break;
    case BULKX_ALGO_REDUCE_BY_KEY:
      if (dtype_size(value_dtype) == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
      *bytes = reduce_by_key_workspace_bytes(dtype, value_dtype, n1);
      break;
    default: return BULKX_ERR_INVALID_ARGUMENT;
  }
  return BULKX_SUCCESS;
}
