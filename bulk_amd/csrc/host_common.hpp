// Host-side plumbing shared by the C-ABI entry points: status handling,
// dtype/op dispatch, device property cache and scratch management.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <string>

#include "../../include/bulkx.h"

namespace bulkx {

// Thread-local text of the last HIP failure (read by bulkx_last_error).
void set_last_error(const char* where, hipError_t e);

#define BULKX_HIP(expr)                                   \
  do {                                                    \
    hipError_t _e = (expr);                               \
    if (_e != hipSuccess) {                               \
      ::bulkx::set_last_error(#expr, _e);                 \
      return _e == hipErrorOutOfMemory ? BULKX_ERR_OUT_OF_MEMORY : BULKX_ERR_HIP; \
    }                                                     \
  } while (0)

#define BULKX_TRY(expr)            \
  do {                             \
    int _s = (expr);               \
    if (_s != BULKX_SUCCESS) return _s; \
  } while (0)

// Launch check: a kernel launch that failed to enqueue.
#define BULKX_LAUNCHED() BULKX_HIP(hipGetLastError())

struct DeviceProps {
  int device = -1;
  int num_cus = 0;
  int lds_per_cu = 0;
  int wave = 64;
  int max_threads = 1024;
  long long hbm_bytes = 0;
};

// Properties of the current device (cached per device id).  Returns
// BULKX_ERR_NO_DEVICE when no GPU is visible.
int current_device_props(DeviceProps* out);

inline size_t dtype_size(int dt) {
  switch (dt) {
    case BULKX_INT32: case BULKX_UINT32: case BULKX_FLOAT32: return 4;
    case BULKX_INT64: case BULKX_UINT64: case BULKX_FLOAT64: return 8;
    default: return 0;
  }
}

inline bool valid_op(int op) { return op == BULKX_PLUS || op == BULKX_MIN || op == BULKX_MAX; }
inline bool valid_cmp(int c) { return c == BULKX_LESS || c == BULKX_GREATER; }

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Scratch for one call: either the caller's workspace (checked against the
// required size) or a stream-ordered temporary allocation released when the
// guard is finished.
class Scratch {
 public:
  Scratch() = default;
  ~Scratch();
  // required == 0 is allowed (nothing is allocated).
  int acquire(void* ws, size_t ws_bytes, size_t required, hipStream_t stream);
  // Release a temporary allocation stream-ordered.
  int release();
  char* ptr() const { return ptr_; }

 private:
  char* ptr_ = nullptr;
  bool owned_ = false;
  hipStream_t stream_ = nullptr;
};

// Dispatch helpers ---------------------------------------------------------
#define BULKX_DISPATCH_DTYPE(dt, T, ...)                         \
  switch (dt) {                                                  \
    case BULKX_INT32: { using T = int32_t; __VA_ARGS__; break; }  \
    case BULKX_UINT32: { using T = uint32_t; __VA_ARGS__; break; } \
    case BULKX_INT64: { using T = int64_t; __VA_ARGS__; break; }  \
    case BULKX_UINT64: { using T = uint64_t; __VA_ARGS__; break; } \
    case BULKX_FLOAT32: { using T = float; __VA_ARGS__; break; }  \
    case BULKX_FLOAT64: { using T = double; __VA_ARGS__; break; } \
    default: return BULKX_ERR_UNSUPPORTED_DTYPE;                 \
  }

#define BULKX_DISPATCH_OP(op, OP, ...)                      \
  switch (op) {                                             \
    case BULKX_PLUS: { using OP = ::bulkx::OpPlus; __VA_ARGS__; break; } \
    case BULKX_MIN: { using OP = ::bulkx::OpMin; __VA_ARGS__; break; }   \
    case BULKX_MAX: { using OP = ::bulkx::OpMax; __VA_ARGS__; break; }   \
    default: return BULKX_ERR_UNSUPPORTED_OP;               \
  }

#define BULKX_DISPATCH_CMP(c, CMP, ...)                          \
  switch (c) {                                                   \
    case BULKX_LESS: { using CMP = ::bulkx::CmpLess; __VA_ARGS__; break; }     \
    case BULKX_GREATER: { using CMP = ::bulkx::CmpGreater; __VA_ARGS__; break; } \
    default: return BULKX_ERR_UNSUPPORTED_OP;                    \
  }

// The look-back guard of the current device (error word + spin limit), a
// device address valid for kernels of every translation unit.
struct LookbackGuard;
int lookback_guard(LookbackGuard** out);
// Spin limit the host last set (bulkx_set_lookback_spin_limit), for guards
// that live in other device memory (the batch executor's).
uint32_t host_spin_limit();

// Whether the batch executor can run job j (kind, dtypes, pointers):
// BULKX_SUCCESS or the status bulkx_batch_set_jobs would return for it.
int batch_check_job(const bulkx_job& j);

// Per-algorithm scratch sizes (defined next to each algorithm).
size_t scan_workspace_bytes(int dtype, int64_t n);
size_t reduce_workspace_bytes(int dtype, int64_t n);
size_t merge_workspace_bytes(int kdtype, int vdtype, int64_t n1, int64_t n2);
size_t sort_workspace_bytes(int kdtype, int vdtype, int64_t n);
size_t reduce_by_key_workspace_bytes(int kdtype, int vdtype, int64_t n);

// Exclusive plus-scan of uint32 values (scan.hip), workspace
// scan_workspace_bytes(BULKX_UINT32, n).
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, char* ws,
                       hipStream_t stream);

}  // namespace bulkx
