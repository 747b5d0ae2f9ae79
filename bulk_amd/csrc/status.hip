// The look-back guard of each device and the host's view of it.
//
// Every kernel of the plain entry points that waits on a predecessor tile
// (scan, single-pass reduce_by_key) gets the address of the current
// device's guard.  A wave that gives up after the spin limit sets the
// guard's error word; bulkx_stream_status turns it into BULKX_ERR_LOOKBACK,
// so the give-up surfaces the way a sticky launch error surfaces at the
// next synchronisation (the reference reports device failures through
// cudaGetLastError after a synchronize, bulk/detail/throw_on_error.hpp:33-49).
#include <atomic>
#include <mutex>

#include "device_common.hpp"
#include "host_common.hpp"

namespace bulkx {

// One guard per device: the code object is loaded on every device that uses
// the library, and each load holds its own copy of this variable.
__device__ LookbackGuard g_lookback_guard = {0u, kSpinLimit, 0u, 0u};

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
LookbackGuard* g_guard[kMaxDevices] = {};
std::atomic<uint32_t> g_spin_limit{kSpinLimit};
}  // namespace

int lookback_guard(LookbackGuard** out) {
  int dev = 0;
  BULKX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return BULKX_ERR_INTERNAL;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_guard[dev]) {
    void* p = nullptr;
    BULKX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(g_lookback_guard)));
    g_guard[dev] = static_cast<LookbackGuard*>(p);
  }
  *out = g_guard[dev];
  return BULKX_SUCCESS;
}

uint32_t host_spin_limit() { return g_spin_limit.load(); }

// Reads and clears the error word in one atomic step, so a give-up that
// lands while the host looks is never lost.
__global__ void guard_take_kernel(LookbackGuard* g) {
  if (threadIdx.x == 0) g->taken = atomicExch(&g->err, 0u);
}

}  // namespace bulkx

using namespace bulkx;

extern "C" int bulkx_stream_status(void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  LookbackGuard* g = nullptr;
  BULKX_TRY(lookback_guard(&g));
  std::lock_guard<std::mutex> lock(g_mu);  // one reader of `taken` at a time
  guard_take_kernel<<<1, 64, 0, stream>>>(g);
  BULKX_LAUNCHED();
  uint32_t taken = 0;
  BULKX_HIP(hipMemcpyAsync(&taken, &g->taken, sizeof(taken), hipMemcpyDeviceToHost, stream));
  BULKX_HIP(hipStreamSynchronize(stream));
  return taken ? BULKX_ERR_LOOKBACK : BULKX_SUCCESS;
}

extern "C" int bulkx_set_lookback_spin_limit(uint32_t polls) {
  LookbackGuard* g = nullptr;
  BULKX_TRY(lookback_guard(&g));
  BULKX_HIP(hipDeviceSynchronize());
  BULKX_HIP(hipMemcpy(&g->spin_limit, &polls, sizeof(polls), hipMemcpyHostToDevice));
  g_spin_limit.store(polls);
  return BULKX_SUCCESS;
}
