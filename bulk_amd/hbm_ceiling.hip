// HBM ceilings of this device for the roofline figures: the best
// hand-written streaming kernels for reading, copying, and reading two
// streams while writing an eighth of them (the shape of reduce_by_key's
// traffic), timed with HIP events on 1 GiB streams.  Prints one JSON object:
// the median GB/s of each (bytes read + written per second).
//
// Every kernel moves one tile of 256 threads x U 16-byte vectors per
// workgroup; the stream length is a multiple of every tile, so no access
// leaves its buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define HIP_OK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);  \
      std::exit(2);                                                                        \
    }                                                                                      \
  } while (0)

typedef unsigned v4u __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ inline v4u load(const v4u* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT>
__device__ inline void store(v4u* p, v4u v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <int U, bool NT>
__global__ __launch_bounds__(256) void read_kernel(const v4u* in, unsigned* sink) {
  const long base = long(blockIdx.x) * 256 * U + threadIdx.x;
  v4u v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) v[u] = load<NT>(in + base + u * 256);
  unsigned x = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) x ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  if (x == 0x9e3779b9u) *sink = x;  // keeps the loads; never true for the fill pattern
}

template <int U, bool NT>
__global__ __launch_bounds__(256) void copy_kernel(const v4u* in, v4u* out) {
  const long base = long(blockIdx.x) * 256 * U + threadIdx.x;
  v4u v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) v[u] = load<NT>(in + base + u * 256);
#pragma unroll
  for (int u = 0; u < U; ++u) store<NT>(out + base + u * 256, v[u]);
}

// reads a and b, writes one vector per eight read from each
template <int U>
__global__ __launch_bounds__(256) void read2_write8_kernel(const v4u* a, const v4u* b, v4u* out) {
  __shared__ v4u s[256];
  const long base = long(blockIdx.x) * 256 * U + threadIdx.x;
  v4u x = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int u = 0; u < U; ++u) x ^= load<true>(a + base + u * 256) ^ load<true>(b + base + u * 256);
  s[threadIdx.x] = x;
  __syncthreads();
  for (int t = threadIdx.x; t < 32 * U; t += 256)
    store<true>(out + long(blockIdx.x) * 32 * U + t, s[(t * 8) % 256]);
}

template <class F>
static double median_ms(F launch, int iters = 20) {
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  launch();
  launch();
  HIP_OK(hipDeviceSynchronize());
  std::vector<float> ts;
  for (int i = 0; i < iters; ++i) {
    HIP_OK(hipEventRecord(a));
    launch();
    HIP_OK(hipEventRecord(b));
    HIP_OK(hipEventSynchronize(b));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    ts.push_back(ms);
  }
  std::sort(ts.begin(), ts.end());
  HIP_OK(hipEventDestroy(a));
  HIP_OK(hipEventDestroy(b));
  return ts[ts.size() / 2];
}

int main() {
  constexpr long n4 = 1L << 26;  // 16-byte vectors per stream: 1 GiB
  constexpr int U = 4;
  static_assert(n4 % (256 * U) == 0, "whole tiles");
  const double bytes = double(n4) * 16;
  v4u *a, *b, *c;
  unsigned* sink;
  HIP_OK(hipMalloc(&a, n4 * 16));
  HIP_OK(hipMalloc(&b, n4 * 16));
  HIP_OK(hipMalloc(&c, n4 * 16));
  HIP_OK(hipMalloc(&sink, 64));
  HIP_OK(hipMemset(a, 1, n4 * 16));
  HIP_OK(hipMemset(b, 2, n4 * 16));
  const unsigned grid = unsigned(n4 / (256 * U));
  const double read_nt = median_ms([&] { read_kernel<U, true><<<grid, 256>>>(a, sink); });
  const double copy_nt = median_ms([&] { copy_kernel<U, true><<<grid, 256>>>(a, c); });
  const double copy_plain = median_ms([&] { copy_kernel<U, false><<<grid, 256>>>(a, c); });
  const double r2w8 = median_ms([&] { read2_write8_kernel<U><<<grid, 256>>>(a, b, c); });
  HIP_OK(hipGetLastError());
  std::printf(
      "{\"stream_bytes\": %.0f, \"read_nt_gbs\": %.1f, \"copy_nt_gbs\": %.1f, "
      "\"copy_plain_gbs\": %.1f, \"read2_write8_gbs\": %.1f}\n",
      bytes, bytes / read_nt / 1e6, 2 * bytes / copy_nt / 1e6, 2 * bytes / copy_plain / 1e6,
      (2 * bytes + bytes / 8) / r2w8 / 1e6);
  HIP_OK(hipFree(a));
  HIP_OK(hipFree(b));
  HIP_OK(hipFree(c));
  HIP_OK(hipFree(sink));
  return 0;
}
