// GPU test of bulk::malloc / bulk::free / bulk::shmalloc / bulk::shfree on
// the group's LDS heap (include/bulkx/bulk.hpp), written like the
// reference's sum.cu (which lets the launcher size the heap, sum.cu:61) and
// against the first-fit semantics of bulk/malloc.hpp:133-199: blocks freed
// in any order, freed space reused, a free block at the end given back.
// Prints "PASS <name>" per check and exits non-zero on the first failure.
#include <bulkx/bulk.hpp>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(cond, name)                                               \
  do {                                                                  \
    if (!(cond)) {                                                      \
      std::fprintf(stderr, "FAIL %s (line %d)\n", name, __LINE__);      \
      std::exit(1);                                                     \
    }                                                                   \
    std::printf("PASS %s\n", name);                                     \
    std::fflush(stdout);                                                \
  } while (0)

#define HIP_OK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(2);                                                                       \
    }                                                                                     \
  } while (0)

// Whether a pointer addresses LDS.
__device__ inline bool on_chip(const void* p) {
  return __builtin_amdgcn_is_shared((const __attribute__((address_space(0))) void*)p);
}

// sum.cu's functor: a halving tree over the group's copy of the data in the heap
struct sum {
  __device__ void operator()(bulk::concurrent_group<>& g, const int* data, int* result,
                             int* lds_flag) {
    unsigned n = unsigned(g.size());
    int* s_data = static_cast<int*>(bulk::malloc(g, n * sizeof(int)));
    if (s_data == nullptr) {  // both heaps exhausted: report instead of writing through null
      if (g.this_exec.index() == 0) *result = -1;
      return;
    }
    if (g.this_exec.index() == 0) *lds_flag = on_chip(s_data) ? 1 : 0;
    bulk::copy_n(g, data, int(n), s_data);
    while (n > 1) {
      const unsigned half = n / 2;
      if (unsigned(g.this_exec.index()) < half) s_data[g.this_exec.index()] += s_data[n - g.this_exec.index() - 1];
      g.wait();
      n -= half;
    }
    if (g.this_exec.index() == 0) *result = s_data[0];
    g.wait();
    bulk::free(g, s_data);
  }
};

// Out-of-order frees.  out[0..7]: flags, each left at 1 when its check held
// (min over the groups of a grid).
struct out_of_order {
  __device__ void operator()(bulk::concurrent_group<>& g, int* out) {
    const int t = g.this_exec.index();
    const int m = 256;  // ints per block
    int* a = static_cast<int*>(bulk::malloc(g, m * sizeof(int)));
    int* b = static_cast<int*>(bulk::malloc(g, m * sizeof(int)));
    for (int i = t; i < m; i += g.size()) b[i] = 1000 + i;
    bulk::free(g, a);  // not the last block: b stays live
    int* c = static_cast<int*>(bulk::malloc(g, m * sizeof(int)));
    for (int i = t; i < m; i += g.size()) c[i] = -i;
    g.wait();
    bool ok_b = true;
    for (int i = t; i < m; i += g.size()) ok_b &= b[i] == 1000 + i;
    if (!ok_b) atomicMin(&out[0], 0);
    if (t == 0) {
      atomicMin(&out[1], (c == a) ? 1 : 0);  // first fit reuses a's block
      const bool overlap = (c < b + m) && (b < c + m);
      atomicMin(&out[2], overlap ? 0 : 1);
      atomicMin(&out[3], (on_chip(a) && on_chip(b)) ? 1 : 0);
    }
    // free in allocation order (reduce_by_key.hpp:211-212 does so): the
    // space is merged and returned, so one block of nearly the whole arena fits
    bulk::free(g, b);
    bulk::free(g, c);
    const std::size_t big = g.arena.bytes - 64;
    int* d = static_cast<int*>(bulk::malloc(g, big));
    if (t == 0) atomicMin(&out[4], (d == a && on_chip(d)) ? 1 : 0);
    bulk::free(g, d);
    // an arena with no room falls back to the device heap
    int* e = static_cast<int*>(bulk::malloc(g, g.arena.bytes + 4096));
    if (t == 0) atomicMin(&out[5], (e != nullptr && !on_chip(e)) ? 1 : 0);
    bulk::free(g, e);
    // three blocks, middle freed first, then the first: the two merge
    int* p = static_cast<int*>(bulk::malloc(g, 1024));
    int* q = static_cast<int*>(bulk::malloc(g, 1024));
    int* r = static_cast<int*>(bulk::malloc(g, 1024));
    bulk::free(g, q);
    bulk::free(g, p);
    int* s = static_cast<int*>(bulk::malloc(g, 2048));  // fits the merged hole
    if (t == 0) atomicMin(&out[6], (s == p) ? 1 : 0);
    bulk::free(g, r);
    bulk::free(g, s);
    if (t == 0) atomicMin(&out[7], 1);
  }
};

// shmalloc by every wavefront's lane 0 at once, then shfree in reverse
struct per_agent {
  __device__ void operator()(bulk::concurrent_group<>& g, int* out) {
    __shared__ int* ptrs[16];
    const int t = g.this_exec.index();
    const int w = t / 64;
    if (t % 64 == 0) {
      int* p = static_cast<int*>(bulk::shmalloc(g, 64 * sizeof(int)));
      ptrs[w] = p;
      for (int i = 0; i < 64; ++i) p[i] = w * 100 + i;
    }
    g.wait();
    const int nw = (g.size() + 63) / 64;
    if (t == 0) {
      bool ok = true;
      for (int i = 0; i < nw; ++i) {
        ok &= on_chip(ptrs[i]);
        for (int j = 0; j < nw; ++j)
          if (i != j) ok &= !(ptrs[i] < ptrs[j] + 64 && ptrs[j] < ptrs[i] + 64);
        for (int k = 0; k < 64; ++k) ok &= ptrs[i][k] == i * 100 + k;
      }
      out[0] = ok ? 1 : 0;
    }
    g.wait();
    if (t % 64 == 0) bulk::shfree(g, ptrs[nw - 1 - w]);
    g.wait();
    if (t == 0) {
      // everything given back: a block of nearly the whole arena fits
      int* big = static_cast<int*>(bulk::shmalloc(g, g.arena.bytes - 64));
      out[1] = on_chip(big) ? 1 : 0;
      bulk::shfree(g, big);
    }
  }
};

int main() {
  const int group_size = 512;
  std::vector<int> ones(group_size, 1);
  int *d_data, *d_result, *d_flag, *d_out;
  HIP_OK(hipMalloc(&d_data, group_size * sizeof(int)));
  HIP_OK(hipMalloc(&d_result, sizeof(int)));
  HIP_OK(hipMalloc(&d_flag, sizeof(int)));
  HIP_OK(hipMalloc(&d_out, 16 * sizeof(int)));
  HIP_OK(hipMemcpy(d_data, ones.data(), group_size * sizeof(int), hipMemcpyHostToDevice));
  using bulk::con;

  // let the runtime size the heap (sum.cu:61): the group's buffer is on chip
  {
    int r = 0, f = 0;
    HIP_OK(hipMemset(d_result, 0, sizeof(int)));
    bulk::async(con(std::size_t(group_size)), sum(), bulk::root, d_data, d_result, d_flag).wait();
    HIP_OK(hipMemcpy(&r, d_result, sizeof(int), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(&f, d_flag, sizeof(int), hipMemcpyDeviceToHost));
    CHECK(r == 512 && f == 1, "sum.cu: con(group_size) default heap, on chip");
  }
  // size the heap ourselves (sum.cu:67)
  {
    int r = 0, f = 0;
    HIP_OK(hipMemset(d_result, 0, sizeof(int)));
    bulk::async(con(std::size_t(group_size), group_size * sizeof(int)), sum(), bulk::root, d_data, d_result,
                d_flag)
        .wait();
    HIP_OK(hipMemcpy(&r, d_result, sizeof(int), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(&f, d_flag, sizeof(int), hipMemcpyDeviceToHost));
    CHECK(r == 512 && f == 1, "sum.cu: con(group_size, heap_size)");
  }
  // no heap: bulk::malloc falls back to the device heap
  {
    int r = 0, f = 1;
    HIP_OK(hipMemset(d_result, 0, sizeof(int)));
    bulk::async(con(std::size_t(group_size), std::size_t(0)), sum(), bulk::root, d_data, d_result, d_flag).wait();
    HIP_OK(hipMemcpy(&r, d_result, sizeof(int), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(&f, d_flag, sizeof(int), hipMemcpyDeviceToHost));
    CHECK(r == 512 && f == 0, "con(group_size, 0): device-heap fallback");
  }
  // out-of-order frees on a heap of 16 KB
  {
    std::vector<int> h(16, 1);
    HIP_OK(hipMemcpy(d_out, h.data(), 16 * sizeof(int), hipMemcpyHostToDevice));
    bulk::async(con(std::size_t(256), std::size_t(16 << 10)), out_of_order(), bulk::root, d_out).wait();
    HIP_OK(hipMemcpy(h.data(), d_out, 16 * sizeof(int), hipMemcpyDeviceToHost));
    CHECK(h[0] == 1, "free(a) keeps the later block b intact");
    CHECK(h[1] == 1 && h[2] == 1, "first fit reuses a freed block without overlapping a live one");
    CHECK(h[3] == 1, "blocks are in LDS");
    CHECK(h[4] == 1, "freed blocks merge and the break moves back");
    CHECK(h[5] == 1, "an oversize request falls back to the device heap");
    CHECK(h[6] == 1 && h[7] == 1, "adjacent free blocks merge");
  }
  // grid of groups with the default heap, each group out of order
  {
    std::vector<int> h(16, 1);
    HIP_OK(hipMemcpy(d_out, h.data(), 16 * sizeof(int), hipMemcpyHostToDevice));
    bulk::async(bulk::grid(64, 256, 16 << 10), out_of_order(), bulk::root.this_exec, d_out).wait();
    HIP_OK(hipMemcpy(h.data(), d_out, 16 * sizeof(int), hipMemcpyDeviceToHost));
    bool ok = true;
    for (int i = 0; i < 8; ++i) ok &= h[i] == 1;
    CHECK(ok, "grid(64, 256, heap): every group's heap behaves");
  }
  // per-agent shmalloc / shfree with concurrent callers
  {
    std::vector<int> h(16, 0);
    HIP_OK(hipMemcpy(d_out, h.data(), 16 * sizeof(int), hipMemcpyHostToDevice));
    bulk::async(con(std::size_t(1024)), per_agent(), bulk::root, d_out).wait();
    HIP_OK(hipMemcpy(h.data(), d_out, 16 * sizeof(int), hipMemcpyDeviceToHost));
    CHECK(h[0] == 1, "shmalloc from every wavefront at once: disjoint LDS blocks");
    CHECK(h[1] == 1, "shfree in reverse order gives the arena back");
  }
  HIP_OK(hipFree(d_data));
  HIP_OK(hipFree(d_result));
  HIP_OK(hipFree(d_flag));
  HIP_OK(hipFree(d_out));
  std::printf("ALL PASS\n");
  return 0;
}
