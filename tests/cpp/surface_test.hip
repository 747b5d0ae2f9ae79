// GPU test of the header-only call surface (include/bulkx/bulk.hpp) and the
// C++ driver facade (include/bulkx/algorithm.hpp), written the way the
// reference's example programs use them (hello_world.cu, saxpy.cu,
// futures.cu, ping_pong.cu, scan.cu, reduce.cu, merge.cu).  Every check
// compares with a host computation; prints "PASS <name>" per check and exits
// non-zero on the first failure.
#include <bulkx/algorithm.hpp>
#include <bulkx/bulk.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <numeric>
#include <vector>

#define CHECK(cond, name)                                  \
  do {                                                     \
    if (!(cond)) {                                         \
      std::fprintf(stderr, "FAIL %s (line %d)\n", name, __LINE__); \
      std::exit(1);                                        \
    }                                                      \
    std::printf("PASS %s\n", name);                        \
    std::fflush(stdout);                                   \
  } while (0)

#define HIP_OK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(2);                                                      \
    }                                                                    \
  } while (0)

template <class T>
struct dvec {
  T* p = nullptr;
  size_t n = 0;
  explicit dvec(size_t n_) : n(n_) { HIP_OK(hipMalloc(&p, sizeof(T) * (n ? n : 1))); }
  dvec(const std::vector<T>& h) : dvec(h.size()) { put(h); }
  ~dvec() { (void)hipFree(p); }
  void put(const std::vector<T>& h) {
    HIP_OK(hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  }
  std::vector<T> get(size_t m = size_t(-1)) const {
    std::vector<T> h(m == size_t(-1) ? n : m);
    HIP_OK(hipMemcpy(h.data(), p, sizeof(T) * h.size(), hipMemcpyDeviceToHost));
    return h;
  }
};

// ------------------------------------------------------------- functors
struct add_op {
  template <class T>
  __host__ __device__ T operator()(const T& a, const T& b) const { return a + b; }
};

struct hello {
  int* out;
  __device__ void operator()(bulk::parallel_group<>& g) {
    out[g.this_exec.index()] = g.this_exec.index() * 10 + 1;
  }
};

struct saxpy {
  __device__ void operator()(bulk::agent<>& self, float a, float* x, float* y) {
    const int i = self.index();
    y[i] = a * x[i] + y[i];
  }
};

struct write_value {
  int* p;
  int v;
  __device__ void operator()() { *p = v; }
};

struct add_one_to {
  int* src;
  int* dst;
  __device__ void operator()() { *dst = *src + 1; }
};

// grid: every group stages its slice in LDS with bulk::malloc, then scans and reduces it
struct group_scan_reduce {
  template <std::size_t G, std::size_t S>
  __device__ void operator()(bulk::concurrent_group<bulk::agent<G>, S>& g, const int* in, int per,
                             int* scanned, int* sums) {
    const int* src = in + g.index() * per;
    int* stage = static_cast<int*>(bulk::malloc(g, per * sizeof(int)));
    bulk::copy_n(g, src, per, stage);
    bulk::inclusive_scan(g, stage, stage + per, scanned + g.index() * per, 7, add_op());
    int s = bulk::reduce(g, stage, stage + per, 0, add_op());
    if (g.this_exec.index() == 0) sums[g.index()] = s;
    bulk::free(g, stage);
  }
};

struct one_group {
  template <std::size_t G, std::size_t S>
  __device__ void operator()(bulk::concurrent_group<bulk::agent<G>, S>& g, const int* a, int na,
                             const int* b, int nb, int* merged, const int* x, int nx, int* ex) {
    bulk::merge(g, a, a + na, b, b + nb, merged, [](int u, int v) { return u < v; });
    bulk::exclusive_scan(g, x, x + nx, ex, 5, add_op());
  }
};

int main() {
  // hello: par(32), bulk::root is the group
  {
    dvec<int> out(32);
    bulk::async(bulk::par(32), hello{out.p}, bulk::root).wait();
    auto h = out.get();
    bool ok = true;
    for (int i = 0; i < 32; ++i) ok &= h[i] == i * 10 + 1;
    CHECK(ok, "par(32) with bulk::root");
  }
  // saxpy: par(n), bulk::root.this_exec is the agent
  {
    const int n = 1 << 20;
    dvec<float> x(std::vector<float>(n, 1.f)), y(std::vector<float>(n, 1.f));
    bulk::async(bulk::par(n), saxpy(), bulk::root.this_exec, 13.f, x.p, y.p).wait();
    auto h = y.get();
    CHECK(std::all_of(h.begin(), h.end(), [](float v) { return v == 14.f; }), "saxpy par(n)");
  }
  // futures: a task on a stream, a second task ordered after it by its future
  {
    hipStream_t s1;
    HIP_OK(hipStreamCreate(&s1));
    dvec<int> a(1), b(1);
    bulk::future<void> t1 = bulk::async(bulk::par(s1, 1), write_value{a.p, 41});
    bulk::future<void> t2 = bulk::async(bulk::par(t1, 1), add_one_to{a.p, b.p});
    t1.wait();
    t2.wait();
    CHECK(t1.valid() && t2.valid() && b.get()[0] == 42, "future ordering par(future, 1)");
    HIP_OK(hipStreamDestroy(s1));
  }
  // grid<256, 4>: LDS arena, copy_n, inclusive_scan with init, reduce
  {
    const int groups = 300, per = 3000;
    std::vector<int> h(groups * per);
    for (size_t i = 0; i < h.size(); ++i) h[i] = int((i * 2654435761u) % 1001) - 500;
    dvec<int> in(h), scanned(h.size()), sums(groups);
    bulk::async(bulk::grid<256, 4>(groups, per * sizeof(int)), group_scan_reduce(),
                bulk::root.this_exec, in.p, per, scanned.p, sums.p)
        .wait();
    auto hs = scanned.get();
    auto hr = sums.get();
    bool ok = true;
    for (int gi = 0; gi < groups && ok; ++gi) {
      int run = 7, tot = 0;
      for (int j = 0; j < per; ++j) {
        run += h[gi * per + j];
        tot += h[gi * per + j];
        ok &= hs[gi * per + j] == run;
      }
      ok &= hr[gi] == tot;
    }
    CHECK(ok, "grid<256,4> malloc + copy_n + inclusive_scan + reduce");
  }
  // con<512, 3>: one group merges and exclusive-scans
  {
    std::vector<int> a(1500), b(2201), x(4000);
    for (size_t i = 0; i < a.size(); ++i) a[i] = int(i / 3);
    for (size_t i = 0; i < b.size(); ++i) b[i] = int(i / 4);
    for (size_t i = 0; i < x.size(); ++i) x[i] = int(i % 7) - 3;
    dvec<int> da(a), db(b), dm(a.size() + b.size()), dx(x), de(x.size());
    bulk::async(bulk::con<512, 3>(0), one_group(), bulk::root, da.p, int(a.size()), db.p,
                int(b.size()), dm.p, dx.p, int(x.size()), de.p)
        .wait();
    std::vector<int> m(a.size() + b.size());
    std::merge(a.begin(), a.end(), b.begin(), b.end(), m.begin());
    std::vector<int> e(x.size());
    int run = 5;
    for (size_t i = 0; i < x.size(); ++i) {
      e[i] = run;
      run += x[i];
    }
    CHECK(dm.get() == m, "con<512,3> merge");
    CHECK(de.get() == e, "con<512,3> exclusive_scan");
  }
  // driver facade over the C-ABI
  {
    const int n = 1000003;
    std::vector<int> h(n);
    for (int i = 0; i < n; ++i) h[i] = (i * 7919) % 2001 - 1000;
    dvec<int> d(h), r(n);
    bulkx::inclusive_scan(d.p, d.p + n, r.p, 13, std::plus<int>());
    std::vector<int> ref(n);
    int run = 13;
    for (int i = 0; i < n; ++i) ref[i] = (run += h[i]);
    CHECK(r.get() == ref, "bulkx::inclusive_scan");
    const long long sum = std::accumulate(h.begin(), h.end(), 13LL);
    CHECK(bulkx::my_reduce(d.p, d.p + n, 13, std::plus<int>()) == int(sum), "bulkx::my_reduce");

    std::vector<float> a(70001), b(50000);
    for (size_t i = 0; i < a.size(); ++i) a[i] = float(i % 977) * 0.5f;
    for (size_t i = 0; i < b.size(); ++i) b[i] = float(i % 613) * 0.5f;
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    std::vector<int> va(a.size()), vb(b.size());
    std::iota(va.begin(), va.end(), 0);
    std::iota(vb.begin(), vb.end(), 1 << 28);
    dvec<float> da(a), db(b), dk(a.size() + b.size());
    dvec<int> dva(va), dvb(vb), dv(a.size() + b.size());
    bulkx::merge_by_key(da.p, da.p + a.size(), dva.p, db.p, db.p + b.size(), dvb.p, dk.p, dv.p,
                        std::less<float>());
    std::vector<std::pair<float, int>> pa, pb, pm(a.size() + b.size());
    for (size_t i = 0; i < a.size(); ++i) pa.push_back({a[i], va[i]});
    for (size_t i = 0; i < b.size(); ++i) pb.push_back({b[i], vb[i]});
    std::merge(pa.begin(), pa.end(), pb.begin(), pb.end(), pm.begin(),
               [](const auto& u, const auto& v) { return u.first < v.first; });
    auto hk = dk.get();
    auto hv = dv.get();
    bool ok = true;
    for (size_t i = 0; i < pm.size(); ++i) ok &= hk[i] == pm[i].first && hv[i] == pm[i].second;
    CHECK(ok, "bulkx::merge_by_key (stable: first range wins ties)");
    bulkx::my_merge(da.p, da.p + a.size(), db.p, db.p + b.size(), dk.p, std::less<float>());
    hk = dk.get();
    ok = true;
    for (size_t i = 0; i < pm.size(); ++i) ok &= hk[i] == pm[i].first;
    CHECK(ok, "bulkx::my_merge");

    std::vector<uint32_t> k(300007), v(300007);
    for (size_t i = 0; i < k.size(); ++i) k[i] = uint32_t((i * 2654435761u) % 5003);
    std::iota(v.begin(), v.end(), 0u);
    dvec<uint32_t> dk2(k), dv2(v);
    bulkx::stable_merge_sort_by_key(dk2.p, dk2.p + k.size(), dv2.p, std::less<uint32_t>());
    std::vector<size_t> order(k.size());
    std::iota(order.begin(), order.end(), size_t(0));
    std::stable_sort(order.begin(), order.end(), [&](size_t i, size_t j) { return k[i] < k[j]; });
    auto sk = dk2.get();
    auto sv = dv2.get();
    ok = true;
    for (size_t i = 0; i < k.size(); ++i) ok &= sk[i] == k[order[i]] && sv[i] == v[order[i]];
    CHECK(ok, "bulkx::stable_merge_sort_by_key");

    std::vector<int> keys(n), vals(n);
    for (int i = 0; i < n; ++i) {
      keys[i] = i / 9;
      vals[i] = i % 5;
    }
    dvec<int> dkk(keys), dvv(vals), ko(n), vo(n);
    auto ends = bulkx::my_reduce_by_key(dkk.p, dkk.p + n, dvv.p, ko.p, vo.p, std::equal_to<int>(),
                                        std::plus<int>());
    const long m = long(ends.first - ko.p);
    std::vector<int> rk, rv;
    for (int i = 0; i < n; ++i) {
      if (i == 0 || keys[i] != keys[i - 1]) {
        rk.push_back(keys[i]);
        rv.push_back(0);
      }
      rv.back() += vals[i];
    }
    CHECK(m == long(rk.size()) && ko.get(m) == rk && vo.get(m) == rv, "bulkx::my_reduce_by_key");

    dvec<int> parts(37);
    bulkx::reduce_intervals(d.p, bulkx::make_aligned_decomposition<long>(n, 37, 896), parts.p,
                            std::plus<int>());
    auto decomp = bulkx::make_aligned_decomposition<long>(n, 37, 896);
    auto hp = parts.get();
    ok = true;
    for (long gi = 0; gi < 37; ++gi) {
      auto rg = decomp[gi];
      long long s = 0;
      for (long i = rg.first; i < rg.second; ++i) s += h[i];
      ok &= hp[gi] == int(s);
    }
    CHECK(ok, "bulkx::reduce_intervals (aligned decomposition)");
  }
  std::printf("ALL PASS\n");
  return 0;
}
