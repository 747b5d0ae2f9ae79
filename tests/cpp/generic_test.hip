// GPU test of the device-wide drivers on operators and types outside the
// C-ABI's enums (include/bulkx/generic.hpp behind include/bulkx/algorithm.hpp)
// and of the call-surface pieces: the no-init group inclusive_scan, the
// bounded-group scans, copy_n, strided_iterator, uninitialized /
// uninitialized_array, choose_sizes and default-sized grids.  Every check
// compares with a host restatement; prints "PASS <name>" per check and exits
// non-zero on the first failure.
#include <bulkx/algorithm.hpp>
#include <bulkx/bulk.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <vector>

#define CHECK(cond, name)                                          \
  do {                                                             \
    if (!(cond)) {                                                 \
      std::fprintf(stderr, "FAIL %s (line %d)\n", name, __LINE__); \
      std::exit(1);                                                \
    }                                                              \
    std::printf("PASS %s\n", name);                                \
    std::fflush(stdout);                                           \
  } while (0)

#define HIP_OK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(2);                                                                       \
    }                                                                                     \
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

// An affine map x -> a*x + b over wrapping uint32; composition is
// associative, not commutative: (f then g)(x) = g(f(x)).
struct affine {
  uint32_t a, b;
  __host__ __device__ bool operator==(const affine& o) const { return a == o.a && b == o.b; }
};
struct compose {
  __host__ __device__ affine operator()(const affine& f, const affine& g) const {
    return affine{g.a * f.a, g.a * f.b + g.b};
  }
};

// A user key: ordered by its `rank`; `tag` rides along and shows stability.
struct item {
  int rank;
  int tag;
};
struct by_rank {
  __host__ __device__ bool operator()(const item& x, const item& y) const { return x.rank < y.rank; }
};
// a comparator outside the enums on plain ints: by the low 6 bits, then ties
struct by_low_bits {
  __host__ __device__ bool operator()(int x, int y) const { return (x & 63) < (y & 63); }
};
// keys equal when they fall in the same block of 4; values combined by max
struct same_quad {
  __host__ __device__ bool operator()(int x, int y) const { return (x >> 2) == (y >> 2); }
};
struct maxop {
  __host__ __device__ long long operator()(long long a, long long b) const { return a < b ? b : a; }
};
struct add_int {  // integer plus as a user functor (not the enum's plus)
  __host__ __device__ int operator()(int a, int b) const { return a + b; }
};

// A decomposition the library does not know: every partition has i + 1 tiles of 100.
struct ramp_decomposition {
  typedef long size_type;
  typedef std::pair<long, long> range;
  long n;
  long parts;
  __host__ __device__ range operator[](long i) const {
    const long b = 100 * i * (i + 1) / 2, e = b + 100 * (i + 1);
    return range(b < n ? b : n, e < n ? e : n);
  }
  __host__ __device__ long size() const { return parts; }
};

// ---------------------------------------------------------- group functors
// no-init inclusive_scan in place, bounded-group scans and copy_n
struct group_scans {
  __device__ void operator()(bulk::concurrent_group<bulk::agent<4>, 128>& g, int* data, int n,
                             int* incl, int* incl_noinit, int* excl, int* copied) {
    bulk::copy_n(g, data, n, copied);
    bulk::inclusive_scan(bulk::bound<512>(g), data, data + n, incl, 7, add_int());
    bulk::inclusive_scan(bulk::bound<512>(g), data, data + n, incl_noinit, add_int());
    bulk::exclusive_scan(bulk::bound<512>(g), data, data + n, excl, 7, add_int());
    // in place, no init, unbounded group form
    bulk::inclusive_scan(g, copied, copied + n, copied, add_int());
  }
};

// strided view: agent t sums elements t, t + S, t + 2S, ... of its group's slice
struct strided_sums {
  __device__ void operator()(bulk::concurrent_group<bulk::agent<1>, 64>& g, const int* x, int n,
                             int* out) {
    const int t = g.this_exec.index();
    auto first = bulk::make_strided_iterator(x + t, 64);
    auto last = bulk::make_strided_iterator(x + t + ((n - t + 63) / 64) * 64, 64);
    int s = 0;
    for (auto it = first; it < last; ++it) s += *it;
    out[t] = s;
    out[64 + t] = int(last - first);  // element count through the iterator difference
  }
};

struct counter {
  int value;
  __device__ counter(int v) : value(v) {}
};
// uninitialized<T> in LDS for a type with a constructor; uninitialized_array as a buffer
struct raw_storage {
  __device__ void operator()(bulk::concurrent_group<>& g, int* out) {
    __shared__ bulk::uninitialized<counter> s_c;
    __shared__ bulk::uninitialized_array<int, 256> s_a;
    const int t = g.this_exec.index();
    if (t == 0) s_c.construct(41);
    s_a[t] = t * 3;
    g.wait();
    if (t == 0) {
      counter& c = s_c;
      c.value += 1;
      int sum = 0;
      for (int v : s_a) sum += v;
      out[0] = s_c.get().value;
      out[1] = sum;
      out[2] = int(s_a.size()) + (s_a.empty() ? 1000 : 0);
      out[3] = s_a.front() + s_a.back();
      s_c.destroy();
    }
  }
};

struct count_groups {
  __device__ void operator()(bulk::concurrent_group<>& g, int* groups, int* size) {
    if (g.this_exec.index() == 0) atomicAdd(groups, 1);
    if (g.index() == 0 && g.this_exec.index() == 0) *size = g.size();
  }
};

int main() {
  std::mt19937 rng(7);

  // ---- scans with a non-commutative user operator and a user type
  {
    const int n = 300001;
    std::vector<affine> h(n);
    for (auto& f : h) f = affine{uint32_t(rng() | 1u), uint32_t(rng())};
    const affine init{3u, 5u};
    dvec<affine> d(h), r(n), e(n);
    bulkx::inclusive_scan(d.p, d.p + n, r.p, init, compose());
    bulkx::exclusive_scan(d.p, d.p + n, e.p, init, compose());
    std::vector<affine> incl(n), excl(n);
    affine acc = init;
    for (int i = 0; i < n; ++i) {
      excl[i] = acc;
      acc = compose()(acc, h[i]);
      incl[i] = acc;
    }
    CHECK(r.get() == incl, "generic inclusive_scan, non-commutative affine composition");
    CHECK(e.get() == excl, "generic exclusive_scan, non-commutative affine composition");
    const affine red = bulkx::my_reduce(d.p, d.p + n, init, compose());
    CHECK(red == acc, "generic my_reduce, non-commutative affine composition");
    auto f = bulkx::async_reduce(d.p, d.p + n, init, compose(), nullptr);
    CHECK(f.get() == acc, "generic async_reduce");
    // an integer plus written as a user functor takes the generic path too
    std::vector<int> xi(n);
    for (auto& v : xi) v = int(rng() % 2001) - 1000;
    dvec<int> dx(xi), dy(n);
    bulkx::inclusive_scan(dx.p, dx.p + n, dy.p, 13, add_int());
    std::vector<int> yi(n);
    int s = 13;
    for (int i = 0; i < n; ++i) yi[i] = s += xi[i];
    CHECK(dy.get() == yi, "generic inclusive_scan, user integer functor");
    CHECK(bulkx::my_reduce(dx.p, dx.p + n, 13, add_int()) == s, "generic my_reduce, user integer functor");
    CHECK(bulkx::my_reduce(dx.p, dx.p, 13, add_int()) == 13, "generic my_reduce, empty range gives init");
  }

  // ---- merge and merge_by_key with user comparators and a user key type
  {
    const int n1 = 70001, n2 = 50003;
    std::vector<int> a(n1), b(n2);
    for (auto& v : a) v = int(rng() % 100000);
    for (auto& v : b) v = int(rng() % 100000);
    std::stable_sort(a.begin(), a.end(), by_low_bits());
    std::stable_sort(b.begin(), b.end(), by_low_bits());
    dvec<int> da(a), db(b), dm(n1 + n2);
    bulkx::my_merge(da.p, da.p + n1, db.p, db.p + n2, dm.p, by_low_bits());
    std::vector<int> m(n1 + n2);
    std::merge(a.begin(), a.end(), b.begin(), b.end(), m.begin(), by_low_bits());
    CHECK(dm.get() == m, "generic my_merge, user comparator with ties (first range wins)");

    std::vector<item> ka(n1), kb(n2);
    for (int i = 0; i < n1; ++i) ka[i] = item{int(rng() % 1000), i};
    for (int i = 0; i < n2; ++i) kb[i] = item{int(rng() % 1000), 1000000 + i};
    std::stable_sort(ka.begin(), ka.end(), by_rank());
    std::stable_sort(kb.begin(), kb.end(), by_rank());
    std::vector<long long> va(n1), vb(n2);
    for (int i = 0; i < n1; ++i) va[i] = ka[i].tag * 7LL;
    for (int i = 0; i < n2; ++i) vb[i] = kb[i].tag * 7LL;
    dvec<item> dka(ka), dkb(kb), dko(n1 + n2);
    dvec<long long> dva(va), dvb(vb), dvo(n1 + n2);
    bulkx::merge_by_key(dka.p, dka.p + n1, dva.p, dkb.p, dkb.p + n2, dvb.p, dko.p, dvo.p, by_rank());
    std::vector<item> km(n1 + n2);
    std::merge(ka.begin(), ka.end(), kb.begin(), kb.end(), km.begin(), by_rank());
    auto ko = dko.get();
    auto vo = dvo.get();
    bool ok = true;
    for (int i = 0; i < n1 + n2; ++i) ok &= ko[i].rank == km[i].rank && ko[i].tag == km[i].tag && vo[i] == km[i].tag * 7LL;
    CHECK(ok, "generic merge_by_key, user key type");
  }

  // ---- stable sort by key with user comparators
  {
    const int n = 200003;
    std::vector<item> k(n);
    std::vector<int> v(n);
    for (int i = 0; i < n; ++i) {
      k[i] = item{int(rng() % 997), i};
      v[i] = i;
    }
    dvec<item> dk(k);
    dvec<int> dv(v);
    bulkx::stable_merge_sort_by_key(dk.p, dk.p + n, dv.p, by_rank());
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return k[x].rank < k[y].rank; });
    auto gk = dk.get();
    auto gv = dv.get();
    bool ok = true;
    for (int i = 0; i < n; ++i) ok &= gk[i].rank == k[order[i]].rank && gk[i].tag == order[i] && gv[i] == order[i];
    CHECK(ok, "generic stable_merge_sort_by_key, user key type (stable)");
    for (int m : {1, 2, 1023, 1024, 1025, 4097}) {
      std::vector<int> ki(m), vi(m);
      for (int i = 0; i < m; ++i) {
        ki[i] = int(rng() % 4096);
        vi[i] = i;
      }
      dvec<int> dki(ki), dvi(vi);
      bulkx::stable_merge_sort_by_key(dki.p, dki.p + m, dvi.p, by_low_bits());
      std::vector<int> ord(m);
      std::iota(ord.begin(), ord.end(), 0);
      std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return by_low_bits()(ki[x], ki[y]); });
      auto rk = dki.get();
      auto rv = dvi.get();
      bool okm = true;
      for (int i = 0; i < m; ++i) okm &= rk[i] == ki[ord[i]] && rv[i] == ord[i];
      CHECK(okm, "generic stable_merge_sort_by_key, int keys by a user comparator");
    }
  }

  // ---- reduce_by_key with a user predicate and operator
  {
    const int n = 250007;
    std::vector<int> k(n);
    std::vector<long long> v(n);
    int cur = 0;
    for (int i = 0; i < n; ++i) {
      if (rng() % 5 == 0) cur += 1 + int(rng() % 3);
      k[i] = cur;
      v[i] = (long long)(rng() % 100000) - 50000;
    }
    dvec<int> dk(k), ko(n);
    dvec<long long> dv(v), vo(n);
    auto ends = bulkx::my_reduce_by_key(dk.p, dk.p + n, dv.p, ko.p, vo.p, same_quad(), maxop());
    std::vector<int> rk;
    std::vector<long long> rv;
    for (int i = 0; i < n; ++i) {
      if (i == 0 || !same_quad()(k[i - 1], k[i])) rv.push_back(v[i]);
      else rv.back() = maxop()(rv.back(), v[i]);
      if (i == n - 1 || !same_quad()(k[i], k[i + 1])) rk.push_back(k[i]);  // the run's last key
    }
    const long m = long(ends.first - ko.p);
    CHECK(m == long(rk.size()) && ends.second - vo.p == m, "generic my_reduce_by_key, run count");
    CHECK(ko.get(m) == rk && vo.get(m) == rv, "generic my_reduce_by_key, last keys and max per run");
    auto fut = bulkx::async_reduce_by_key(dk.p, dk.p + n, dv.p, ko.p, vo.p, same_quad(), maxop(), nullptr);
    CHECK(fut.num_runs() == m, "generic async_reduce_by_key");
  }

  // ---- reduce_intervals over every decomposition
  {
    const long n = 100000;
    std::vector<int> x(n);
    for (auto& v : x) v = int(rng() % 2001) - 1000;
    dvec<int> dx(x);
    auto host = [&](long b, long e) {
      long long s = 0;
      for (long i = b; i < e; ++i) s += x[i];
      return int(s);
    };
    {
      auto dd = bulkx::make_blocked_decomposition<long>(n, 777);
      dvec<int> out(dd.size());
      bulkx::reduce_intervals(dx.p, dd, out.p, std::plus<int>());
      auto h = out.get();
      bool ok = dd.size() == (n + 776) / 777;
      for (long i = 0; i < dd.size(); ++i) ok &= h[i] == host(dd[i].first, dd[i].second);
      CHECK(ok, "reduce_intervals, blocked decomposition (library kernel)");
      dvec<int> out2(dd.size());
      bulkx::reduce_intervals(dx.p, dx.p + n, 777L, out2.p, add_int());
      CHECK(out2.get() == h, "reduce_intervals(first, last, interval_size), user functor");
    }
    {
      auto ud = bulkx::make_uniform_decomposition<long>(n, 33);
      dvec<int> out(ud.size());
      bulkx::reduce_intervals(dx.p, ud, out.p, std::plus<int>());
      auto h = out.get();
      bool ok = ud.size() == (n + n / 33 - 1) / (n / 33);
      for (long i = 0; i < ud.size(); ++i) ok &= h[i] == host(ud[i].first, ud[i].second);
      CHECK(ok, "reduce_intervals, uniform decomposition");
    }
    {
      auto td = bulkx::make_trivial_decomposition<long>(n);
      dvec<int> out(1);
      bulkx::reduce_intervals(dx.p, td, out.p, std::plus<int>());
      CHECK(out.get()[0] == host(0, n), "reduce_intervals, trivial decomposition");
    }
    {
      ramp_decomposition rd{n, 40};
      dvec<int> out(rd.size());
      bulkx::reduce_intervals(dx.p, rd, out.p, std::plus<int>());
      auto h = out.get();
      bool ok = true;
      for (long i = 0; i < rd.size(); ++i)
        if (rd[i].second > rd[i].first) ok &= h[i] == host(rd[i].first, rd[i].second);
      CHECK(ok, "reduce_intervals, a user decomposition (generic path)");
    }
  }

  // ---- group scans: no-init, bounded forms, copy_n
  {
    const int n = 500;  // within the bound of 512 = 128 agents x 4
    std::vector<int> h(n);
    for (auto& v : h) v = int(rng() % 200) - 100;
    dvec<int> d(h), incl(n), noinit(n), excl(n), copied(n);
    bulk::async(bulk::con<128, 4>(0), group_scans(), bulk::root, d.p, n, incl.p, noinit.p, excl.p,
                copied.p)
        .wait();
    std::vector<int> ri(n), rn(n), re(n);
    int s = 7;
    for (int i = 0; i < n; ++i) {
      re[i] = s;
      s += h[i];
      ri[i] = s;
    }
    s = 0;
    for (int i = 0; i < n; ++i) rn[i] = s += h[i];
    CHECK(incl.get() == ri, "bounded-group inclusive_scan with init");
    CHECK(noinit.get() == rn, "bounded-group inclusive_scan without init");
    CHECK(excl.get() == re, "bounded-group exclusive_scan (exclusive)");
    CHECK(copied.get() == rn, "group copy_n, then no-init inclusive_scan in place");
  }

  // ---- strided_iterator
  {
    const int n = 1000;
    std::vector<int> h(n);
    for (int i = 0; i < n; ++i) h[i] = i * i % 97;
    dvec<int> d(h), out(128);
    bulk::async(bulk::con<64, 1>(0), strided_sums(), bulk::root, d.p, n, out.p).wait();
    auto o = out.get();
    bool ok = true;
    for (int t = 0; t < 64; ++t) {
      int sum = 0, cnt = 0;
      for (int i = t; i < n; i += 64) {
        sum += h[i];
        ++cnt;
      }
      ok &= o[t] == sum && o[64 + t] == cnt;
    }
    CHECK(ok, "strided_iterator: strided sums and distances");
    int arr[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
    auto a = bulk::make_strided_iterator(arr, 3);
    auto b = a + 3;
    CHECK(*b == 9 && b - a == 3 && a[2] == 6 && (bulk::make_strided_iterator(arr + 10, 3) - a) == 4,
          "strided_iterator arithmetic (distance rounds up, as the reference)");
  }

  // ---- uninitialized / uninitialized_array
  {
    dvec<int> out(4);
    bulk::async(bulk::con(std::size_t(256), std::size_t(0)), raw_storage(), bulk::root, out.p).wait();
    auto o = out.get();
    CHECK(o[0] == 42 && o[1] == 3 * 255 * 256 / 2 && o[2] == 256 && o[3] == 255 * 3,
          "uninitialized<T> in LDS (construct / destroy) and uninitialized_array");
  }

  // ---- choose_sizes and default-sized grids
  {
    dvec<int> groups(1), size(1);
    HIP_OK(hipMemset(groups.p, 0, sizeof(int)));
    auto sizes = bulk::choose_sizes(bulk::grid(), count_groups(), bulk::root.this_exec, groups.p, size.p);
    int cus = 0, per_cu = 0, dev = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_OK(hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxThreadsPerMultiProcessor, dev));
    CHECK(sizes.second >= 64 && sizes.second <= 1024 && sizes.first == (per_cu / sizes.second) * cus,
          "choose_sizes(grid(), f, args...): occupancy group size, subscription x CUs groups");
    bulk::async(bulk::grid(), count_groups(), bulk::root.this_exec, groups.p, size.p).wait();
    CHECK(groups.get()[0] == sizes.first && size.get()[0] == sizes.second,
          "bulk::async(grid()) launches the sizes choose_sizes reports");
    auto s2 = bulk::choose_sizes(bulk::grid(100, 128), count_groups(), bulk::root.this_exec, groups.p, size.p);
    CHECK(s2.first == 100 && s2.second == 128, "choose_sizes keeps requested sizes");
  }

  std::printf("ALL PASS\n");
  return 0;
}
