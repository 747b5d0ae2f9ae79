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

// bounded-agent bodies: each agent works on its own grain of at most 8 elements
struct lcg_op {  // sequential fold that is neither commutative nor associative
  __host__ __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a * 3u + b; }
};

struct agent_bodies {
  __device__ void operator()(bulk::agent<>& self, const int* in, int n, int* incl, int* excl,
                             int* sums, uint32_t* folds, int* skeys, int* svals, const int* map,
                             const int* stencil, int* gathered, int* scattered) {
    const int t = self.index();
    const int lo = t * 8;
    const int cnt = n - lo < 8 ? n - lo : 8;
    auto& b = bulk::bound<8>(self);
    bulk::inclusive_scan(b, in + lo, in + lo + cnt, incl + lo, 3, add_op());
    bulk::exclusive_scan(b, in + lo, in + lo + cnt, excl + lo, 3, add_op());
    sums[t] = bulk::reduce(b, in + lo, in + lo + cnt, 0, add_op());
    folds[t] = bulk::accumulate(b, in + lo, in + lo + cnt, 7u, lcg_op());
    bulk::stable_sort_by_key(b, skeys + lo, skeys + lo + cnt, svals + lo,
                             [](int u, int v) { return u < v; });
    bulk::gather(b, map + lo, map + lo + cnt, in, gathered + lo);
    bulk::scatter_if(b, in + lo, in + lo + cnt, map + lo, stencil + lo, scattered);
  }
};

// bounded-agent merge and merge_by_key of two sorted 4- and 5-element grains
struct agent_merges {
  __device__ void operator()(bulk::agent<>& self, const int* a, const int* b, const int* va,
                             const int* vb, int* m, int* mk, int* mv) {
    const int t = self.index();
    auto& e = bulk::bound<9>(self);
    auto less = [](int u, int v) { return u < v; };
    bulk::merge(e, a + 4 * t, a + 4 * t + 4, b + 5 * t, b + 5 * t + 5, m + 9 * t, less);
    bulk::merge_by_key(e, a + 4 * t, a + 4 * t + 4, b + 5 * t, b + 5 * t + 5, va + 4 * t,
                       vb + 5 * t, mk + 9 * t, mv + 9 * t, less);
  }
};

// bounded group: stable_sort_by_key staged in the LDS arena, inplace_merge in global memory
struct group_sort_merge {
  template <std::size_t G, std::size_t S>
  __device__ void operator()(bulk::concurrent_group<bulk::agent<G>, S>& g, int* keys, int* vals,
                             int n, int* x, int mid, int nx) {
    int* sk = static_cast<int*>(bulk::malloc(g, n * sizeof(int)));
    int* sv = static_cast<int*>(bulk::malloc(g, n * sizeof(int)));
    bulk::copy_n(g, keys, n, sk);
    bulk::copy_n(g, vals, n, sv);
    auto less = [](int u, int v) { return u < v; };
    bulk::stable_sort_by_key(bulk::bound<S * G>(g), sk, sk + n, sv, less);
    bulk::copy_n(g, sk, n, keys);
    bulk::copy_n(g, sv, n, vals);
    bulk::free(g, sk);
    bulk::inplace_merge(bulk::bound<S * G>(g), x, x + mid, x + nx, less);
  }
};

struct affine {
  uint32_t a, b;
};
struct compose {  // apply f, then h: associative, not commutative
  __host__ __device__ affine operator()(affine f, affine h) const {
    return affine{f.a * h.a, f.b * h.a + h.b};
  }
};

struct rbk_result {
  long long nout;
  int carry_key;
  int carry_val;
};

// group: merge_by_key, accumulate, reduce_by_key, scatter_if, gather, flag ranges
struct group_more {
  template <std::size_t G, std::size_t S>
  __device__ void operator()(bulk::concurrent_group<bulk::agent<G>, S>& g, const int* a, int na,
                             const int* va, const int* b, int nb, const int* vb, int* mk, int* mv,
                             const affine* f, int nf, affine* acc, const int* keys,
                             const int* vals, int n, int init_key, int* ko, int* vo,
                             rbk_result* res, const int* map, const int* stencil, int* gathered,
                             int* scattered, int* flag_counts) {
    auto less = [](int u, int v) { return u < v; };
    bulk::merge_by_key(g, a, a + na, b, b + nb, va, vb, mk, mv, less);
    const affine r = bulk::accumulate(g, f, f + nf, affine{1u, 0u}, compose());
    if (g.this_exec.index() == 0) *acc = r;
    auto eq = [](int u, int v) { return u == v; };
    auto t = bulk::reduce_by_key(g, keys, keys + n, vals, ko, vo, init_key, 100, eq, add_op());
    if (g.this_exec.index() == 0)
      *res = rbk_result{(long long)(thrust::get<0>(t) - ko), thrust::get<2>(t), thrust::get<3>(t)};
    bulk::gather(g, map, map + n, vals, gathered);
    bulk::scatter_if(g, vals, vals + n, map, stencil, scattered);
    auto hf = bulk::make_head_flags(keys, keys + n);
    auto tf = bulk::make_tail_flags(keys, keys + n, eq);
    bulk::head_flags_with_init<const int*> hfi(keys, keys + n, init_key);
    const int heads = bulk::reduce(g, hf.begin(), hf.end(), 0, add_op());
    const int tails = bulk::reduce(g, tf.begin(), tf.end(), 0, add_op());
    const int heads_init = bulk::reduce(g, hfi.begin(), hfi.end(), 0, add_op());
    if (g.this_exec.index() == 0) {
      flag_counts[0] = heads;
      flag_counts[1] = tails;
      flag_counts[2] = heads_init;
    }
  }
};

// adjacent_difference: a group over several tiles, out of place with init and
// in place without; one agent over its own 16 elements
struct group_adjdiff {
  template <std::size_t G, std::size_t S>
  __device__ void operator()(bulk::concurrent_group<bulk::agent<G>, S>& g, const int* x, int n,
                             int* d, int* inplace) {
    auto sub = [](int a, int b) { return a - b; };
    bulk::adjacent_difference(g, x, x + n, d, 100, sub);
    bulk::adjacent_difference(g, inplace, inplace + n, inplace, sub);
  }
};

struct agent_adjdiff {
  __device__ void operator()(bulk::agent<>& self, const int* x, int* d) {
    const int t = self.index();
    bulk::adjacent_difference(self, x + 16 * t, x + 16 * t + 16, d + 16 * t, -1,
                              [](int a, int b) { return a - b; });
  }
};

// for_each_n on a group and on bounded agents: add 1 to every element
struct add_one {
  __device__ void operator()(int& x) const { x += 1; }
};
struct group_for_each {
  template <std::size_t G, std::size_t S>
  __device__ void operator()(bulk::concurrent_group<bulk::agent<G>, S>& g, int* x, int n) {
    bulk::for_each_n(g, x, n, add_one());
  }
};
struct agent_for_each {
  __device__ void operator()(bulk::agent<>& self, int* x, int n) {
    const int lo = self.index() * 4;
    bulk::for_each_n(bulk::bound<4>(self), x + lo, n - lo < 4 ? n - lo : 4, add_one());
  }
};

// join_iterator: a group copies and reduces two ranges viewed as one
struct group_join {
  template <std::size_t G, std::size_t S>
  __device__ void operator()(bulk::concurrent_group<bulk::agent<G>, S>& g, const int* a, int na,
                             const int* b, int nb, int* out, int* sum) {
    auto j = bulk::make_join_iterator(a, na, b);
    bulk::copy_n(g, j, na + nb, out);
    const int s = bulk::reduce(g, j, j + (na + nb), 0, add_op());
    if (g.this_exec.index() == 0) *sum = s;
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
  // bounded-agent bodies over grains of 8 (the last one ragged)
  {
    const int n = 8 * 123 + 5, agents = (n + 7) / 8;
    std::vector<int> in(n), skeys(n), svals(n), map(n), stencil(n);
    for (int i = 0; i < n; ++i) {
      in[i] = int((i * 2654435761u) % 201) - 100;
      skeys[i] = in[i] & 3;
      svals[i] = i;
      map[i] = int((uint64_t(i) * 7919u) % uint64_t(n));  // a permutation (7919 is prime, n is not a multiple)
      stencil[i] = i % 3 != 0;
    }
    dvec<int> din(in), dincl(n), dexcl(n), dsums(agents), dsk(skeys), dsv(svals), dmap(map),
        dst(stencil), dga(n), dsc(std::vector<int>(n, -7));
    dvec<uint32_t> dfold(agents);
    bulk::async(bulk::par(agents), agent_bodies(), bulk::root.this_exec, din.p, n, dincl.p, dexcl.p,
                dsums.p, dfold.p, dsk.p, dsv.p, dmap.p, dst.p, dga.p, dsc.p)
        .wait();
    auto hi = dincl.get(), he = dexcl.get(), hs = dsums.get(), hk = dsk.get(), hv = dsv.get(),
         hg = dga.get(), hsc = dsc.get();
    auto hf = dfold.get();
    bool ok_scan = true, ok_red = true, ok_sort = true;
    for (int t = 0; t < agents; ++t) {
      const int lo = t * 8, cnt = std::min(8, n - lo);
      int run = 3, sum = 0;
      uint32_t fold = 7u;
      std::vector<std::pair<int, int>> grain;
      for (int j = 0; j < cnt; ++j) {
        ok_scan &= he[lo + j] == run;
        run += in[lo + j];
        ok_scan &= hi[lo + j] == run;
        sum += in[lo + j];
        fold = fold * 3u + uint32_t(in[lo + j]);
        grain.push_back({skeys[lo + j], svals[lo + j]});
      }
      ok_red &= hs[t] == sum && hf[t] == fold;
      std::stable_sort(grain.begin(), grain.end(),
                       [](const auto& u, const auto& v) { return u.first < v.first; });
      for (int j = 0; j < cnt; ++j)
        ok_sort &= hk[lo + j] == grain[j].first && hv[lo + j] == grain[j].second;
    }
    CHECK(ok_scan, "bounded agent inclusive_scan / exclusive_scan");
    CHECK(ok_red, "bounded agent reduce / accumulate (sequential left fold)");
    CHECK(ok_sort, "bounded agent stable_sort_by_key (odd-even transposition, ties keep order)");
    bool ok_g = true, ok_s = true;
    for (int i = 0; i < n; ++i) ok_g &= hg[i] == in[map[i]];
    std::vector<int> sc(n, -7);
    for (int i = 0; i < n; ++i)
      if (stencil[i]) sc[map[i]] = in[i];
    for (int i = 0; i < n; ++i) ok_s &= hsc[i] == sc[i];
    CHECK(ok_g, "bounded agent gather");
    CHECK(ok_s, "bounded agent scatter_if");
  }
  // bounded-agent merge / merge_by_key with ties (the first range wins)
  {
    const int agents = 200;
    std::vector<int> a(4 * agents), b(5 * agents), va(a.size()), vb(b.size());
    for (int t = 0; t < agents; ++t) {
      for (int j = 0; j < 4; ++j) a[4 * t + j] = (t % 7) + j * (t % 3);
      for (int j = 0; j < 5; ++j) b[5 * t + j] = (t % 5) + j * (t % 2);
    }
    for (size_t i = 0; i < a.size(); ++i) va[i] = int(i);
    for (size_t i = 0; i < b.size(); ++i) vb[i] = 100000 + int(i);
    dvec<int> da(a), db(b), dva(va), dvb(vb), dm(9 * agents), dmk(9 * agents), dmv(9 * agents);
    bulk::async(bulk::par(agents), agent_merges(), bulk::root.this_exec, da.p, db.p, dva.p, dvb.p,
                dm.p, dmk.p, dmv.p)
        .wait();
    auto hm = dm.get(), hk = dmk.get(), hv = dmv.get();
    bool ok = true;
    for (int t = 0; t < agents; ++t) {
      std::vector<std::pair<int, int>> pa, pb, pm(9);
      for (int j = 0; j < 4; ++j) pa.push_back({a[4 * t + j], va[4 * t + j]});
      for (int j = 0; j < 5; ++j) pb.push_back({b[5 * t + j], vb[5 * t + j]});
      std::merge(pa.begin(), pa.end(), pb.begin(), pb.end(), pm.begin(),
                 [](const auto& u, const auto& v) { return u.first < v.first; });
      for (int j = 0; j < 9; ++j)
        ok &= hm[9 * t + j] == pm[j].first && hk[9 * t + j] == pm[j].first &&
              hv[9 * t + j] == pm[j].second;
    }
    CHECK(ok, "bounded agent merge / merge_by_key");
  }
  // con<256, 4>: bounded-group stable_sort_by_key in LDS, inplace_merge
  {
    const int n = 1000, nx = 990, mid = 400;
    std::vector<int> keys(n), vals(n), x(nx);
    for (int i = 0; i < n; ++i) {
      keys[i] = int((i * 2654435761u) % 37);
      vals[i] = i;
    }
    for (int i = 0; i < mid; ++i) x[i] = i / 3;
    for (int i = mid; i < nx; ++i) x[i] = (i - mid) / 5;
    dvec<int> dk(keys), dv(vals), dx(x);
    bulk::async(bulk::con<256, 4>(2 * n * sizeof(int)), group_sort_merge(), bulk::root, dk.p, dv.p,
                n, dx.p, mid, nx)
        .wait();
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), size_t(0));
    std::stable_sort(order.begin(), order.end(), [&](size_t i, size_t j) { return keys[i] < keys[j]; });
    auto hk = dk.get(), hv = dv.get();
    bool ok = true;
    for (int i = 0; i < n; ++i) ok &= hk[i] == keys[order[i]] && hv[i] == vals[order[i]];
    CHECK(ok, "bounded group stable_sort_by_key (LDS arena)");
    std::vector<int> xm(x);
    std::inplace_merge(xm.begin(), xm.begin() + mid, xm.end());
    CHECK(dx.get() == xm, "bounded group inplace_merge");
  }
  // con<256, 4>: group merge_by_key, accumulate, reduce_by_key, scatter_if, gather, flags
  for (int variant = 0; variant < 2; ++variant) {
    const int na = 3001, nb = 2500, nf = 5000, n = 5003;
    std::vector<int> a(na), va(na), b(nb), vb(nb);
    for (int i = 0; i < na; ++i) {
      a[i] = i / 4;
      va[i] = i;
    }
    for (int i = 0; i < nb; ++i) {
      b[i] = i / 3;
      vb[i] = 1 << 20 | i;
    }
    std::vector<affine> f(nf);
    for (int i = 0; i < nf; ++i) f[i] = affine{uint32_t(i * 2654435761u) | 1u, uint32_t(i * 40503u)};
    std::vector<int> keys(n), vals(n), map(n), stencil(n);
    {
      int key = 11, left = 1;
      uint32_t r = 12345u;
      for (int i = 0; i < n; ++i) {
        if (--left == 0) {
          ++key;
          r = r * 1664525u + 1013904223u;
          left = 1 + int((r >> 16) % 17);
        }
        keys[i] = key;
        vals[i] = int((i * 2654435761u) % 1001) - 500;
        map[i] = int((uint64_t(i) * 7919u) % uint64_t(n));
        stencil[i] = (i % 5) != 2;
      }
    }
    // variant 0: the carried run continues into the first keys; 1: it ends before them
    const int init_key = variant == 0 ? keys[0] : -5;
    dvec<int> da(a), dva(va), db(b), dvb(vb), dmk(na + nb), dmv(na + nb), dkeys(keys), dvals(vals),
        dko(n + 1), dvo(n + 1), dmap(map), dst(stencil), dga(n), dsc(std::vector<int>(n, -9)),
        dfc(3);
    dvec<affine> df(f), dacc(1);
    dvec<rbk_result> dres(1);
    bulk::async(bulk::con<256, 4>(0), group_more(), bulk::root, da.p, na, dva.p, db.p, nb, dvb.p,
                dmk.p, dmv.p, df.p, nf, dacc.p, dkeys.p, dvals.p, n, init_key, dko.p, dvo.p,
                dres.p, dmap.p, dst.p, dga.p, dsc.p, dfc.p)
        .wait();
    if (variant == 0) {
      std::vector<std::pair<int, int>> pa, pb, pm(na + nb);
      for (int i = 0; i < na; ++i) pa.push_back({a[i], va[i]});
      for (int i = 0; i < nb; ++i) pb.push_back({b[i], vb[i]});
      std::merge(pa.begin(), pa.end(), pb.begin(), pb.end(), pm.begin(),
                 [](const auto& u, const auto& v) { return u.first < v.first; });
      auto hk = dmk.get(), hv = dmv.get();
      bool ok = true;
      for (int i = 0; i < na + nb; ++i) ok &= hk[i] == pm[i].first && hv[i] == pm[i].second;
      CHECK(ok, "group merge_by_key (ties take the first range)");
      affine e{1u, 0u};
      for (int i = 0; i < nf; ++i) e = compose()(e, f[i]);
      const affine got = dacc.get()[0];
      CHECK(got.a == e.a && got.b == e.b, "group accumulate (in order, non-commutative op)");
    }
    // host restatement: the runs of (init_key, 100) followed by the input; all
    // but the last are written (key = the run's last key), the last is the carry
    std::vector<int> rk, rv;
    int ck = init_key, cv = 100;
    for (int i = 0; i < n; ++i) {
      if (keys[i] != ck) {
        rk.push_back(ck);
        rv.push_back(cv);
        cv = vals[i];
      } else {
        cv += vals[i];
      }
      ck = keys[i];
    }
    const rbk_result res = dres.get()[0];
    const bool ok_rbk = res.nout == (long long)rk.size() && res.carry_key == ck &&
                        res.carry_val == cv && dko.get(rk.size()) == rk && dvo.get(rv.size()) == rv;
    CHECK(ok_rbk, variant == 0 ? "group reduce_by_key (carried run continues)"
                               : "group reduce_by_key (carried run written first)");
    if (variant == 0) {
      auto hg = dga.get(), hsc = dsc.get();
      bool ok_g = true, ok_s = true;
      std::vector<int> sc(n, -9);
      for (int i = 0; i < n; ++i) {
        ok_g &= hg[i] == vals[map[i]];
        if (stencil[i]) sc[map[i]] = vals[i];
      }
      for (int i = 0; i < n; ++i) ok_s &= hsc[i] == sc[i];
      CHECK(ok_g, "group gather");
      CHECK(ok_s, "group scatter_if");
    }
    int heads = 0, tails = 0;
    for (int i = 0; i < n; ++i) {
      heads += i == 0 || keys[i] != keys[i - 1];
      tails += i == n - 1 || keys[i] != keys[i + 1];
    }
    const int heads_init = heads - 1 + (init_key != keys[0]);
    auto hfc = dfc.get();
    CHECK(hfc[0] == heads && hfc[1] == tails && hfc[2] == heads_init,
          "head_flags / tail_flags / head_flags_with_init");
  }
  // adjacent_difference (group, several tiles; in place; one agent)
  {
    const int n = 5003, agents = 40;
    std::vector<int> x(n);
    for (int i = 0; i < n; ++i) x[i] = int((i * 2654435761u) % 1999) - 999;
    dvec<int> dx(x), dd(n), din(x), da(16 * agents);
    bulk::async(bulk::con<256, 4>(0), group_adjdiff(), bulk::root, dx.p, n, dd.p, din.p).wait();
    bulk::async(bulk::par(agents), agent_adjdiff(), bulk::root.this_exec, dx.p, da.p).wait();
    auto hd = dd.get(), hin = din.get(), ha = da.get();
    bool ok1 = true, ok2 = true, ok3 = true;
    for (int i = 0; i < n; ++i) {
      ok1 &= hd[i] == x[i] - (i ? x[i - 1] : 100);
      ok2 &= hin[i] == (i ? x[i] - x[i - 1] : x[0]);
    }
    for (int i = 0; i < 16 * agents; ++i) ok3 &= ha[i] == x[i] - (i % 16 ? x[i - 1] : -1);
    CHECK(ok1, "group adjacent_difference with init (5 tiles)");
    CHECK(ok2, "group adjacent_difference in place, no init");
    CHECK(ok3, "agent adjacent_difference");
  }
  // for_each_n: the group adds 1 to all, then the agents add 1 again
  {
    const int n = 3001;
    std::vector<int> x(n);
    for (int i = 0; i < n; ++i) x[i] = i * 3;
    dvec<int> dx(x);
    bulk::async(bulk::con<256, 4>(0), group_for_each(), bulk::root, dx.p, n).wait();
    bulk::async(bulk::par((n + 3) / 4), agent_for_each(), bulk::root.this_exec, dx.p, n).wait();
    auto h = dx.get();
    bool ok = true;
    for (int i = 0; i < n; ++i) ok &= h[i] == i * 3 + 2;
    CHECK(ok, "group and bounded agent for_each_n");
  }
  // join_iterator over two ranges
  {
    std::vector<int> a(1234), b(777);
    for (size_t i = 0; i < a.size(); ++i) a[i] = int(i) * 2 + 1;
    for (size_t i = 0; i < b.size(); ++i) b[i] = -int(i);
    dvec<int> da(a), db(b), dout(a.size() + b.size()), dsum(1);
    bulk::async(bulk::con<256, 4>(0), group_join(), bulk::root, da.p, int(a.size()), db.p,
                int(b.size()), dout.p, dsum.p)
        .wait();
    std::vector<int> ref(a);
    ref.insert(ref.end(), b.begin(), b.end());
    CHECK(dout.get() == ref && dsum.get()[0] == std::accumulate(ref.begin(), ref.end(), 0),
          "join_iterator (copy_n and reduce over two ranges)");
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
    bulkx::exclusive_scan(d.p, d.p + n, r.p, 5, std::plus<int>());
    run = 5;
    for (int i = 0; i < n; ++i) {
      ref[i] = run;
      run += h[i];
    }
    CHECK(r.get() == ref, "bulkx::exclusive_scan");
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

    // the scalar results without the blocking read: a future per call on a stream
    {
      hipStream_t s;
      HIP_OK(hipStreamCreate(&s));
      bulkx::scalar_future<int> f = bulkx::async_reduce(d.p, d.p + n, 13, std::plus<int>(), s);
      auto g = bulkx::async_reduce_by_key(dkk.p, dkk.p + n, dvv.p, ko.p, vo.p,
                                          std::equal_to<int>(), std::plus<int>(), s);
      bulkx::scalar_future<int> f2 = std::move(f);
      CHECK(!f.valid() && f2.valid() && f2.get() == int(sum) && f2.ready(), "bulkx::async_reduce");
      const auto e2 = g.get();
      CHECK(g.ready() && g.num_runs() == m && e2.first - ko.p == m && e2.second - vo.p == m &&
                ko.get(m) == rk && vo.get(m) == rv,
            "bulkx::async_reduce_by_key");
      HIP_OK(hipStreamDestroy(s));
    }
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
