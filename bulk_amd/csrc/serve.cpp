// Serving loop: continuous batching of device-resident jobs under an HBM
// budget.
//
// submit() only records the request (arrival time, device footprint) in a
// FIFO queue.  A worker thread owns `max_inflight` slots, each a batch
// executor on a stream of its own.  Whenever a slot is free it takes the
// requests at the head of the queue that fit the slot's capacity and the
// remaining budget (admitted-but-unfinished bytes <= budget), uploads them
// and replays the slot's graph; requests that arrive while earlier batches
// run are therefore admitted as slots and memory free up, and the batches
// of different slots overlap on the device.  Completion is detected with a
// hipEvent per slot; every request gets arrival, launch and completion
// timestamps (microseconds on one steady clock).  In per-call mode every
// request is launched alone through the plain entry points.
//
// Host-staged requests (bulkx_server_submit_host) name host buffers: when
// such a request is admitted, the slot allocates its device buffers
// (stream-ordered), copies its inputs in ahead of the batch and its outputs
// back behind it, and frees the buffers, all on the slot's stream; the copies
// of one slot overlap the batches of the others.  Its footprint charged to
// the budget is the size of those device buffers.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "host_common.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  bulkx_job job;
  int64_t bytes = 0;
  bool host = false;  // job names host buffers, staged through the device
  double t_arrival = 0, t_launch = 0, t_done = 0;
};

struct Copy {
  void* dst;
  const void* src;
  size_t bytes;
};

struct Slot {
  bulkx_batch* batch = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  int64_t bytes = 0;
  std::vector<int64_t> reqs;
  std::vector<void*> staged;  // device buffers of the slot's host-staged requests
  std::vector<Copy> back;     // their device -> host copies, after the batch
};

// The device buffers of a host-staged job, in bulkx_job order: for each of
// in0..in3 and out0..out2, its size in bytes (0 when the job has none) and
// whether it is copied in (input, or sorted in place) and back (output).
struct Staging {
  size_t bytes[7] = {0, 0, 0, 0, 0, 0, 0};
  bool in[7] = {false, false, false, false, false, false, false};
  bool out[7] = {false, false, false, false, false, false, false};
  size_t total = 0;
};

constexpr size_t kStageAlign = 256;

static bool staging_of(const bulkx_job& j, Staging* g) {
  const size_t ks = bulkx_dtype_size(j.dtype);
  const size_t vs = j.value_dtype >= 0 ? bulkx_dtype_size(j.value_dtype) : 0;
  if (ks == 0 || j.n0 < 0 || j.n1 < 0) return false;
  const size_t n0 = size_t(j.n0), n1 = size_t(j.n1);
  auto set = [&](int i, size_t b, bool in, bool out) {
    g->bytes[i] = b;
    g->in[i] = in;
    g->out[i] = out;
  };
  switch (j.kind) {
    case BULKX_JOB_INCLUSIVE_SCAN:
      set(0, n0 * ks, true, false);
      set(4, n0 * ks, false, true);
      break;
    case BULKX_JOB_REDUCE:
      set(0, n0 * ks, true, false);
      set(4, ks, false, true);
      break;
    case BULKX_JOB_REDUCE_BY_KEY:  // outputs sized for n0 runs, copied back whole
      if (vs == 0) return false;
      set(0, n0 * ks, true, false);
      set(1, n0 * vs, true, false);
      set(4, n0 * ks, false, true);
      set(5, n0 * vs, false, true);
      set(6, sizeof(int64_t), false, true);
      break;
    case BULKX_JOB_MERGE_BY_KEY:
      set(0, n0 * ks, true, false);
      set(2, n1 * ks, true, false);
      set(4, (n0 + n1) * ks, false, true);
      if (j.in1 || j.in3 || j.out1) {
        if (vs == 0) return false;
        set(1, n0 * vs, true, false);
        set(3, n1 * vs, true, false);
        set(5, (n0 + n1) * vs, false, true);
      }
      break;
    case BULKX_JOB_SORT_BY_KEY:  // in place
      set(4, n0 * ks, true, true);
      if (j.out1) {
        if (vs == 0) return false;
        set(5, n0 * vs, true, true);
      }
      break;
    default:
      return false;
  }
  g->total = 0;
  for (size_t b : g->bytes) g->total += (b + kStageAlign - 1) / kStageAlign * kStageAlign;
  if (g->total == 0) g->total = kStageAlign;
  return true;
}

static const void* job_ptr(const bulkx_job& j, int i) {
  const void* p[7] = {j.in0, j.in1, j.in2, j.in3, j.out0, j.out1, j.out2};
  return p[i];
}

// Allocates the device buffers of host job `hj` on `stream`, enqueues the
// copies of its inputs, and returns the device job in *dj; the copies back
// are appended to `back` (enqueued after the batch).
static int stage_in(const bulkx_job& hj, hipStream_t stream, bulkx_job* dj,
                    std::vector<void*>& staged, std::vector<Copy>& back) {
  Staging g;
  if (!staging_of(hj, &g)) return BULKX_ERR_INVALID_ARGUMENT;
  char* base = nullptr;
  hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&base), g.total, stream);
  if (e != hipSuccess) {
    bulkx::set_last_error("host-staged request", e);
    return BULKX_ERR_OUT_OF_MEMORY;
  }
  staged.push_back(base);
  void* dp[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  size_t off = 0;
  for (int i = 0; i < 7; ++i) {
    if (g.bytes[i] == 0) continue;
    const void* hp = job_ptr(hj, i);
    if (hp == nullptr) return BULKX_ERR_INVALID_ARGUMENT;
    dp[i] = base + off;
    off += (g.bytes[i] + kStageAlign - 1) / kStageAlign * kStageAlign;
    if (g.in[i]) {
      e = hipMemcpyAsync(dp[i], hp, g.bytes[i], hipMemcpyHostToDevice, stream);
      if (e != hipSuccess) {
        bulkx::set_last_error("host-staged request", e);
        return BULKX_ERR_HIP;
      }
    }
    if (g.out[i]) back.push_back(Copy{const_cast<void*>(hp), dp[i], g.bytes[i]});
  }
  *dj = hj;
  dj->in0 = dp[0];
  dj->in1 = dp[1];
  dj->in2 = dp[2];
  dj->in3 = dp[3];
  dj->out0 = dp[4];
  dj->out1 = dp[5];
  dj->out2 = dp[6];
  return BULKX_SUCCESS;
}

}  // namespace

struct bulkx_server {
  int64_t budget = 0;
  int max_batch_jobs = 0;
  int64_t max_batch_elements = 0;
  int flags = 0;
  int device = 0;
  std::vector<Slot> slots;
  std::deque<Request> reqs;  // stable references; index = request id
  std::deque<int64_t> pending;
  int64_t admitted_bytes = 0, peak_bytes = 0;
  int64_t batches = 0, completed = 0;
  int error = BULKX_SUCCESS;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread worker;
  bool stop = false;
  Clock::time_point t0;

  double now_us() const {
    return std::chrono::duration<double, std::micro>(Clock::now() - t0).count();
  }
};

static int64_t job_elements(const bulkx_job& j) {
  return j.n0 + (j.kind == BULKX_JOB_MERGE_BY_KEY ? j.n1 : 0);
}

// Fails every request still queued (after an error): each counts as
// completed now, so drain() returns the error instead of waiting forever.
// Called with the mutex held.
static void fail_pending(bulkx_server* s) {
  const double t = s->now_us();
  for (int64_t id : s->pending) s->reqs[size_t(id)].t_done = t;
  s->completed += int64_t(s->pending.size());
  s->pending.clear();
  s->cv_done.notify_all();
}

static void set_error(bulkx_server* s, int st) {
  if (s->error == BULKX_SUCCESS) s->error = st;
  fail_pending(s);
}

// Enqueues one slot's batch on its stream: staging of host requests, the
// job list, the launch, the copies back and the completion event.  Runs
// without the server mutex (only the worker touches a slot).
static int launch_slot(bulkx_server* s, Slot& sl, const std::vector<Request>& taken) {
  std::vector<bulkx_job> jobs;
  jobs.reserve(taken.size());
  int st = BULKX_SUCCESS;
  for (const Request& r : taken) {
    if (r.host) {
      bulkx_job dj;
      if (st == BULKX_SUCCESS) st = stage_in(r.job, sl.stream, &dj, sl.staged, sl.back);
      jobs.push_back(st == BULKX_SUCCESS ? dj : r.job);
    } else {
      jobs.push_back(r.job);
    }
  }
  if (st == BULKX_SUCCESS) st = bulkx_batch_set_jobs(sl.batch, jobs.data(), int(jobs.size()));
  if (st == BULKX_SUCCESS) {
    if (s->flags & BULKX_SERVE_PER_CALL) st = bulkx_batch_launch_per_call(sl.batch);
    else st = bulkx_batch_launch(sl.batch, (s->flags & BULKX_SERVE_NO_GRAPH) ? 0 : 1);
  }
  // host-staged results back, then their device buffers released, in stream order
  for (const Copy& c : sl.back) {
    if (st != BULKX_SUCCESS) break;
    const hipError_t e = hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, sl.stream);
    if (e != hipSuccess) {
      bulkx::set_last_error("host-staged request", e);
      st = BULKX_ERR_HIP;
    }
  }
  for (void* p : sl.staged) (void)hipFreeAsync(p, sl.stream);
  sl.staged.clear();
  sl.back.clear();
  if (st == BULKX_SUCCESS && hipEventRecord(sl.done, sl.stream) != hipSuccess) st = BULKX_ERR_HIP;
  return st;
}

// One pass of the worker: retire finished slots, then fill free slots.
// Called with the mutex held; releases it while a slot's batch is enqueued.
static void serve_step(bulkx_server* s, std::unique_lock<std::mutex>& lock) {
  for (Slot& sl : s->slots) {
    if (!sl.busy) continue;
    hipError_t q = hipEventQuery(sl.done);
    if (q == hipErrorNotReady) continue;
    int st = q == hipSuccess ? BULKX_SUCCESS : BULKX_ERR_HIP;
    // a look-back that gave up inside the batch makes its results wrong
    if (st == BULKX_SUCCESS) st = bulkx_batch_status(sl.batch);
    const double t = s->now_us();
    for (int64_t id : sl.reqs) s->reqs[size_t(id)].t_done = t;
    s->completed += int64_t(sl.reqs.size());
    s->admitted_bytes -= sl.bytes;
    sl.reqs.clear();
    sl.bytes = 0;
    sl.busy = false;
    if (st != BULKX_SUCCESS) set_error(s, st);
    s->cv_done.notify_all();
  }
  for (Slot& sl : s->slots) {
    if (sl.busy || s->pending.empty() || s->error != BULKX_SUCCESS) continue;
    const double now = s->now_us();
    std::vector<Request> taken;
    int64_t bytes = 0, elems = 0;
    const int cap = (s->flags & BULKX_SERVE_PER_CALL) ? 1 : s->max_batch_jobs;
    while (!s->pending.empty() && int(taken.size()) < cap) {
      const Request& r = s->reqs[size_t(s->pending.front())];
      if (r.t_arrival > now) break;  // not arrived yet
      const int64_t e = job_elements(r.job);
      if (s->admitted_bytes + bytes + r.bytes > s->budget) break;  // wait for memory
      if (!taken.empty() && elems + e > s->max_batch_elements) break;
      taken.push_back(r);
      sl.reqs.push_back(s->pending.front());
      bytes += r.bytes;
      elems += e;
      s->pending.pop_front();
    }
    if (taken.empty()) break;  // the head request waits for budget
    // charge the budget before the lock is released
    s->admitted_bytes += bytes;
    if (s->admitted_bytes > s->peak_bytes) s->peak_bytes = s->admitted_bytes;
    lock.unlock();
    const int st = launch_slot(s, sl, taken);
    lock.lock();
    const double t = s->now_us();
    for (int64_t id : sl.reqs) s->reqs[size_t(id)].t_launch = t;
    if (st != BULKX_SUCCESS) {
      // fail the slot's requests and everything queued, so drain() returns
      s->admitted_bytes -= bytes;
      for (int64_t id : sl.reqs) s->reqs[size_t(id)].t_done = t;
      s->completed += int64_t(sl.reqs.size());
      sl.reqs.clear();
      set_error(s, st);
      continue;
    }
    sl.busy = true;
    sl.bytes = bytes;
    s->batches += 1;
  }
}

static void serve_loop(bulkx_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lock(s->mu);
  while (true) {
    serve_step(s, lock);
    bool any_busy = false;
    for (const Slot& sl : s->slots) any_busy |= sl.busy;
    if (s->stop && s->pending.empty() && !any_busy) break;
    const bool waiting_arrival =
        !s->pending.empty() && s->reqs[size_t(s->pending.front())].t_arrival > s->now_us();
    if (any_busy || waiting_arrival) {
      // poll completions at ~10 us granularity while work is on the device
      lock.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(10));
      lock.lock();
    } else if (s->pending.empty() || s->error != BULKX_SUCCESS) {
      if (s->error != BULKX_SUCCESS) fail_pending(s);
      s->cv_work.wait(lock, [&] { return s->stop || !s->pending.empty(); });
    } else {
      // requests have arrived but none fits yet: nothing is on the device,
      // so the head request needs more than the free budget can ever give
      // (submit() rejects such requests; this keeps the loop from spinning)
      set_error(s, BULKX_ERR_BUDGET);
    }
  }
}

using namespace bulkx;

extern "C" int bulkx_server_create(int64_t hbm_budget_bytes, int max_inflight, int max_batch_jobs,
                                   int64_t max_batch_elements, int flags, bulkx_server** out) {
  if (!out || hbm_budget_bytes <= 0 || max_inflight <= 0 || max_batch_jobs <= 0 ||
      max_batch_elements <= 0)
    return BULKX_ERR_INVALID_ARGUMENT;
  DeviceProps props;
  BULKX_TRY(current_device_props(&props));
  bulkx_server* s = new bulkx_server();
  s->budget = hbm_budget_bytes;
  s->max_batch_jobs = max_batch_jobs;
  s->max_batch_elements = max_batch_elements;
  s->flags = flags;
  s->device = props.device;
  s->t0 = Clock::now();
  s->slots.resize(size_t(max_inflight));
  for (Slot& sl : s->slots) {
    hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    int st = e == hipSuccess ? BULKX_SUCCESS : BULKX_ERR_HIP;
    if (st == BULKX_SUCCESS)
      st = bulkx_batch_create(max_batch_jobs, max_batch_elements, sl.stream, &sl.batch);
    // scratch for the largest batch now: serving never allocates or frees it
    if (st == BULKX_SUCCESS) st = bulkx_batch_reserve(sl.batch, max_batch_elements);
    // graphs of every merge-round count up front: no request waits on a capture
    if (st == BULKX_SUCCESS && !(flags & (BULKX_SERVE_PER_CALL | BULKX_SERVE_NO_GRAPH)))
      st = bulkx_batch_prepare_graphs(sl.batch);
    if (st != BULKX_SUCCESS) {
      if (e != hipSuccess) set_last_error("server slot", e);
      bulkx_server_destroy(s);
      return st;
    }
  }
  s->worker = std::thread(serve_loop, s);
  *out = s;
  return BULKX_SUCCESS;
}

extern "C" int bulkx_server_destroy(bulkx_server* s) {
  if (!s) return BULKX_SUCCESS;
  if (s->worker.joinable()) {
    {
      std::lock_guard<std::mutex> lock(s->mu);
      s->stop = true;
    }
    s->cv_work.notify_all();
    s->worker.join();
  }
  for (Slot& sl : s->slots) {
    if (sl.batch) bulkx_batch_destroy(sl.batch);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  delete s;
  return BULKX_SUCCESS;
}

static int submit_impl(bulkx_server* s, const bulkx_job* job, int64_t bytes, double arrival,
                       int64_t* request_id, bool host = false) {
  if (!s || !job || bytes < 0) return BULKX_ERR_INVALID_ARGUMENT;
  if (bytes > s->budget) return BULKX_ERR_BUDGET;
  if (job_elements(*job) > s->max_batch_elements) return BULKX_ERR_BUDGET;
  // a job the batch cannot run is refused here, never queued
  BULKX_TRY(batch_check_job(*job));
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->error != BULKX_SUCCESS) return s->error;
  const double now = s->now_us();
  if (arrival < 0) arrival = now;
  if (!s->pending.empty() && arrival < s->reqs[size_t(s->pending.back())].t_arrival)
    return BULKX_ERR_INVALID_ARGUMENT;  // arrivals must be submitted in order
  Request r;
  r.job = *job;
  r.bytes = bytes;
  r.host = host;
  r.t_arrival = arrival;
  const int64_t id = int64_t(s->reqs.size());
  s->reqs.push_back(r);
  s->pending.push_back(id);
  if (request_id) *request_id = id;
  s->cv_work.notify_one();
  return BULKX_SUCCESS;
}

extern "C" int bulkx_server_submit(bulkx_server* s, const bulkx_job* job, int64_t bytes,
                                   int64_t* request_id) {
  return submit_impl(s, job, bytes, -1.0, request_id);
}

extern "C" int bulkx_server_submit_at(bulkx_server* s, const bulkx_job* job, int64_t bytes,
                                      double arrival_us, int64_t* request_id) {
  if (arrival_us < 0) return BULKX_ERR_INVALID_ARGUMENT;
  return submit_impl(s, job, bytes, arrival_us, request_id);
}

extern "C" int bulkx_server_submit_host(bulkx_server* s, const bulkx_job* job, double arrival_us,
                                        int64_t* request_id, int64_t* device_bytes) {
  if (!s || !job) return BULKX_ERR_INVALID_ARGUMENT;
  Staging g;
  if (!staging_of(*job, &g)) return BULKX_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 7; ++i)
    if (g.bytes[i] != 0 && job_ptr(*job, i) == nullptr) return BULKX_ERR_INVALID_ARGUMENT;
  if (device_bytes) *device_bytes = int64_t(g.total);
  return submit_impl(s, job, int64_t(g.total), arrival_us < 0 ? -1.0 : arrival_us, request_id,
                     true);
}

extern "C" double bulkx_server_now_us(bulkx_server* s) { return s ? s->now_us() : 0.0; }

extern "C" int bulkx_server_drain(bulkx_server* s) {
  if (!s) return BULKX_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lock(s->mu);
  s->cv_done.wait(lock, [&] { return s->completed == int64_t(s->reqs.size()); });
  return s->error;
}

extern "C" int bulkx_server_request_times(bulkx_server* s, int64_t id, double* t_arrival,
                                          double* t_launch, double* t_done) {
  if (!s) return BULKX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(s->mu);
  if (id < 0 || id >= int64_t(s->reqs.size())) return BULKX_ERR_INVALID_ARGUMENT;
  const Request& r = s->reqs[size_t(id)];
  if (t_arrival) *t_arrival = r.t_arrival;
  if (t_launch) *t_launch = r.t_launch;
  if (t_done) *t_done = r.t_done;
  return BULKX_SUCCESS;
}

extern "C" int bulkx_server_stats(bulkx_server* s, int64_t* batches, int64_t* completed,
                                  int64_t* peak_bytes) {
  if (!s) return BULKX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(s->mu);
  if (batches) *batches = s->batches;
  if (completed) *completed = s->completed;
  if (peak_bytes) *peak_bytes = s->peak_bytes;
  return BULKX_SUCCESS;
}
