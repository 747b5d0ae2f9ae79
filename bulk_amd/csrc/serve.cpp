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
  double t_arrival = 0, t_launch = 0, t_done = 0;
};

struct Slot {
  bulkx_batch* batch = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  int64_t bytes = 0;
  std::vector<int64_t> reqs;
};

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

// One pass of the worker: retire finished slots, then fill free slots.
// Called with the mutex held.
static void serve_step(bulkx_server* s) {
  for (Slot& sl : s->slots) {
    if (!sl.busy) continue;
    hipError_t q = hipEventQuery(sl.done);
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) s->error = BULKX_ERR_HIP;
    const double t = s->now_us();
    for (int64_t id : sl.reqs) s->reqs[size_t(id)].t_done = t;
    s->completed += int64_t(sl.reqs.size());
    s->admitted_bytes -= sl.bytes;
    sl.reqs.clear();
    sl.bytes = 0;
    sl.busy = false;
    s->cv_done.notify_all();
  }
  const double now = s->now_us();
  for (Slot& sl : s->slots) {
    if (sl.busy || s->pending.empty() || s->error != BULKX_SUCCESS) continue;
    std::vector<bulkx_job> jobs;
    int64_t bytes = 0, elems = 0;
    const int cap = (s->flags & BULKX_SERVE_PER_CALL) ? 1 : s->max_batch_jobs;
    while (!s->pending.empty() && int(jobs.size()) < cap) {
      const Request& r = s->reqs[size_t(s->pending.front())];
      if (r.t_arrival > now) break;  // not arrived yet
      const int64_t e = job_elements(r.job);
      if (s->admitted_bytes + bytes + r.bytes > s->budget) break;  // wait for memory
      if (!jobs.empty() && elems + e > s->max_batch_elements) break;
      jobs.push_back(r.job);
      sl.reqs.push_back(s->pending.front());
      bytes += r.bytes;
      elems += e;
      s->pending.pop_front();
    }
    if (jobs.empty()) break;  // the head request waits for budget
    int st = bulkx_batch_set_jobs(sl.batch, jobs.data(), int(jobs.size()));
    if (st == BULKX_SUCCESS) {
      if (s->flags & BULKX_SERVE_PER_CALL) st = bulkx_batch_launch_per_call(sl.batch);
      else st = bulkx_batch_launch(sl.batch, (s->flags & BULKX_SERVE_NO_GRAPH) ? 0 : 1);
    }
    if (st == BULKX_SUCCESS && hipEventRecord(sl.done, sl.stream) != hipSuccess) st = BULKX_ERR_HIP;
    const double t = s->now_us();
    for (int64_t id : sl.reqs) s->reqs[size_t(id)].t_launch = t;
    if (st != BULKX_SUCCESS) {
      s->error = st;
      // fail the slot's requests so drain() returns
      for (int64_t id : sl.reqs) s->reqs[size_t(id)].t_done = t;
      s->completed += int64_t(sl.reqs.size());
      sl.reqs.clear();
      s->cv_done.notify_all();
      continue;
    }
    sl.busy = true;
    sl.bytes = bytes;
    s->admitted_bytes += bytes;
    if (s->admitted_bytes > s->peak_bytes) s->peak_bytes = s->admitted_bytes;
    s->batches += 1;
  }
}

static void serve_loop(bulkx_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lock(s->mu);
  while (true) {
    serve_step(s);
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
    } else {
      s->cv_work.wait(lock, [&] { return s->stop || !s->pending.empty(); });
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
                       int64_t* request_id) {
  if (!s || !job || bytes < 0) return BULKX_ERR_INVALID_ARGUMENT;
  if (bytes > s->budget) return BULKX_ERR_BUDGET;
  if (job_elements(*job) > s->max_batch_elements) return BULKX_ERR_BUDGET;
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->error != BULKX_SUCCESS) return s->error;
  const double now = s->now_us();
  if (arrival < 0) arrival = now;
  if (!s->pending.empty() && arrival < s->reqs[size_t(s->pending.back())].t_arrival)
    return BULKX_ERR_INVALID_ARGUMENT;  // arrivals must be submitted in order
  Request r;
  r.job = *job;
  r.bytes = bytes;
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
