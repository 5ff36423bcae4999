// Serving loop: a device-resident tree behind a request queue with continuous batching, admission
// under a device-memory budget and update epochs.
//
// Worker threads (one HIP stream each) take every pending read request (within the batch and memory
// limits), coalesce lookups into one batched search and ranges into one batched scan, run them and
// write each request's results back.  While one worker's batch is on the GPU the others keep
// admitting whatever arrived meanwhile, so batches overlap on the device.
//
// Updates: when the oldest pending request is an update, a worker takes every pending update and
// applies them as one epoch (in submission order, so the last write of a key wins).  Reads and
// updates that are pending at the same time are concurrent operations and may be ordered either
// way; a request submitted after another one completed always sees its effect.  The update worker
// stages the batch and prepares the epoch while read batches keep running on the current one (a
// re-layout is built in standby buffers); only the commit -- an enqueue of the in-place merge after
// every earlier search, or a pointer swap -- holds the epoch lock exclusively.  A read batch packs
// its inputs first and then holds that lock shared only while its device work is enqueued, so every
// read sees one whole epoch, which is reported with its result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hm_tree.h"

using namespace hm;
using Clock = std::chrono::steady_clock;

namespace {

#define HM_TRY_RC(expr)       \
  do {                        \
    const int rc_ = (expr);   \
    if (rc_) return rc_;      \
  } while (0)

enum ReqType { REQ_LOOKUP = 0, REQ_RANGE = 1, REQ_UPDATE = 2 };

// The caller's buffers (inputs included) stay valid until the request completes, so a request
// holds pointers only: the keys are copied once, straight into the worker's pinned staging buffer.
struct Request {
  uint64_t id = 0;
  int type = REQ_LOOKUP;
  const uint64_t* keys = nullptr;  // lookup keys / range starts / update keys
  const uint32_t* lens = nullptr;  // range lengths
  const uint64_t* vals = nullptr;  // update values
  uint32_t n = 0;
  uint64_t* out_vals = nullptr;
  uint8_t* out_found = nullptr;
  uint64_t* out_keys = nullptr;
  uint32_t* out_counts = nullptr;
  uint64_t range_total = 0;  // sum of lens
  Clock::time_point t_submit;
  int status = HM_OK;
  uint64_t epoch = 0;
  bool done = false;
};
using ReqPtr = std::shared_ptr<Request>;

// Grow-only host (pinned) and device buffers of one worker.
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  bool pinned = false;
  int ensure(size_t b) {
    if (b <= bytes) return HM_OK;
    const size_t old = bytes;
    release();  // grows geometrically: a hipFree synchronizes the whole device
    size_t want = std::max(std::max(b, 2 * old), (size_t)1 << 20);
    if (pinned) {
      const int rc = host_alloc_mapped(&p, want);
      if (rc) return rc;
    } else {
      HM_TRY(hipMalloc(&p, want));
    }
    bytes = want;
    return HM_OK;
  }
  void release() {
    if (p) (void)(pinned ? hipHostFree(p) : hipFree(p));
    p = nullptr;
    bytes = 0;
  }
};

struct Worker {
  hipStream_t s = nullptr;
  // mapped pinned buffers: the search and range kernels read the packed inputs and write the
  // results here directly, so a batch issues no copy commands at all
  Buf hq{nullptr, 0, true}, hv{nullptr, 0, true}, hf{nullptr, 0, true}, hl{nullptr, 0, true},
      hc{nullptr, 0, true}, hk{nullptr, 0, true};
  Buf doff;  // range offsets stay on the device
  std::thread th;
};

size_t lookup_bytes(uint64_t n) { return n * 17 + pipe_bytes(n, 32, HM_OUT_UNPERMUTE); }
size_t range_bytes(uint64_t n, uint64_t total) { return n * 24 + total * 16; }
size_t request_bytes(const Request& r) {
  if (r.type == REQ_LOOKUP) return lookup_bytes(r.n);
  if (r.type == REQ_RANGE) return range_bytes(r.n, r.range_total);
  return (uint64_t)r.n * 16;
}

}  // namespace

struct hm_server {
  hm_tree* t = nullptr;
  hm_server_params p{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ReqPtr> pending;
  std::unordered_map<uint64_t, ReqPtr> live;
  uint64_t next_id = 1;
  bool stop = false;
  uint64_t inflight_bytes = 0;
  uint64_t outstanding = 0;  // submitted, not yet completed
  uint32_t idle = 0;         // workers waiting for work
  bool updating = false;     // a worker is applying an update batch
  std::shared_mutex epoch_mu;
  std::vector<std::unique_ptr<Worker>> workers;
  // statistics
  std::vector<double> lat_us;
  std::vector<double> lat_kind_us[3];  // by request kind (REQ_LOOKUP, REQ_RANGE, REQ_UPDATE)
  uint64_t n_lookups = 0, n_ranges = 0, n_updates = 0, n_queries = 0, n_rejected = 0, n_launches = 0,
           n_epochs = 0;
  bool have_first = false;
  Clock::time_point first_submit, last_complete;
  double max_read_batch_us = 0, max_update_us = 0, max_prepare_us = 0, max_commit_wait_us = 0;
};

namespace {

// Mark a launch's requests done under one lock acquisition and wake the waiters once.
void complete_all(hm_server* sv, const std::vector<ReqPtr>& rs, const int* status, int status_all, uint64_t epoch) {
  if (rs.empty()) return;
  const auto now = Clock::now();
  {
    std::lock_guard<std::mutex> g(sv->mu);
    for (size_t i = 0; i < rs.size(); ++i) {
      Request& r = *rs[i];
      r.status = status ? status[i] : status_all;
      r.epoch = epoch;
      r.done = true;
      const double us = std::chrono::duration<double, std::micro>(now - r.t_submit).count();
      sv->lat_us.push_back(us);
      sv->lat_kind_us[r.type].push_back(us);
      if (r.type == REQ_LOOKUP) ++sv->n_lookups, sv->n_queries += r.n;
      if (r.type == REQ_RANGE) ++sv->n_ranges, sv->n_queries += r.n;
      if (r.type == REQ_UPDATE) ++sv->n_updates;
    }
    sv->last_complete = now;
    sv->outstanding -= rs.size();
  }
  sv->cv_done.notify_all();
}

uint64_t tree_epoch(hm_tree* t) {
  hm_stats st;
  hm_tree_stats(t, &st);
  return st.epoch;
}

// HM_SERVER_TRACE=1 prints every step slower than 2 ms; HM_SERVER_TRACE=<n> (n > 1) every step
// slower than n microseconds.  Times are since the first traced step.
double trace_threshold_us() {
  static const double th = [] {
    const char* e = std::getenv("HM_SERVER_TRACE");
    if (!e || !*e || *e == '0') return -1.0;
    const double v = std::atof(e);
    return v > 1.0 ? v : 2000.0;
  }();
  return th;
}
bool trace_on() { return trace_threshold_us() > 0; }
void trace_step(const char* what, Clock::time_point a, Clock::time_point b, uint64_t size) {
  static const Clock::time_point t0 = Clock::now();
  const double us = std::chrono::duration<double, std::micro>(b - a).count();
  if (us > trace_threshold_us())
    std::fprintf(stderr, "[hm_server] t=%.3f ms %s took %.0f us (size %llu)\n",
                 std::chrono::duration<double, std::milli>(a - t0).count(), what, us, (unsigned long long)size);
}

// Run one coalesced read batch on worker w.  Called with the epoch lock held shared until the
// device work is enqueued.
int run_reads(hm_server* sv, Worker& w, std::vector<ReqPtr>& lookups, std::vector<ReqPtr>& ranges,
              std::shared_lock<std::shared_mutex>& epoch_lock, uint64_t* epoch) {
  hm_tree* t = sv->t;
  uint64_t nq = 0;
  for (auto& r : lookups) nq += r->n;
  uint64_t nr = 0, rtot = 0;
  for (auto& r : ranges) nr += r->n, rtot += r->range_total;
  int rc = HM_OK;
  const auto t_pack = Clock::now();
  // pack the inputs into pinned staging first: host copies do not hold the epoch lock
  if (nq) {
    if ((rc = w.hq.ensure(nq * 8)) || (rc = w.hv.ensure(nq * 8)) || (rc = w.hf.ensure(nq))) return rc;
    uint64_t o = 0;
    for (auto& r : lookups) {
      std::memcpy((uint64_t*)w.hq.p + o, r->keys, (size_t)r->n * 8);
      o += r->n;
    }
  }
  if (nr) {
    if ((rc = w.hl.ensure(nr * 12)) || (rc = w.hc.ensure(nr * 4)) || (rc = w.hk.ensure(std::max<uint64_t>(rtot, 1) * 16)) ||
        (rc = w.doff.ensure((nr + 1) * 8)))
      return rc;
    uint64_t* hs = (uint64_t*)w.hl.p;
    uint32_t* hlen = (uint32_t*)((char*)w.hl.p + nr * 8);
    uint64_t o = 0;
    for (auto& r : ranges) {
      std::memcpy(hs + o, r->keys, (size_t)r->n * 8);
      std::memcpy(hlen + o, r->lens, (size_t)r->n * 4);
      o += r->n;
    }
  }
  if (trace_on()) trace_step("read pack", t_pack, Clock::now(), nq + rtot);
  // then enqueue the device work under the shared epoch lock: the batch sees one whole epoch
  const auto t_lock = Clock::now();
  epoch_lock.lock();
  if (trace_on()) trace_step("read wait for the epoch lock", t_lock, Clock::now(), nq + rtot);
  const auto t_enq = Clock::now();
  *epoch = tree_epoch(t);
  if (nq) {
    const auto a0 = Clock::now();
    if ((rc = search_impl(t, (const uint64_t*)w.hq.p, nq, (uint64_t*)w.hv.p, (uint8_t*)w.hf.p, nullptr, w.s, true)))
      return rc;
    if (trace_on()) trace_step("  search enqueue", a0, Clock::now(), nq);
  }
  if (nr) {
    const auto a0 = Clock::now();
    uint64_t* hkeys = (uint64_t*)w.hk.p;
    uint64_t* hvals = hkeys + rtot;
    if ((rc = range_impl(t, (const uint64_t*)w.hl.p, (const uint32_t*)((char*)w.hl.p + nr * 8), nr, (uint64_t*)w.doff.p,
                         hkeys, hvals, (uint32_t*)w.hc.p, rtot, w.s, true)))
      return rc;
    if (trace_on()) trace_step("  range enqueue", a0, Clock::now(), nr);
  }
  epoch_lock.unlock();  // the device work is ordered against later publishes from here on
  const auto t_sync = Clock::now();
  HM_TRY(hipStreamSynchronize(w.s));
  if (trace_on()) {
    trace_step("read enqueue (epoch lock held)", t_enq, t_sync, nq + rtot);
    trace_step("read device work", t_sync, Clock::now(), nq + rtot);
  }
  // hand results back
  const auto t_scatter = Clock::now();
  uint64_t o = 0;
  for (auto& r : lookups) {
    const size_t n = r->n;
    std::memcpy(r->out_vals, (uint64_t*)w.hv.p + o, n * 8);
    std::memcpy(r->out_found, (uint8_t*)w.hf.p + o, n);
    o += n;
  }
  uint64_t ro = 0, rk = 0;
  for (auto& r : ranges) {
    const size_t n = r->n;
    std::memcpy(r->out_counts, (uint32_t*)w.hc.p + ro, n * 4);
    std::memcpy(r->out_keys, (uint64_t*)w.hk.p + rk, r->range_total * 8);
    std::memcpy(r->out_vals, (uint64_t*)w.hk.p + rtot + rk, r->range_total * 8);
    ro += n;
    rk += r->range_total;
  }
  if (trace_on()) trace_step("read scatter", t_scatter, Clock::now(), nq + rtot);
  return HM_OK;
}

// Apply a coalesced update batch as one epoch.  Staging and the prepare phase (sort, locate, plan,
// and a full re-layout into the standby buffers when some leaf overflows) run while read batches
// keep going on the current epoch; only the commit (an enqueue or a pointer swap) holds the epoch
// lock exclusively.  A request that cannot be staged (e.g. it holds the reserved key) fails alone;
// status[i] is the outcome of ups[i].
void run_updates(hm_server* sv, Worker& w, std::vector<ReqPtr>& ups, std::vector<int>& status, uint64_t* epoch) {
  hm_tree* t = sv->t;
  status.assign(ups.size(), HM_OK);
  std::lock_guard<std::mutex> pub(t->pub_mu);
  const auto ts = Clock::now();
  uint64_t staged = 0;
  for (size_t i = 0; i < ups.size(); ++i) {
    status[i] = stage_updates(t, ups[i]->keys, ups[i]->vals, ups[i]->n, w.s);
    staged += ups[i]->n;
  }
  PublishPlan plan;
  const auto t0 = Clock::now();
  int rc = publish_prepare(t, w.s, &plan);
  const auto t1 = Clock::now();
  if (trace_on()) {
    trace_step("update staging", ts, t0, staged);
    trace_step(plan.relayout ? "publish prepare (re-layout)" : "publish prepare", t0, t1, staged);
  }
  double wait_us = 0;
  if (!rc) {
    // no new read batch may start while the commit waits for the ones still enqueueing
    std::unique_lock<std::mutex> lk(sv->mu);
    std::unique_lock<std::shared_mutex> ex(sv->epoch_mu);
    wait_us = std::chrono::duration<double, std::micro>(Clock::now() - t1).count();
    if (trace_on()) trace_step("commit wait for the epoch lock", t1, Clock::now(), staged);
    sv->max_prepare_us = std::max(sv->max_prepare_us, std::chrono::duration<double, std::micro>(t1 - t0).count());
    sv->max_commit_wait_us = std::max(sv->max_commit_wait_us, wait_us);
    lk.unlock();
    int rebuilt = 0;
    rc = publish_commit(t, w.s, plan, &rebuilt);
    *epoch = tree_epoch(t);
  }
  // a failed epoch is dropped (its requests are reported failed below), so it is not applied with
  // the next one
  if (rc) discard_staged(t);
  if (!rc && hipStreamSynchronize(w.s) != hipSuccess) rc = set_err(HM_EDEVICE, "update stream failed");
  if (rc)
    for (int& st : status)
      if (!st) st = rc;
}

int reserve_worker(hm_server* sv, Worker& w) {
  const uint64_t b = sv->p.max_batch;
  for (Buf* x : {&w.hq, &w.hv}) HM_TRY_RC(x->ensure(b * 8));
  HM_TRY_RC(w.hf.ensure(b));
  HM_TRY_RC(w.hl.ensure(b * 12));
  HM_TRY_RC(w.hc.ensure(b * 4));
  HM_TRY_RC(w.hk.ensure(b * 16));
  HM_TRY_RC(w.doff.ensure((b + 1) * 8));
  // the tree's workspace for this stream: the largest search pipeline or range scan of a batch
  const size_t ws = std::max(pipe_bytes(b, 32, HM_OUT_UNPERMUTE), range_ws_bytes(b));
  return reserve_workspace(sv->t, w.s, ws);
}

// A backlog is split among idle workers down to batches of this many queries.
constexpr uint64_t kMinShare = 1u << 15;

// The first pending read, or nullptr.
Request* first_read(hm_server* sv) {
  for (auto& r : sv->pending)
    if (r->type != REQ_UPDATE) return r.get();
  return nullptr;
}

void worker_loop(hm_server* sv, Worker* w) {
  (void)hipSetDevice(sv->t->device);
  for (;;) {
    std::vector<ReqPtr> lookups, ranges, ups;
    uint64_t bytes = 0;
    std::shared_lock<std::shared_mutex> epoch_lock(sv->epoch_mu, std::defer_lock);
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      ++sv->idle;
      // one update batch at a time: while it is being applied, the other workers serve reads
      // (pending updates wait for the next epoch instead of parking workers on the publish lock)
      auto update_ready = [&] { return !sv->updating && sv->pending.front()->type == REQ_UPDATE; };
      sv->cv_work.wait(lk, [&] {
        if (sv->pending.empty()) return sv->stop;  // on stop, every pending request is served first
        if (update_ready()) return true;
        if (!sv->p.coalesce) return sv->inflight_bytes == 0 && sv->pending.front()->type != REQ_UPDATE;
        const Request* h = first_read(sv);
        return h && (sv->inflight_bytes == 0 || sv->inflight_bytes + request_bytes(*h) <= sv->p.mem_budget);
      });
      --sv->idle;
      if (sv->pending.empty()) return;  // stopping
      if (!sv->p.coalesce) {  // one request per launch, in arrival order
        ReqPtr h = sv->pending.front();
        sv->pending.pop_front();
        if (h->type == REQ_UPDATE) {
          ups.push_back(h);
          sv->updating = true;
        } else {
          (h->type == REQ_LOOKUP ? lookups : ranges).push_back(h);
          bytes = request_bytes(*h);
          sv->inflight_bytes += bytes;
        }
      } else if (update_ready()) {
        // every pending update, in submission order, as one epoch
        std::deque<ReqPtr> rest;
        for (auto& r : sv->pending) (r->type == REQ_UPDATE ? ups.push_back(r) : rest.push_back(r));
        sv->pending.swap(rest);
        sv->updating = true;
      } else {
        // every pending read that fits the batch and memory limits -- but a backlog is shared with
        // the workers still idle: each worker packs and unpacks its batch on the host, and one
        // worker taking it all would serialize that work (and grow the next backlog)
        uint64_t pending_q = 0;
        for (auto& h : sv->pending)
          if (h->type != REQ_UPDATE) pending_q += h->n + h->range_total;
        const uint64_t share = (pending_q + sv->idle) / (sv->idle + 1);
        const uint64_t cap = std::min<uint64_t>(sv->p.max_batch, std::max<uint64_t>(share, kMinShare));
        uint64_t nq = 0;
        std::deque<ReqPtr> rest;
        bool full = false;
        for (auto& h : sv->pending) {
          if (h->type == REQ_UPDATE || full) {
            rest.push_back(h);
            continue;
          }
          const uint64_t b = request_bytes(*h);
          const uint64_t m = h->n + h->range_total;
          if ((!lookups.empty() || !ranges.empty()) &&
              (nq + m > cap || sv->inflight_bytes + bytes + b > sv->p.mem_budget)) {
            full = true;  // keep arrival order among the reads left behind
            rest.push_back(h);
            continue;
          }
          (h->type == REQ_LOOKUP ? lookups : ranges).push_back(h);
          nq += m;
          bytes += b;
        }
        sv->pending.swap(rest);
        sv->inflight_bytes += bytes;
      }
      ++sv->n_launches;
      if (!sv->pending.empty()) sv->cv_work.notify_one();
    }
    uint64_t epoch = 0;
    const auto tb = Clock::now();
    if (!ups.empty()) {
      std::vector<int> status;
      run_updates(sv, *w, ups, status, &epoch);
      {
        std::lock_guard<std::mutex> g(sv->mu);
        ++sv->n_epochs;
        sv->updating = false;
        sv->max_update_us =
            std::max(sv->max_update_us, std::chrono::duration<double, std::micro>(Clock::now() - tb).count());
      }
      complete_all(sv, ups, status.data(), HM_OK, epoch);
      sv->cv_work.notify_all();
    } else {
      if (trace_on()) {
        Clock::time_point oldest = tb;
        for (auto* v : {&lookups, &ranges})
          for (auto& r : *v) oldest = std::min(oldest, r->t_submit);
        trace_step("queue wait of a batch's oldest request", oldest, tb, lookups.size() + ranges.size());
      }
      const int rc = run_reads(sv, *w, lookups, ranges, epoch_lock, &epoch);
      if (epoch_lock.owns_lock()) epoch_lock.unlock();
      {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->inflight_bytes -= bytes;
        sv->max_read_batch_us =
            std::max(sv->max_read_batch_us, std::chrono::duration<double, std::micro>(Clock::now() - tb).count());
      }
      sv->cv_work.notify_all();
      complete_all(sv, lookups, nullptr, rc, epoch);
      complete_all(sv, ranges, nullptr, rc, epoch);
    }
  }
}

int submit(hm_server* sv, ReqPtr r, uint64_t* ticket) {
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> g(sv->mu);
  if (sv->stop) return set_err(HM_EINVAL, "server is shutting down");
  if (!sv->have_first) {
    sv->first_submit = r->t_submit;
    sv->have_first = true;
  }
  r->id = sv->next_id++;
  if (r->type != REQ_UPDATE && request_bytes(*r) > sv->p.mem_budget) {
    ++sv->n_rejected;
    return set_err(HM_EBUDGET, "request needs %llu device bytes, budget is %llu",
                   (unsigned long long)request_bytes(*r), (unsigned long long)sv->p.mem_budget);
  }
  sv->live[r->id] = r;
  sv->pending.push_back(r);
  ++sv->outstanding;
  if (ticket) *ticket = r->id;
  if (sv->idle) sv->cv_work.notify_one();  // busy workers look at the queue when they finish
  return HM_OK;
}

}  // namespace

extern "C" {

int hm_server_create(hm_tree* t, const hm_server_params* p, hm_server** out) {
  if (!t || !out) return set_err(HM_EINVAL, "NULL argument");
  *out = nullptr;
  hm_server* sv = new hm_server();
  sv->t = t;
  if (p) sv->p = *p;
  if (!sv->p.mem_budget) sv->p.mem_budget = 1ull << 30;
  if (!sv->p.max_batch) sv->p.max_batch = 1u << 20;
  if (!sv->p.max_inflight) sv->p.max_inflight = 4;
  if (sv->p.max_inflight > 16) sv->p.max_inflight = 16;
  for (uint32_t i = 0; i < sv->p.max_inflight; ++i) {
    auto w = std::make_unique<Worker>();
    if (hipStreamCreateWithFlags(&w->s, hipStreamNonBlocking) != hipSuccess) {
      for (auto& x : sv->workers) (void)hipStreamDestroy(x->s);
      delete sv;
      return set_err(HM_EDEVICE, "hipStreamCreate failed");
    }
    sv->workers.push_back(std::move(w));
  }
  // a re-layout while serving builds into standby buffers: allocate them now if the budget allows
  (void)reserve_relayout(t, sv->workers[0]->s);
  // and size the update path for coalesced epochs (a grow frees device memory: a device-wide sync)
  (void)reserve_updates(t, std::max<size_t>(sv->p.max_batch, 1u << 17), sv->workers[0]->s);
  // size every worker's buffers for a full batch now: a grow while serving would allocate pinned
  // memory (slow) and free device memory (synchronizes the device), stalling every request
  for (auto& w : sv->workers) {
    const int rc = reserve_worker(sv, *w);
    if (rc) {
      hm_server_destroy(sv);
      return rc;
    }
  }
  // run a one-key lookup and a one-key range now: the first launch of a kernel in a process loads
  // the code object (milliseconds), which would otherwise land on the first requests
  {
    Worker& w = *sv->workers[0];
    uint64_t* hq = (uint64_t*)w.hq.p;
    uint32_t* hlen = (uint32_t*)((char*)w.hl.p + 8);
    hq[0] = 0;
    ((uint64_t*)w.hl.p)[0] = 0;
    hlen[0] = 1;
    int rc = search_impl(t, hq, 1, (uint64_t*)w.hv.p, (uint8_t*)w.hf.p, nullptr, w.s, true);
    if (!rc)
      rc = range_impl(t, (const uint64_t*)w.hl.p, hlen, 1, (uint64_t*)w.doff.p, (uint64_t*)w.hk.p,
                      (uint64_t*)w.hk.p + 1, (uint32_t*)w.hc.p, 1, w.s, true);
    if (!rc && hipStreamSynchronize(w.s) != hipSuccess) rc = set_err(HM_EDEVICE, "warm-up failed");
    if (rc) {
      hm_server_destroy(sv);
      return rc;
    }
  }
  for (auto& w : sv->workers) w->th = std::thread(worker_loop, sv, w.get());
  *out = sv;
  return HM_OK;
}

int hm_server_lookup(hm_server* s, const uint64_t* keys, uint32_t n, uint64_t* out_vals, uint8_t* out_found,
                     uint64_t* ticket) {
  if (!s || (n && (!keys || !out_vals || !out_found))) return set_err(HM_EINVAL, "NULL argument");
  auto r = std::make_shared<Request>();
  r->type = REQ_LOOKUP;
  r->keys = keys;
  r->n = n;
  r->out_vals = out_vals;
  r->out_found = out_found;
  return submit(s, r, ticket);
}

int hm_server_range(hm_server* s, const uint64_t* starts, const uint32_t* lens, uint32_t n, uint64_t* out_keys,
                    uint64_t* out_vals, uint32_t* out_counts, uint64_t capacity, uint64_t* ticket) {
  if (!s || (n && (!starts || !lens || !out_keys || !out_vals || !out_counts)))
    return set_err(HM_EINVAL, "NULL argument");
  auto r = std::make_shared<Request>();
  r->type = REQ_RANGE;
  r->keys = starts;
  r->lens = lens;
  r->n = n;
  for (uint32_t i = 0; i < n; ++i) r->range_total += lens[i];
  if (r->range_total > capacity) return set_err(HM_EINVAL, "capacity %llu < sum of lengths %llu",
                                                (unsigned long long)capacity, (unsigned long long)r->range_total);
  r->out_keys = out_keys;
  r->out_vals = out_vals;
  r->out_counts = out_counts;
  return submit(s, r, ticket);
}

int hm_server_update(hm_server* s, const uint64_t* keys, const uint64_t* vals, uint32_t n, uint64_t* ticket) {
  if (!s || (n && (!keys || !vals))) return set_err(HM_EINVAL, "NULL argument");
  auto r = std::make_shared<Request>();
  r->type = REQ_UPDATE;
  r->keys = keys;
  r->vals = vals;
  r->n = n;
  return submit(s, r, ticket);
}

int hm_server_wait(hm_server* s, uint64_t ticket, int* status, uint64_t* epoch) {
  if (!s) return set_err(HM_EINVAL, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) return set_err(HM_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
  ReqPtr r = it->second;
  s->cv_done.wait(lk, [&] { return r->done; });
  if (status) *status = r->status;
  if (epoch) *epoch = r->epoch;
  s->live.erase(ticket);
  return HM_OK;
}

int hm_server_drain(hm_server* s) {
  if (!s) return set_err(HM_EINVAL, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
  return HM_OK;
}

int hm_server_get_stats(hm_server* s, hm_server_stats* st) {
  if (!s || !st) return set_err(HM_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> g(s->mu);
  std::memset(st, 0, sizeof(*st));
  st->completed = s->lat_us.size();
  st->lookups = s->n_lookups;
  st->ranges = s->n_ranges;
  st->updates = s->n_updates;
  st->queries = s->n_queries;
  st->rejected = s->n_rejected;
  st->launches = s->n_launches;
  st->epochs = s->n_epochs;
  if (!s->lat_us.empty()) {
    std::vector<double> v = s->lat_us;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) { return v[std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5))]; };
    st->p50_us = pct(0.50);
    st->p99_us = pct(0.99);
    double sum = 0;
    for (double x : v) sum += x;
    st->mean_us = sum / (double)v.size();
    st->max_us = v.back();
    st->span_s = std::chrono::duration<double>(s->last_complete - s->first_submit).count();
  }
  double* p99_kind[3] = {&st->p99_lookup_us, &st->p99_range_us, &st->p99_update_us};
  for (int k = 0; k < 3; ++k) {
    *p99_kind[k] = 0;
    if (s->lat_kind_us[k].empty()) continue;
    std::vector<double> v = s->lat_kind_us[k];
    std::sort(v.begin(), v.end());
    *p99_kind[k] = v[std::min(v.size() - 1, (size_t)(0.99 * (double)(v.size() - 1) + 0.5))];
  }
  st->max_read_batch_us = s->max_read_batch_us;
  st->max_update_us = s->max_update_us;
  st->max_prepare_us = s->max_prepare_us;
  st->max_commit_wait_us = s->max_commit_wait_us;
  return HM_OK;
}

int hm_server_reset_stats(hm_server* s) {
  if (!s) return set_err(HM_EINVAL, "server is NULL");
  std::lock_guard<std::mutex> g(s->mu);
  s->lat_us.clear();
  for (auto& v : s->lat_kind_us) v.clear();
  s->n_lookups = s->n_ranges = s->n_updates = s->n_queries = s->n_rejected = s->n_launches = s->n_epochs = 0;
  s->max_read_batch_us = s->max_update_us = s->max_prepare_us = s->max_commit_wait_us = 0;
  s->have_first = false;
  return HM_OK;
}

int hm_server_destroy(hm_server* s) {
  if (!s) return set_err(HM_EINVAL, "server is NULL");
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& w : s->workers)
    if (w->th.joinable()) w->th.join();
  for (auto& w : s->workers) {
    (void)hipStreamSynchronize(w->s);
    for (Buf* b : {&w->hq, &w->hv, &w->hf, &w->hl, &w->hc, &w->hk, &w->doff})
      b->release();
    (void)hipStreamDestroy(w->s);
  }
  delete s;
  return HM_OK;
}

}  // extern "C"
