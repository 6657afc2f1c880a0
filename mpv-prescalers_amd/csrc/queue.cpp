// The request queue of include/mpx.h (mpx_queue_create / mpx_submit / mpx_poll):
// continuous batching of upscale requests over one context, in the library.
//
// Client threads submit requests (device pointers + variant + sizes); a scheduler thread
// takes what is queued, in arrival order, groups it by plan and launches each group (up to
// max_batch frames at independent addresses) on the least loaded of the queue's streams;
// a completion thread waits for each launch's event, in launch order, and finishes its
// requests.  Everything the threads share is guarded by one mutex, and every change of
// state is announced on one condition variable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "internal.h"

namespace {

using Clock = std::chrono::steady_clock;
constexpr int kLaunchesPerStream = 2;   // in flight per stream: later arrivals batch meanwhile

struct Key {
  std::string name;
  int in_w, in_h, tw, th, in_type, bits, out_type, precision, sized;
  bool operator<(const Key& o) const {
    return std::tie(name, in_w, in_h, tw, th, in_type, bits, out_type, precision, sized) <
           std::tie(o.name, o.in_w, o.in_h, o.tw, o.th, o.in_type, o.bits, o.out_type, o.precision, o.sized);
  }
};

struct Req {
  uint64_t ticket = 0;
  std::string name;
  mpx_request r{};
  Key key;
  int out_w = 0;
  hipEvent_t ready = nullptr;  // recorded on the caller's stream at submit
  Clock::time_point arrival;
};

struct PlanSlot {
  mpx_plan* plan = nullptr;
  int inflight = 0;        // launches of this plan not finished yet
  uint64_t last_use = 0;
};

struct Launch {
  hipEvent_t done = nullptr;
  int stream = 0;
  std::pair<Key, int> plan_key;
  std::vector<uint64_t> tickets;
  std::vector<hipEvent_t> ready;  // returned to the pool when the launch finished
};

int default_scale(const std::string& name) { return name.rfind("ravu-3x", 0) == 0 ? 3 : 2; }

}  // namespace

struct mpx_queue {
  mpx_ctx* ctx = nullptr;
  int max_batch = 0, max_w = 0, max_h = 0;
  std::vector<hipStream_t> streams;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::unique_ptr<Req>> queued;
  std::deque<Launch> inflight;
  std::vector<int> load;                      // launches in flight per stream
  std::map<uint64_t, int> status;             // MPX_PENDING, MPX_OK or the request's error
  std::map<uint64_t, std::string> errors;
  std::map<uint64_t, Clock::time_point> arrival;   // submit time of unfinished tickets
  std::map<uint64_t, int64_t> latency;        // submit -> finish, microseconds
  std::vector<hipEvent_t> events;             // free events
  std::map<std::pair<Key, int>, PlanSlot> plans;
  uint64_t next_ticket = 1, launches = 0, clock = 0;
  bool stop = false, sched_done = false, held = false;
  int broken = 0;
  std::string broken_msg;
  std::thread scheduler, completer;

  // ---- helpers (mu held) -------------------------------------------------------------
  hipEvent_t take_event() {
    if (!events.empty()) {
      hipEvent_t e = events.back();
      events.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
  }
  void finish(uint64_t t, int st, const std::string& msg) {
    status[t] = st;
    if (st < 0) errors[t] = msg;
    auto a = arrival.find(t);
    if (a != arrival.end()) {
      latency[t] = std::chrono::duration_cast<std::chrono::microseconds>(Clock::now() - a->second).count();
      arrival.erase(a);
    }
  }
  void break_queue(const std::string& msg) {
    if (!broken) {
      broken = MPX_ERR_HIP;
      broken_msg = msg;
    }
    for (auto& r : queued) {
      finish(r->ticket, MPX_ERR_HIP, "the queue stopped after a device error: " + broken_msg);
      if (r->ready) events.push_back(r->ready);
    }
    queued.clear();
  }

  void schedule();
  void complete();
  void launch_group(std::vector<std::unique_ptr<Req>>& grp, std::unique_lock<std::mutex>& lk);
  mpx_plan* plan_for(const std::pair<Key, int>& k, std::unique_lock<std::mutex>& lk, int* rc, std::string* msg);
};

// The plan of (key, stream), created on first use; when the context's budget is spent,
// idle plans are freed, least recently used first (waiting for launches if none is idle).
mpx_plan* mpx_queue::plan_for(const std::pair<Key, int>& k, std::unique_lock<std::mutex>& lk, int* rc,
                              std::string* msg) {
  auto it = plans.find(k);
  if (it != plans.end()) {
    it->second.last_use = ++clock;
    return it->second.plan;
  }
  const Key& key = k.first;
  for (;;) {
    mpx_plan* p = nullptr;
    lk.unlock();
    const int c = mpx_plan_create(ctx, key.name.c_str(), key.in_w, key.in_h, key.tw, key.th, key.in_type, key.bits,
                                  key.out_type, key.precision, max_batch, &p);
    const std::string err = c < 0 ? mpx_host::last_error() : std::string();
    lk.lock();
    if (c >= 0) {
      plans[k] = PlanSlot{p, 0, ++clock};
      return p;
    }
    if (c != MPX_ERR_BUDGET) {
      *rc = c;
      *msg = err;
      return nullptr;
    }
    auto victim = plans.end();
    for (auto pi = plans.begin(); pi != plans.end(); ++pi)
      if (pi->second.inflight == 0 && (victim == plans.end() || pi->second.last_use < victim->second.last_use))
        victim = pi;
    if (victim != plans.end()) {
      mpx_plan_destroy(victim->second.plan);
      plans.erase(victim);
      continue;
    }
    if (inflight.empty()) {  // nothing to wait for and nothing to free
      *rc = c;
      *msg = err;
      return nullptr;
    }
    const uint64_t seen = launches;
    const size_t n = inflight.size();
    cv.wait(lk, [&] { return inflight.size() < n || launches != seen || broken; });
    if (broken) {
      *rc = MPX_ERR_HIP;
      *msg = broken_msg;
      return nullptr;
    }
  }
}

void mpx_queue::launch_group(std::vector<std::unique_ptr<Req>>& grp, std::unique_lock<std::mutex>& lk) {
  // the least loaded stream with a free slot (waiting for one)
  int s = -1;
  cv.wait(lk, [&] {
    if (broken) return true;
    for (size_t i = 0; i < load.size(); ++i)
      if (load[i] < kLaunchesPerStream && (s < 0 || load[i] < load[s])) s = static_cast<int>(i);
    return s >= 0;
  });
  auto fail_all = [&](int rc, const std::string& msg) {
    for (auto& r : grp) {
      finish(r->ticket, rc, msg);
      if (r->ready) events.push_back(r->ready);
    }
    cv.notify_all();
  };
  if (broken) return fail_all(MPX_ERR_HIP, "the queue stopped after a device error: " + broken_msg);
  const std::pair<Key, int> pk(grp[0]->key, s);
  int rc = MPX_OK;
  std::string msg;
  mpx_plan* plan = plan_for(pk, lk, &rc, &msg);
  if (!plan) return fail_all(rc, msg);
  const int n = static_cast<int>(grp.size());
  std::vector<const void*> ins(n);
  std::vector<void*> outs(n);
  std::vector<int> ws(n), hs(n);
  std::vector<int64_t> ips(n), ops(n);
  Launch L;
  L.stream = s;
  L.plan_key = pk;
  for (int i = 0; i < n; ++i) {
    const mpx_request& r = grp[i]->r;
    ins[i] = r.in;
    outs[i] = r.out;
    ws[i] = r.in_w;
    hs[i] = r.in_h;
    ips[i] = r.in_w;
    ops[i] = grp[i]->out_w;
    L.tickets.push_back(grp[i]->ticket);
    if (grp[i]->ready) L.ready.push_back(grp[i]->ready);
  }
  hipEvent_t done = take_event();
  PlanSlot& slot = plans[pk];
  ++slot.inflight;
  ++load[s];
  lk.unlock();
  hipStream_t st = streams[s];
  std::string herr;
  int lrc = MPX_OK;
  for (hipEvent_t e : L.ready)
    if (hipStreamWaitEvent(st, e, 0) != hipSuccess) {
      lrc = MPX_ERR_HIP;
      herr = "hipStreamWaitEvent failed";
    }
  if (lrc == MPX_OK) {
    lrc = pk.first.sized ? mpx_run_frames_sized(plan, n, ins.data(), outs.data(), ws.data(), hs.data(), ips.data(),
                                                ops.data(), st)
                         : mpx_run_frames(plan, n, ins.data(), outs.data(), ips[0], ops[0], st);
    if (lrc < 0) herr = mpx_host::last_error();
  }
  if (lrc == MPX_OK && (!done || hipEventRecord(done, st) != hipSuccess)) {
    lrc = MPX_ERR_HIP;
    herr = "hipEventRecord failed";
  }
  lk.lock();
  if (lrc < 0) {
    --slot.inflight;
    --load[s];
    if (done) events.push_back(done);
    L.ready.clear();
    fail_all(lrc, herr);
    for (auto& r : grp) r->ready = nullptr;
    if (lrc == MPX_ERR_HIP) break_queue(herr);
    return;
  }
  L.done = done;
  inflight.push_back(std::move(L));
  ++launches;
  cv.notify_all();
}

void mpx_queue::schedule() {
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv.wait(lk, [&] { return stop || (!queued.empty() && !held); });
    if (queued.empty()) break;  // stop, and everything queued has been launched
    // everything queued now, in arrival order, grouped by plan (groups in the order of
    // their first request); expired requests fail without running
    std::deque<std::unique_ptr<Req>> take;
    take.swap(queued);
    const auto now = Clock::now();
    std::vector<Key> order;
    std::map<Key, std::vector<std::unique_ptr<Req>>> groups;
    for (auto& r : take) {
      if (r->r.deadline_us > 0 &&
          std::chrono::duration_cast<std::chrono::microseconds>(now - r->arrival).count() > r->r.deadline_us) {
        finish(r->ticket, MPX_ERR_DEADLINE, "deadline passed before the request was launched");
        if (r->ready) events.push_back(r->ready);
        continue;
      }
      auto& g = groups[r->key];
      if (g.empty()) order.push_back(r->key);
      g.push_back(std::move(r));
    }
    cv.notify_all();
    for (const Key& k : order) {
      auto& g = groups[k];
      for (size_t i = 0; i < g.size(); i += max_batch) {
        std::vector<std::unique_ptr<Req>> chunk;
        for (size_t j = i; j < std::min(g.size(), i + max_batch); ++j) chunk.push_back(std::move(g[j]));
        launch_group(chunk, lk);
      }
    }
  }
  sched_done = true;
  cv.notify_all();
}

void mpx_queue::complete() {
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv.wait(lk, [&] { return !inflight.empty() || sched_done; });
    if (inflight.empty()) break;  // the scheduler has finished and every launch is done
    hipEvent_t done = inflight.front().done;
    lk.unlock();
    const hipError_t e = hipEventSynchronize(done);
    lk.lock();
    Launch L = std::move(inflight.front());
    inflight.pop_front();
    --load[L.stream];
    auto ps = plans.find(L.plan_key);
    if (ps != plans.end()) --ps->second.inflight;
    for (uint64_t t : L.tickets)
      finish(t, e == hipSuccess ? MPX_OK : MPX_ERR_HIP,
             e == hipSuccess ? std::string() : std::string("device error: ") + hipGetErrorString(e));
    events.push_back(L.done);
    for (hipEvent_t r : L.ready) events.push_back(r);
    if (e != hipSuccess) {
      const std::string m = std::string("device error: ") + hipGetErrorString(e);
      for (auto& f : inflight)
        for (uint64_t t : f.tickets) finish(t, MPX_ERR_HIP, m);
      break_queue(m);
    }
    cv.notify_all();
  }
}

extern "C" {

int mpx_queue_create(mpx_ctx* ctx, int streams, int max_batch, int max_w, int max_h, mpx_queue** out) {
  if (!ctx || !out) return fail(MPX_ERR_ARG, "ctx/out is NULL");
  *out = nullptr;
  if (streams < 1 || streams > 16) return fail(MPX_ERR_ARG, "streams outside 1..16");
  if (max_batch < 1 || max_batch > MPX_MAX_FRAMES) return fail(MPX_ERR_ARG, "max_batch outside 1..16");
  if (max_w < 1 || max_h < 1 || max_w > 32768 || max_h > 32768) return fail(MPX_ERR_SHAPE, "max size outside 1..32768");
  mpx_host::DeviceScope ds;
  if (int rc = ds.enter(ctx->device)) return rc;
  std::unique_ptr<mpx_queue> q(new mpx_queue());
  q->ctx = ctx;
  q->max_batch = max_batch;
  q->max_w = max_w;
  q->max_h = max_h;
  for (int i = 0; i < streams; ++i) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
      for (hipStream_t t : q->streams) (void)hipStreamDestroy(t);
      return fail(MPX_ERR_HIP, "hipStreamCreateWithFlags failed");
    }
    q->streams.push_back(s);
  }
  q->load.assign(streams, 0);
  q->scheduler = std::thread([p = q.get()] { p->schedule(); });
  q->completer = std::thread([p = q.get()] { p->complete(); });
  *out = q.release();
  return MPX_OK;
}

void mpx_queue_destroy(mpx_queue* q) {
  if (!q) return;
  {
    std::lock_guard<std::mutex> lk(q->mu);
    q->stop = true;
    q->held = false;
  }
  q->cv.notify_all();
  q->scheduler.join();
  q->completer.join();
  mpx_host::DeviceScope ds;
  (void)ds.enter(q->ctx->device);
  for (auto& kv : q->plans) mpx_plan_destroy(kv.second.plan);
  for (hipStream_t s : q->streams) (void)hipStreamDestroy(s);
  for (hipEvent_t e : q->events) (void)hipEventDestroy(e);
  delete q;
}

int mpx_submit(mpx_queue* q, const mpx_request* r, uint64_t* ticket) {
  if (!q || !r || !ticket) return fail(MPX_ERR_ARG, "queue/request/ticket is NULL");
  if (!r->name) return fail(MPX_ERR_UNKNOWN_VARIANT, "variant name is NULL");
  if (!r->in || !r->out) return fail(MPX_ERR_ARG, "in/out is NULL");
  if (r->in_type < MPX_U8 || r->in_type > MPX_F32) return fail(MPX_ERR_ARG, "bad in_type");
  if (r->out_type != MPX_F16 && r->out_type != MPX_F32) return fail(MPX_ERR_ARG, "out_type must be MPX_F16 or MPX_F32");
  if (r->in_type == MPX_U16 && (r->bits < 1 || r->bits > 16)) return fail(MPX_ERR_ARG, "bits must be 1..16 for MPX_U16");
  const int prec = r->precision & ~MPX_STORE_FP16;
  if (prec < MPX_PREC_FP32 || prec > MPX_PREC_FP16X3) return fail(MPX_ERR_ARG, "bad precision");
  if (r->in_w < 1 || r->in_h < 1 || r->in_w > 32768 || r->in_h > 32768) return fail(MPX_ERR_SHAPE, "sizes out of range");
  if ((r->target_w == 0) != (r->target_h == 0) || r->target_w < 0 || r->target_h < 0)
    return fail(MPX_ERR_ARG, "target must be both 0 (default) or both positive");
  if (r->deadline_us < 0) return fail(MPX_ERR_ARG, "negative deadline");
  const std::string name(r->name);
  const int k = default_scale(name);
  const int tw = r->target_w ? r->target_w : r->in_w * k, th = r->target_h ? r->target_h : r->in_h * k;
  mpx_variant_info info;
  if (int rc = mpx_query(r->name, r->in_w, r->in_h, tw, th, &info)) return rc;
  if (!info.available)
    if (int rc = mpx_weights_verify(q->ctx->wdir.c_str(), r->name)) return rc;
  if (info.family != MPX_FAMILY_NNEDI3 && prec != MPX_PREC_FP32)
    return fail(MPX_ERR_UNSUPPORTED, "RAVU variants compute in FP32 only");
  std::unique_ptr<Req> req(new Req());
  req->name = name;
  req->r = *r;
  req->r.name = nullptr;  // the copy in req->name is used
  req->out_w = info.out_w;
  const bool default_target = tw == r->in_w * k && th == r->in_h * k;
  const bool sized = info.applies && info.family != MPX_FAMILY_RAVU_ZOOM && default_target && r->in_w <= q->max_w &&
                     r->in_h <= q->max_h;
  req->key = sized ? Key{name, q->max_w, q->max_h, q->max_w * k, q->max_h * k, r->in_type, r->bits, r->out_type,
                         r->precision, 1}
                   : Key{name, r->in_w, r->in_h, tw, th, r->in_type, r->bits, r->out_type, r->precision, 0};
  if (r->in_type != MPX_U16) req->key.bits = 0;
  req->arrival = Clock::now();
  std::unique_lock<std::mutex> lk(q->mu);
  if (q->broken) return fail(MPX_ERR_HIP, "the queue stopped after a device error: " + q->broken_msg);
  if (q->stop) return fail(MPX_ERR_ARG, "the queue is being destroyed");
  if (r->ready_stream) {
    hipEvent_t e = q->take_event();
    mpx_host::DeviceScope ds;
    if (!e || ds.enter(q->ctx->device) || hipEventRecord(e, static_cast<hipStream_t>(r->ready_stream)) != hipSuccess) {
      if (e) q->events.push_back(e);
      return fail(MPX_ERR_HIP, "cannot record the ready event on ready_stream");
    }
    req->ready = e;
  }
  req->ticket = q->next_ticket++;
  *ticket = req->ticket;
  q->status[req->ticket] = MPX_PENDING;
  q->arrival[req->ticket] = req->arrival;
  q->queued.push_back(std::move(req));
  lk.unlock();
  q->cv.notify_all();
  return MPX_OK;
}

int mpx_poll(mpx_queue* q, uint64_t ticket, int64_t timeout_us, int64_t* latency_us) {
  if (!q) return fail(MPX_ERR_ARG, "queue is NULL");
  std::unique_lock<std::mutex> lk(q->mu);
  auto it = q->status.find(ticket);
  if (it == q->status.end()) return fail(MPX_ERR_ARG, "unknown or retired ticket " + std::to_string(ticket));
  auto finished = [&] {
    auto f = q->status.find(ticket);
    return f == q->status.end() || f->second != MPX_PENDING;
  };
  if (timeout_us < 0) q->cv.wait(lk, finished);
  else if (timeout_us > 0) q->cv.wait_for(lk, std::chrono::microseconds(timeout_us), finished);
  it = q->status.find(ticket);
  if (it == q->status.end()) return fail(MPX_ERR_ARG, "ticket " + std::to_string(ticket) + " retired by another poll");
  const int st = it->second;
  if (st == MPX_PENDING) return MPX_PENDING;
  q->status.erase(it);
  auto lt = q->latency.find(ticket);
  if (lt != q->latency.end()) {
    if (latency_us) *latency_us = lt->second;
    q->latency.erase(lt);
  }
  if (st < 0) {
    const std::string msg = q->errors[ticket];
    q->errors.erase(ticket);
    return fail(st, msg);
  }
  return MPX_OK;
}

int mpx_queue_hold(mpx_queue* q, int hold) {
  if (!q) return fail(MPX_ERR_ARG, "queue is NULL");
  {
    std::lock_guard<std::mutex> lk(q->mu);
    q->held = hold != 0;
  }
  q->cv.notify_all();
  return MPX_OK;
}

uint64_t mpx_queue_launches(const mpx_queue* q) {
  if (!q) return 0;
  std::lock_guard<std::mutex> lk(const_cast<mpx_queue*>(q)->mu);
  return q->launches;
}

}  // extern "C"
