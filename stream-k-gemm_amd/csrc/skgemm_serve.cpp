// Continuous-batching request loop over systems whose Psi stays resident in
// HBM.  No counterpart in the reference (its drivers run one problem per
// call); this is the serving side of the batched launch.
//
// Threads: callers submit requests (overlap Ms = Psi^T Psi or rotation
// Psi * Q) into a queue; a dispatcher thread takes whatever is queued (up to
// max_batch requests whose split-form workspace fits the free HBM budget)
// as soon as one of max_inflight launch slots is free and issues it as one
// skgemm_batched launch on that slot's stream, so work that arrives while
// earlier launches run joins the next launch; a completion thread polls the
// slots' events and stamps each request's completion time.
//
// Slots run concurrently on separate default-priority streams.  No launch
// of a slot waits between its workgroups: plain FP64 slots run the
// persistent Stream-K kernel with the wait-free (last-arriver) fix-up and
// split slots only data-parallel INT8 schedules, so every launch takes the
// whole device's worth of workgroups, a launch running alone has all of it,
// and launches in flight together share the CUs as their workgroups
// retire.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/skgemm.h"
#include "skgemm_ctx.h"

using namespace skg;

namespace {

constexpr int kServeCfg = 8;  // 128 x 128 tiles, 8 waves, LDS-DMA (square: SYRK)
constexpr int kMaxInflight = 8;

double now_s() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct System {
  const double* psi;
  int Nd, Ns;
};

struct Request {
  int sys = 0, kind = 0;
  const double* q = nullptr;
  double* out = nullptr;
  double t_submit = 0.0, t_done = 0.0;
  size_t bytes = 0;  // split-form workspace estimate
  double flops = 0.0;
  bool done = false;
  int rc = SKGEMM_OK;
  std::string err;  // error text, set in the thread that saw the error
};

struct Slot {
  skgemm_ctx* ctx = nullptr;
  hipEvent_t ev = nullptr;
  bool busy = false;   // taken by the dispatcher
  bool armed = false;  // launched and its completion event recorded
  std::vector<int64_t> tickets;
  size_t bytes = 0;
  double flops = 0.0;
};

}  // namespace

struct skgemm_server {
  int device = 0;
  size_t budget = 0;
  int max_batch = 1, max_inflight = 1, precision = SKGEMM_FP64, slices = 0;
  std::vector<System> systems;
  size_t resident = 0;  // bytes of registered Psi

  std::mutex mu;
  std::condition_variable cv_work;  // dispatcher: queue or slot changed
  std::condition_variable cv_busy;  // completer: a slot became busy
  std::condition_variable cv_done;  // waiters: a request completed
  std::deque<int64_t> queue;
  std::unordered_map<int64_t, Request> reqs;
  int64_t next_ticket = 1;
  std::vector<Slot> slots;
  size_t inflight_bytes = 0;
  bool stop = false;
  std::thread dispatcher, completer;

  int64_t completed = 0, launches = 0;
  double flops = 0.0, busy_s = 0.0, busy_since = 0.0;
  int inflight = 0, peak_inflight = 0;

  void dispatch_loop();
  void complete_loop();
};

namespace {

// Split-form workspace of one request: residue planes of Psi (and of Q),
// chunk residue planes of the output, norms.  Plain FP64 needs none beyond
// the slot's fixed Stream-K workspace.
size_t request_bytes(const skgemm_server* s, const System& y, int kind) {
  if (s->precision != SKGEMM_OZAKI_INT8) return 0;
  const size_t L = s->slices == 0 ? 14 : (size_t)s->slices;
  auto up = [](size_t x, size_t q) { return (x + q - 1) / q * q; };
  if (kind == 0) {  // SYRK: K = Nd
    const size_t kpad = up((size_t)y.Nd, 128), o = up((size_t)y.Ns, 256);
    return L * o * kpad + 16 * L * (size_t)y.Ns * up((size_t)y.Ns, 16);
  }
  const size_t kpad = up((size_t)y.Ns, 128);  // rotation: K = Ns
  return L * (up((size_t)y.Nd, 256) + up((size_t)y.Ns, 256)) * kpad +
         L * (size_t)y.Ns * up((size_t)y.Nd, 16);
}

double request_flops(const System& y, int kind) {
  const double Nd = y.Nd, Ns = y.Ns;
  return kind == 0 ? Ns * (Ns + 1) * Nd : 2.0 * Nd * Ns * Ns;
}

}  // namespace

void skgemm_server::dispatch_loop() {
  hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    int si = -1;
    cv_work.wait(lk, [&] {
      if (stop && queue.empty()) return true;
      if (queue.empty()) return false;
      for (int i = 0; i < (int)slots.size(); ++i)
        if (!slots[i].busy) {
          si = i;
          return true;
        }
      return false;
    });
    if (stop && queue.empty()) return;
    // Admission: take queued requests in order while the batch is under
    // max_batch and its workspace fits the budget left by other launches
    // (the head request always goes when nothing else is in flight).
    std::vector<int64_t> batch;
    size_t bytes = 0;
    double fl = 0.0;
    auto fits = [&](size_t b) { return resident + inflight_bytes + b <= budget; };
    while (!queue.empty() && (int)batch.size() < max_batch) {
      const Request& r = reqs[queue.front()];
      const size_t nb = std::max(bytes, r.bytes);  // the slot reuses its workspace
      if (!fits(nb) && (!batch.empty() || inflight > 0)) break;
      bytes = nb;
      fl += r.flops;
      batch.push_back(queue.front());
      queue.pop_front();
    }
    if (batch.empty()) {  // wait for a launch to finish and free budget
      cv_work.wait(lk, [&] {
        return stop || inflight == 0 || fits(reqs[queue.front()].bytes);
      });
      continue;
    }
    Slot& sl = slots[si];
    sl.busy = true;
    sl.tickets = batch;
    sl.bytes = bytes;
    sl.flops = fl;
    inflight_bytes += bytes;
    if (inflight++ == 0) busy_since = now_s();
    peak_inflight = std::max(peak_inflight, inflight);
    std::vector<skgemm_problem> probs(batch.size());
    for (size_t i = 0; i < batch.size(); ++i) {
      const Request& r = reqs[batch[i]];
      const System& y = systems[r.sys];
      skgemm_problem& p = probs[i];
      p = skgemm_problem{};
      if (r.kind == 0) {  // Ms = Psi^T Psi, Psi Nd x Ns column-major
        p.kind = 1;
        p.layout_a = SKGEMM_COL_MAJOR;
        p.M = p.N = y.Ns;
        p.K = y.Nd;
        p.alpha = 1.0;
        p.A = y.psi;
        p.lda = y.Nd;
        p.D = r.out;
        p.ldd = y.Ns;
      } else {  // out = Psi Q, Q Ns x Ns row-major, out column-major
        p.kind = 0;
        p.layout_a = SKGEMM_COL_MAJOR;
        p.layout_b = SKGEMM_ROW_MAJOR;
        p.layout_d = SKGEMM_COL_MAJOR;
        p.M = y.Nd;
        p.N = y.Ns;
        p.K = y.Ns;
        p.alpha = 1.0;
        p.A = y.psi;
        p.lda = y.Nd;
        p.B = r.q;
        p.ldb = y.Ns;
        p.D = r.out;
        p.ldd = y.Nd;
      }
    }
    ++launches;
    lk.unlock();
    int rc = skgemm_batched(sl.ctx, probs.data(), (int)probs.size(), precision,
                            slices, precision == SKGEMM_FP64 ? 1 : 0);
    if (rc == SKGEMM_OK && hipEventRecord(sl.ev, sl.ctx->stream) != hipSuccess)
      rc = set_error(SKGEMM_ERR_HIP, "server event record");
    lk.lock();
    if (rc == SKGEMM_OK) sl.armed = true;
    if (rc != SKGEMM_OK) {  // complete the batch with the error
      const double t = now_s();
      for (int64_t tk : batch) {
        Request& r = reqs[tk];
        r.done = true;
        r.rc = rc;
        r.err = skgemm_last_error();
        r.t_done = t;
      }
      sl.busy = false;
      sl.tickets.clear();
      inflight_bytes -= bytes;
      if (--inflight == 0) busy_s += now_s() - busy_since;
      cv_done.notify_all();
      continue;
    }
    cv_busy.notify_one();
  }
}

void skgemm_server::complete_loop() {
  hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_busy.wait(lk, [&] {
      if (stop && queue.empty() && inflight == 0) return true;
      for (const Slot& s : slots)
        if (s.armed) return true;
      return false;
    });
    if (stop && queue.empty() && inflight == 0) return;
    bool any = false;
    for (Slot& s : slots) {
      if (!s.armed) continue;
      const hipError_t q = hipEventQuery(s.ev);
      if (q == hipErrorNotReady) continue;
      const double t = now_s();
      int rc = q == hipSuccess ? s.ctx->check_status()
                               : set_error(SKGEMM_ERR_HIP, "server launch: %s",
                                           hipGetErrorString(q));
      for (int64_t tk : s.tickets) {
        Request& r = reqs[tk];
        r.done = true;
        r.rc = rc;
        if (rc != SKGEMM_OK) r.err = skgemm_last_error();
        r.t_done = t;
      }
      completed += (int64_t)s.tickets.size();
      flops += s.flops;
      s.tickets.clear();
      s.armed = false;
      s.busy = false;
      inflight_bytes -= s.bytes;
      if (--inflight == 0) busy_s += t - busy_since;
      any = true;
    }
    if (any) {
      cv_done.notify_all();
      cv_work.notify_one();
      continue;
    }
    lk.unlock();
    std::this_thread::sleep_for(std::chrono::microseconds(20));
    lk.lock();
  }
}

extern "C" int skgemm_server_create(int device, size_t hbm_budget, int max_batch,
                                    int max_inflight, int precision, int slices,
                                    skgemm_server** out) {
  if (out == nullptr) return set_error(SKGEMM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (max_batch < 1 || max_inflight < 1 || max_inflight > kMaxInflight)
    return set_error(SKGEMM_ERR_ARG, "max_batch >= 1, 1 <= max_inflight <= %d",
                     kMaxInflight);
  if (precision != SKGEMM_FP64 && precision != SKGEMM_OZAKI_INT8)
    return set_error(SKGEMM_ERR_ARG, "precision must be SKGEMM_FP64 or "
                                     "SKGEMM_OZAKI_INT8");
  if (slices != 0 && (slices < 8 || slices > 16))
    return set_error(SKGEMM_ERR_ARG, "slices must be 0 or 8..16");
  auto* s = new skgemm_server();
  s->device = device;
  s->budget = hbm_budget;
  s->max_batch = max_batch;
  s->max_inflight = max_inflight;
  s->precision = precision;
  s->slices = slices;
  s->slots.resize(max_inflight);
  for (Slot& sl : s->slots) {
    int rc = skgemm_create(device, &sl.ctx);
    if (rc == SKGEMM_OK &&
        hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess)
      rc = set_error(SKGEMM_ERR_HIP, "server event");
    if (rc == SKGEMM_OK) {
      // Nothing waits across workgroups, so each launch may take the
      // whole device whatever else is in flight.
      rc = skgemm_set_tuning(sl.ctx, kServeCfg, sl.ctx->capacity(kServeCfg));
      sl.ctx->waitfree = true;
      sl.ctx->oz_no_wait = true;
    }
    if (rc != SKGEMM_OK) {
      for (Slot& x : s->slots) {
        if (x.ev) hipEventDestroy(x.ev);
        if (x.ctx) skgemm_destroy(x.ctx);
      }
      delete s;
      return rc;
    }
  }
  s->dispatcher = std::thread([s] { s->dispatch_loop(); });
  s->completer = std::thread([s] { s->complete_loop(); });
  *out = s;
  return SKGEMM_OK;
}

extern "C" int skgemm_server_destroy(skgemm_server* s) {
  if (s == nullptr) return set_error(SKGEMM_ERR_ARG, "server is NULL");
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;  // queued requests are still run
  }
  s->cv_work.notify_all();
  s->cv_busy.notify_all();
  if (s->dispatcher.joinable()) s->dispatcher.join();
  {
    std::lock_guard<std::mutex> g(s->mu);
  }
  s->cv_busy.notify_all();
  if (s->completer.joinable()) s->completer.join();
  for (Slot& x : s->slots) {
    if (x.ctx) skgemm_synchronize(x.ctx);
    if (x.ev) hipEventDestroy(x.ev);
    if (x.ctx) skgemm_destroy(x.ctx);
  }
  s->cv_done.notify_all();
  delete s;
  return SKGEMM_OK;
}

extern "C" int skgemm_server_add_system(skgemm_server* s, const double* psi,
                                        int Nd, int Ns, int* id) {
  if (s == nullptr || id == nullptr)
    return set_error(SKGEMM_ERR_ARG, "server or id is NULL");
  if (Nd <= 0 || Ns <= 0 || psi == nullptr)
    return set_error(SKGEMM_ERR_ARG, "Psi must be a non-empty device matrix");
  SKG_TRY(check_dev(psi, "psi"));
  std::lock_guard<std::mutex> g(s->mu);
  const size_t bytes = sizeof(double) * (size_t)Nd * Ns;
  if (s->resident + bytes > s->budget)
    return set_error(SKGEMM_ERR_NOMEM,
                     "resident Psi pool would exceed the HBM budget "
                     "(%zu + %zu > %zu bytes)", s->resident, bytes, s->budget);
  s->resident += bytes;  // the pool and the launch workspace share the budget
  s->systems.push_back(System{psi, Nd, Ns});
  *id = (int)s->systems.size() - 1;
  return SKGEMM_OK;
}

extern "C" int skgemm_server_submit(skgemm_server* s, int system, int kind,
                                    const double* q, double* out,
                                    int64_t* ticket) {
  if (s == nullptr || ticket == nullptr)
    return set_error(SKGEMM_ERR_ARG, "server or ticket is NULL");
  if (kind != 0 && kind != 1)
    return set_error(SKGEMM_ERR_ARG, "kind must be 0 (overlap) or 1 (rotation)");
  SKG_TRY(check_dev(out, "out"));
  if (out == nullptr) return set_error(SKGEMM_ERR_ARG, "out is NULL");
  if (kind == 1) {
    if (q == nullptr) return set_error(SKGEMM_ERR_ARG, "Q is NULL");
    SKG_TRY(check_dev(q, "Q"));
  }
  std::lock_guard<std::mutex> g(s->mu);
  if (s->stop) return set_error(SKGEMM_ERR_STATE, "server is shutting down");
  if (system < 0 || system >= (int)s->systems.size())
    return set_error(SKGEMM_ERR_ARG, "unknown system %d", system);
  const System& y = s->systems[system];
  Request r;
  r.sys = system;
  r.kind = kind;
  r.q = q;
  r.out = out;
  r.bytes = request_bytes(s, y, kind);
  if (s->resident + r.bytes > s->budget)
    return set_error(SKGEMM_ERR_NOMEM,
                     "request workspace %zu bytes exceeds the HBM budget left "
                     "by the resident pool", r.bytes);
  r.flops = request_flops(y, kind);
  r.t_submit = now_s();
  const int64_t t = s->next_ticket++;
  s->reqs.emplace(t, r);
  s->queue.push_back(t);
  *ticket = t;
  s->cv_work.notify_one();
  return SKGEMM_OK;
}

extern "C" int skgemm_server_wait(skgemm_server* s, int64_t ticket,
                                  double* latency_us) {
  if (s == nullptr) return set_error(SKGEMM_ERR_ARG, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end())
    return set_error(SKGEMM_ERR_ARG, "unknown or already collected ticket");
  s->cv_done.wait(lk, [&] { return s->reqs[ticket].done; });
  const Request r = s->reqs[ticket];
  s->reqs.erase(ticket);
  if (latency_us) *latency_us = (r.t_done - r.t_submit) * 1e6;
  if (r.rc != SKGEMM_OK) return set_error(r.rc, "%s", r.err.c_str());
  return r.rc;
}

extern "C" int skgemm_server_stats(skgemm_server* s, int64_t* completed,
                                   int64_t* launches, double* flops,
                                   double* busy_s, int* peak_inflight) {
  if (s == nullptr) return set_error(SKGEMM_ERR_ARG, "server is NULL");
  std::lock_guard<std::mutex> g(s->mu);
  if (completed) *completed = s->completed;
  if (launches) *launches = s->launches;
  if (flops) *flops = s->flops;
  if (busy_s)
    *busy_s = s->busy_s + (s->inflight > 0 ? now_s() - s->busy_since : 0.0);
  if (peak_inflight) *peak_inflight = s->peak_inflight;
  return SKGEMM_OK;
}
