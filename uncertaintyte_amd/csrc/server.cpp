// server.cpp -- request loop for many reconstructions of different sizes.
//
// Requests are submitted with host buffers and a ticket comes back at once.
// A fixed set of slots, each a context with its own HIP stream and cached
// device buffers, takes requests off the queue: a request is admitted when a
// slot is idle and the device bytes all slots would then hold stay within the
// HBM budget (the dense Schur complement and the inverse buffers dominate:
// 5 * 8 * (cp * cams)^2 bytes).  Admission is FIFO; a later request may pass a
// head that does not fit yet for at most `bypass_ms`, after which the queue
// waits for memory to free up (no starvation).  When the head does not fit,
// idle slots drop their cached buffers first.  Requests on different slots
// overlap on the device; each request runs the same kernels as
// unc_compute, so its result does not depend on what else is in flight.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/uncertainty_amd.h"
#include "internal.h"

using Clock = std::chrono::steady_clock;

struct unc_server {
  struct Req {
    long long ticket = 0;
    unc_options o;
    int nrows = 0, ncols = 0;
    const int *rows = nullptr, *cols = nullptr;
    const double *vals = nullptr;
    double *cam = nullptr, *pt = nullptr;
    unc_stats *st = nullptr;
    size_t ws_need = 0, io_need = 0;
    Clock::time_point arrival, start, end;
    int status = UNC_OK;
    bool done = false;
  };
  struct Slot {
    unc_context *ctx = nullptr;
    bool busy = false;
    size_t held = 0;  // bytes the slot holds (or will hold for its running request)
    std::thread th;
  };
  unc_server_config cfg;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req *> queue;
  std::map<long long, Req *> reqs;
  std::vector<Slot> slots;
  size_t held_total = 0, peak_held = 0;
  long long next_ticket = 1, completed = 0, admitted_out_of_order = 0;
  int max_concurrent = 0, running = 0;
  bool stop = false;

  size_t grown(const Slot &s, const Req *r) const {
    const size_t ws = unc_ctx_ws_bytes(s.ctx), io = unc_ctx_io_bytes(s.ctx);
    return (ws > r->ws_need ? ws : r->ws_need) + (io > r->io_need ? io : r->io_need);
  }

  // with mu held; slot i idle
  Req *pick(int i) {
    Slot &me = slots[i];
    const auto now = Clock::now();
    for (size_t q = 0; q < queue.size(); ++q) {
      Req *r = queue[q];
      size_t after = grown(me, r);
      size_t total = held_total - me.held + after;
      if (total > cfg.hbm_budget && q == 0) {
        // the head does not fit: idle slots drop their cached buffers
        for (size_t j = 0; j < slots.size() && total > cfg.hbm_budget; ++j) {
          Slot &o = slots[j];
          if ((int)j == i || o.busy || o.held == 0) continue;
          unc_ctx_release(o.ctx);
          held_total -= o.held;
          total -= o.held;
          o.held = 0;
        }
        if (total > cfg.hbm_budget && me.held > 0) {
          // start this slot from empty buffers sized to the request
          unc_ctx_release(me.ctx);
          held_total -= me.held;
          me.held = 0;
          after = r->ws_need + r->io_need;
          total = held_total + after;
        }
      }
      if (total <= cfg.hbm_budget) {
        queue.erase(queue.begin() + (long)q);
        if (q > 0) ++admitted_out_of_order;
        held_total = total;
        me.held = after;
        if (held_total > peak_held) peak_held = held_total;
        return r;
      }
      // a later request may pass the head only while the head is young
      const double waited =
          std::chrono::duration<double, std::milli>(now - queue[0]->arrival).count();
      if (waited > cfg.bypass_ms) return nullptr;
    }
    return nullptr;
  }

  void worker(int i) {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      Req *r = nullptr;
      for (;;) {
        if (!queue.empty() && (r = pick(i))) break;
        if (stop && queue.empty()) return;
        // re-check periodically: the bypass window of the head expires by time
        cv_work.wait_for(lk, std::chrono::milliseconds(2));
      }
      Slot &me = slots[i];
      me.busy = true;
      ++running;
      if (running > max_concurrent) max_concurrent = running;
      lk.unlock();
      r->start = Clock::now();
      const int rc = unc_compute(me.ctx, &r->o, r->nrows, r->ncols, r->rows, r->cols, r->vals,
                                 r->cam, r->pt, r->st);
      r->end = Clock::now();
      lk.lock();
      --running;
      me.busy = false;
      // the admitted estimate is the layout the run allocates; should the
      // context hold more than was admitted, its buffers are dropped so that
      // the budget holds between requests too
      const size_t actual = unc_ctx_held_bytes(me.ctx);
      held_total -= me.held;
      if (actual > me.held) {
        unc_ctx_release(me.ctx);
        me.held = 0;
      } else {
        me.held = actual;
      }
      held_total += me.held;
      if (held_total > peak_held) peak_held = held_total;
      r->status = rc;
      r->done = true;
      ++completed;
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
};

constexpr int SERVER_FORK_MAX_SLOTS = 2;  // two streams per slot, four hardware queues

extern "C" {

void unc_server_default_config(unc_server_config *c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->device = 0;
  c->slots = 2;
  c->hbm_budget = (size_t)64 << 30;
  c->bypass_ms = 50.0;
}

int unc_server_create(const unc_server_config *cfg, unc_server **out) {
  if (!cfg || !out || cfg->slots < 1 || cfg->slots > 64 || cfg->hbm_budget == 0)
    return UNC_ERR_BAD_INPUT;
  *out = nullptr;
  unc_server *s = new unc_server();
  s->cfg = *cfg;
  s->slots.resize(cfg->slots);
  for (int i = 0; i < cfg->slots; ++i) {
    const int rc = unc_context_create(cfg->device, &s->slots[i].ctx);
    if (rc) {
      for (int j = 0; j < i; ++j) unc_context_destroy(s->slots[j].ctx);
      delete s;
      return rc;
    }
    // each slot's recursion forks onto a second stream only while the slots'
    // streams do not outnumber the device's four hardware queues (HIP's
    // default): shared queues serialise unrelated streams behind the
    // cross-stream waits of the fork
    if (cfg->slots > SERVER_FORK_MAX_SLOTS) unc_ctx_set_fork(s->slots[i].ctx, false);
  }
  for (int i = 0; i < cfg->slots; ++i) s->slots[i].th = std::thread(&unc_server::worker, s, i);
  *out = s;
  return UNC_OK;
}

int unc_server_submit(unc_server *s, const unc_options *opts, int num_rows, int num_cols,
                      const int *rows, const int *cols, const double *values, double *cam_unc,
                      double *pt_unc, unc_stats *stats, long long *ticket) {
  if (!s || !opts || !rows || !cols || !values || !cam_unc || !pt_unc || !ticket)
    return UNC_ERR_BAD_INPUT;
  if (num_rows <= 0 || opts->cam_params < 1 || opts->cam_params > 9) return UNC_ERR_BAD_INPUT;
  auto *r = new unc_server::Req();
  r->o = *opts;
  r->nrows = num_rows;
  r->ncols = num_cols;
  r->rows = rows;
  r->cols = cols;
  r->vals = values;
  r->cam = cam_unc;
  r->pt = pt_unc;
  r->st = stats;
  r->ws_need = unc_workspace_bytes(opts, num_rows);
  r->io_need = unc_io_bytes_for(opts, num_rows, (size_t)rows[num_rows]);
  if (r->ws_need == 0) {
    delete r;
    return UNC_ERR_BAD_INPUT;
  }
  if (r->ws_need + r->io_need > s->cfg.hbm_budget) {  // could never be admitted
    delete r;
    return UNC_ERR_OOM;
  }
  r->arrival = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  r->ticket = s->next_ticket++;
  s->reqs[r->ticket] = r;
  s->queue.push_back(r);
  *ticket = r->ticket;
  s->cv_work.notify_all();
  return UNC_OK;
}

int unc_server_wait(unc_server *s, long long ticket, int *status, double *latency_ms,
                    double *queue_ms) {
  if (!s) return UNC_ERR_BAD_INPUT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return UNC_ERR_BAD_INPUT;  // unknown or already claimed
  unc_server::Req *r = it->second;
  s->reqs.erase(it);  // claimed: a second waiter on this ticket gets BAD_INPUT
  s->cv_done.wait(lk, [&] { return r->done; });
  if (status) *status = r->status;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->end - r->arrival).count();
  if (queue_ms) *queue_ms = std::chrono::duration<double, std::milli>(r->start - r->arrival).count();
  delete r;
  return UNC_OK;
}

int unc_server_query(unc_server *s, unc_server_stats *out) {
  if (!s || !out) return UNC_ERR_BAD_INPUT;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof(*out));
  out->queued = (long long)s->queue.size();
  out->completed = s->completed;
  out->admitted_out_of_order = s->admitted_out_of_order;
  out->held_bytes = s->held_total;
  out->peak_held_bytes = s->peak_held;
  out->max_concurrent = s->max_concurrent;
  return UNC_OK;
}

void unc_server_destroy(unc_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  for (auto &sl : s->slots)
    if (sl.th.joinable()) sl.th.join();
  for (auto &kv : s->reqs) delete kv.second;
  for (auto &sl : s->slots) unc_context_destroy(sl.ctx);
  delete s;
}

}  // extern "C"
