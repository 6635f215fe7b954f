// Continuous-batching request server over one Engine (mixed or FP64 PCG).
//
// A worker thread owns the engine and loops:
//   1. admit queued requests (FIFO) while a slot is free and the device
//      memory budget allows their resident storage (coefficients and Krylov
//      vectors of every admitted system stay in HBM until it completes);
//   2. run one replay of the captured iteration graph for all active systems;
//   3. for every system that finished its current time step: start the next
//      step (new right-hand side, warm start from the current solution) in
//      the same slot, or complete the request and free its storage.
// Requests that arrive while others run are admitted between replays, so
// their iterations share the same kernel launches.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"

namespace {

struct Req {
  long long id = 0;
  aphpois_request r{};
  int slot = -1;
  int step = 0;
  aphpois_result res{};
  bool done = false;
  bool cancel = false;
};

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct aphpois_server {
  aph::Engine eng;
  aphpois_server_desc desc{};
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Req>> queue;
  std::map<long long, std::shared_ptr<Req>> reqs;  // all not yet collected
  std::map<int, std::shared_ptr<Req>> by_slot;     // admitted
  long long next_id = 1;
  bool stop = false;
  std::string fatal;
  long long completed = 0;
  std::thread worker;

  void finish(const std::shared_ptr<Req>& q, int status) {
    q->res.status = status;
    q->res.t_done = now_s(t0);
    q->res.steps_done = q->step;
    const int slot = q->slot;
    if (slot >= 0) eng.remove(slot);
    std::lock_guard<std::mutex> lk(mu);
    if (slot >= 0) by_slot.erase(slot);
    q->slot = -1;
    q->done = true;
    ++completed;
    cv_done.notify_all();
  }

  // Loads the next step of an admitted request; returns false on error.
  bool start_step(const std::shared_ptr<Req>& q) {
    const aphpois_request& r = q->r;
    const long long ly = r.nx, lz = (long long)r.nx * r.ny;
    int rc = APHPOIS_OK;
    if (q->step > 0) {
      const double* e7 = r.e7 ? r.e7[q->step] : nullptr;
      if (e7) {
        rc = eng.set_rhs(q->slot, e7, false, 0, ly, lz);
      } else {
        const unsigned long long key =
            r.seed * 0x9E3779B97F4A7C15ull + (unsigned long long)q->step;
        rc = eng.perturb_rhs(q->slot, r.perturb, key);
      }
    }
    if (rc == APHPOIS_OK) rc = eng.prepare(q->slot, nullptr, false, 0, ly, lz, q->step > 0);
    if (rc != APHPOIS_OK) {
      finish(q, rc);
      return false;
    }
    return true;
  }

  void admit() {
    for (;;) {
      std::shared_ptr<Req> q;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (queue.empty()) return;
        if (desc.max_active > 0 && (int)by_slot.size() >= desc.max_active) return;
        q = queue.front();
        if (q->cancel) {
          queue.pop_front();
          q->done = true;
          q->res.status = APHPOIS_ECANCEL;
          q->res.t_done = now_s(t0);
          cv_done.notify_all();
          continue;
        }
      }
      const aphpois_request& r = q->r;
      int slot = -1;
      int rc = eng.add_system(r.nx, r.ny, r.nz, r.h3, &slot);
      if (rc == APHPOIS_OK) rc = eng.set_system(slot, r.expr, false, 0, r.nx, (long long)r.nx * r.ny);
      if (rc == APHPOIS_EBUDGET) {
        if (slot >= 0) eng.remove(slot);
        if (!by_slot.empty()) return;  // wait for memory to free up
        // nothing else resident: the request can never fit
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        queue.pop_front();
      }
      if (rc != APHPOIS_OK) {
        if (slot >= 0 && rc != APHPOIS_EBUDGET) eng.remove(slot);
        q->slot = -1;
        q->res.t_admit = now_s(t0);
        finish(q, rc);
        continue;
      }
      q->slot = slot;
      q->res.t_admit = now_s(t0);
      {
        std::lock_guard<std::mutex> lk(mu);
        by_slot[slot] = q;
      }
      start_step(q);
    }
  }

  void loop() {
    std::vector<int> fin;
    while (true) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !queue.empty() || !by_slot.empty(); });
        if (stop) break;
      }
      admit();
      // cancellation of admitted requests
      std::vector<std::shared_ptr<Req>> cancelled;
      {
        std::lock_guard<std::mutex> lk(mu);
        for (auto& kv : by_slot)
          if (kv.second->cancel) cancelled.push_back(kv.second);
      }
      for (auto& q : cancelled) finish(q, APHPOIS_ECANCEL);
      if (by_slot.empty()) continue;
      const int rc = eng.step(1, &fin);
      if (rc != APHPOIS_OK) {
        std::lock_guard<std::mutex> lk(mu);
        fatal = eng.err();
        // fail every admitted request; the device state is not trusted
        for (auto& kv : by_slot) {
          kv.second->res.status = rc;
          kv.second->res.t_done = now_s(t0);
          kv.second->done = true;
        }
        by_slot.clear();
        for (auto& q : queue) {
          q->res.status = rc;
          q->done = true;
        }
        queue.clear();
        cv_done.notify_all();
        stop = true;
        break;
      }
      for (int slot : fin) {
        auto it = by_slot.find(slot);
        if (it == by_slot.end()) continue;
        std::shared_ptr<Req> q = it->second;
        aphpois_info info{};
        eng.fetch(slot, nullptr, false, 0, 0, 0, &info);
        q->res.iters_total += info.iter;
        q->res.last_relres = info.true_relres;
        q->res.refinements_total += info.refinements;
        q->step += 1;
        if (info.status != APHPOIS_OK || !info.converged) {
          finish(q, info.status != APHPOIS_OK ? info.status : APHPOIS_ENUMERIC);
          continue;
        }
        if (q->step < q->r.nsteps) {
          start_step(q);
        } else {
          if (q->r.x_out) {
            const int frc = eng.fetch(slot, q->r.x_out, false, 0, q->r.nx,
                                      (long long)q->r.nx * q->r.ny, nullptr);
            if (frc != APHPOIS_OK) {
              finish(q, frc);
              continue;
            }
          }
          finish(q, APHPOIS_OK);
        }
      }
    }
  }
};

namespace {
thread_local std::string g_srv_err;
int srv_err(int code, const std::string& m) {
  g_srv_err = m;
  return code;
}
}  // namespace

extern "C" {

const char* aphpois_server_last_error(void) { return g_srv_err.c_str(); }

int aphpois_server_create(const aphpois_server_desc* d, aphpois_server** out) {
  if (!d || !out) return srv_err(APHPOIS_EINVAL, "NULL argument");
  *out = nullptr;
  aphpois_server* s = new (std::nothrow) aphpois_server();
  if (!s) return srv_err(APHPOIS_ENOMEM, "host allocation failed");
  s->desc = *d;
  const int rc = s->eng.init(d->device, APHPOIS_PCG, d->precision, APHPOIS_RES_REL_TRUE, d->tol,
                             0, d->maxiter, APHPOIS_NULL_AUTO, d->delta, d->check_every,
                             d->mem_budget, d->max_active > 0 ? d->max_active : 1024,
                             d->precond, d->mg_w);
  if (rc != APHPOIS_OK) {
    srv_err(rc, s->eng.err());
    delete s;
    return rc;
  }
  s->worker = std::thread([s] { s->loop(); });
  *out = s;
  return APHPOIS_OK;
}

int aphpois_server_submit(aphpois_server* s, const aphpois_request* r, long long* id) {
  if (!s || !r || !id) return srv_err(APHPOIS_EINVAL, "NULL argument");
  if (r->nx < 1 || r->ny < 1 || r->nz < 1 || !r->expr || r->nsteps < 1 || !(r->h3 > 0))
    return srv_err(APHPOIS_EINVAL, "bad request");
  auto q = std::make_shared<Req>();
  q->r = *r;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return srv_err(APHPOIS_ESTATE, "server stopped: " + s->fatal);
  q->id = s->next_id++;
  q->res.id = q->id;
  q->res.t_submit = now_s(s->t0);
  s->queue.push_back(q);
  s->reqs[q->id] = q;
  *id = q->id;
  s->cv_work.notify_one();
  return APHPOIS_OK;
}

int aphpois_server_wait(aphpois_server* s, long long id, double timeout_s,
                        aphpois_result* out) {
  if (!s) return srv_err(APHPOIS_EINVAL, "NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return srv_err(APHPOIS_EINVAL, "unknown request id");
  auto q = it->second;
  const auto pred = [&] { return q->done; };
  if (timeout_s < 0) {
    s->cv_done.wait(lk, pred);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
    return srv_err(APHPOIS_ESTATE, "timeout");
  }
  if (out) *out = q->res;
  s->reqs.erase(it);
  return APHPOIS_OK;
}

int aphpois_server_cancel(aphpois_server* s, long long id) {
  if (!s) return srv_err(APHPOIS_EINVAL, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return srv_err(APHPOIS_EINVAL, "unknown request id");
  it->second->cancel = true;
  s->cv_work.notify_one();
  return APHPOIS_OK;
}

int aphpois_server_get_stats(aphpois_server* s, aphpois_server_stats* st) {
  if (!s || !st) return srv_err(APHPOIS_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  st->queued = (int)s->queue.size();
  st->active = (int)s->by_slot.size();
  st->completed = s->completed;
  st->device_bytes = s->eng.bytes();
  st->time = now_s(s->t0);
  return APHPOIS_OK;
}

int aphpois_server_destroy(aphpois_server* s) {
  if (!s) return APHPOIS_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto& q : s->queue) {
      q->res.status = APHPOIS_ECANCEL;
      q->done = true;
    }
    s->queue.clear();
    s->cv_work.notify_all();
    s->cv_done.notify_all();
  }
  if (s->worker.joinable()) s->worker.join();
  delete s;
  return APHPOIS_OK;
}

}  // extern "C"
