// Continuous-batching request server (mixed or FP64 PCG) on one GPU.
//
// Requests are routed by size to lanes: systems of at most kLaneSplitCells
// cells to the small lane, larger ones to the large lane.  Each lane owns one
// Engine (its own HIP stream and captured iteration graph) and a worker
// thread that loops:
//   1. admit queued requests (FIFO) while a slot is free and the device
//      memory budget (shared by the lanes) allows their resident storage:
//      coefficients and Krylov vectors of every admitted system stay in HBM
//      until its request completes;
//   2. run one replay of the captured iteration graph for the lane's active
//      systems;
//   3. for every system that finished its current time step: start the next
//      step (new right-hand side, warm start from the current solution) in
//      the same slot, or complete the request and free its storage.
// Requests that arrive while others run are admitted between replays, so
// their iterations share the same kernel launches; the two lanes' streams
// run concurrently on the device, so a small system does not advance at the
// pace of the large ones it would otherwise share every launch with.  A
// system's result does not depend on which systems share its launches.
#include <atomic>
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

// 128^3 = 2^21 cells: systems up to this size form the small lane
constexpr long long kLaneSplitCells = 1LL << 21;
constexpr int kLanes = 2;

struct Req {
  long long id = 0;
  aphpois_request r{};
  int lane = 0;
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

struct aphpois_server;

// One size class: an engine and the worker thread that alone touches it.
struct ServerLane {
  aphpois_server* srv = nullptr;
  int index = 0;
  aph::Engine eng;
  std::deque<std::shared_ptr<Req>> queue;          // guarded by srv->mu
  std::map<int, std::shared_ptr<Req>> by_slot;     // written under srv->mu
  std::condition_variable cv_work;
  std::atomic<long long> dev_bytes{0};
  long long mem_epoch_seen = -1;  // admission blocked by the budget at this epoch
  std::thread worker;

  void finish(const std::shared_ptr<Req>& q, int status);
  bool start_step(const std::shared_ptr<Req>& q);
  void admit();
  void loop();
};

struct aphpois_server {
  aphpois_server_desc desc{};
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::mutex mu;
  std::condition_variable cv_done;
  std::condition_variable cv_quiet;  // waiters drained (destroy)
  std::map<long long, std::shared_ptr<Req>> reqs;  // all not yet collected
  // declared before the lanes: their engines release into `used` when they
  // are destroyed
  std::atomic<long long> used{0};       // device bytes reserved by all lanes
  std::atomic<long long> mem_epoch{0};  // bumped whenever a lane frees storage
  // held by a lane for every engine call except its graph replays, and by an
  // engine while it captures (Engine::share_api_lock): lock order api_mu, mu
  std::mutex api_mu;
  ServerLane lanes[kLanes];
  long long next_id = 1;
  bool stop = false;
  std::string fatal;
  long long completed = 0;
  int waiters = 0;  // threads inside aphpois_server_wait

  int resident_locked() const {
    int n = 0;
    for (const ServerLane& l : lanes) n += (int)l.by_slot.size();
    return n;
  }
  void memory_freed() {
    mem_epoch.fetch_add(1);
    std::lock_guard<std::mutex> lk(mu);
    for (ServerLane& l : lanes) l.cv_work.notify_one();
  }
};

void ServerLane::finish(const std::shared_ptr<Req>& q, int status) {
  q->res.status = status;
  q->res.t_done = now_s(srv->t0);
  q->res.steps_done = q->step;
  const int slot = q->slot;
  if (slot >= 0) eng.remove(slot);
  dev_bytes = eng.bytes();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (slot >= 0) by_slot.erase(slot);
    q->slot = -1;
    q->done = true;
    ++srv->completed;
    srv->cv_done.notify_all();
  }
  if (slot >= 0) srv->memory_freed();
}

// Loads the next step of an admitted request; returns false on error.
bool ServerLane::start_step(const std::shared_ptr<Req>& q) {
  const aphpois_request& r = q->r;
  const long long ly = r.nx, lz = (long long)r.nx * r.ny;
  int rc = APHPOIS_OK;
  if (q->step > 0) {
    const double* e7 = r.e7 ? r.e7[q->step] : nullptr;
    if (e7) {
      rc = eng.set_rhs(q->slot, e7, false, 0, ly, lz);
    } else {
      const unsigned long long key = r.seed * 0x9E3779B97F4A7C15ull + (unsigned long long)q->step;
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

void ServerLane::admit() {
  for (;;) {
    std::shared_ptr<Req> q;
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      if (queue.empty()) return;
      if (srv->desc.max_active > 0 && srv->resident_locked() >= srv->desc.max_active) return;
      q = queue.front();
      if (q->cancel) {
        queue.pop_front();
        q->done = true;
        q->res.status = APHPOIS_ECANCEL;
        q->res.t_done = now_s(srv->t0);
        srv->cv_done.notify_all();
        continue;
      }
    }
    // after a refusal by the budget, try again only once storage was freed
    const long long epoch = srv->mem_epoch.load();
    if (mem_epoch_seen == epoch) return;
    const aphpois_request& r = q->r;
    int slot = -1;
    int rc = eng.add_system(r.nx, r.ny, r.nz, r.h3, &slot);
    if (rc == APHPOIS_OK)
      rc = eng.set_system(slot, r.expr, false, 0, r.nx, (long long)r.nx * r.ny);
    if (rc == APHPOIS_EBUDGET) {
      if (slot >= 0) eng.remove(slot);
      dev_bytes = eng.bytes();
      bool others;
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        others = srv->resident_locked() > 0;
      }
      if (others) {
        mem_epoch_seen = epoch;  // wait for memory to free up
        return;
      }
      // nothing resident anywhere: the request can never fit
    }
    mem_epoch_seen = -1;
    dev_bytes = eng.bytes();
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      queue.pop_front();
    }
    if (rc != APHPOIS_OK) {
      if (slot >= 0 && rc != APHPOIS_EBUDGET) eng.remove(slot);
      dev_bytes = eng.bytes();
      q->slot = -1;
      q->res.t_admit = now_s(srv->t0);
      finish(q, rc);
      continue;
    }
    q->slot = slot;
    q->res.t_admit = now_s(srv->t0);
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      by_slot[slot] = q;
    }
    start_step(q);
  }
}

void ServerLane::loop() {
  std::vector<int> fin;
  while (true) {
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      cv_work.wait(lk, [&] {
        if (srv->stop || !by_slot.empty()) return true;
        return !queue.empty() && mem_epoch_seen != srv->mem_epoch.load();
      });
      if (srv->stop) break;
    }
    {
      std::lock_guard<std::mutex> al(srv->api_mu);
      admit();
      // cancellation of admitted requests
      std::vector<std::shared_ptr<Req>> cancelled;
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        for (auto& kv : by_slot)
          if (kv.second->cancel) cancelled.push_back(kv.second);
      }
      for (auto& q : cancelled) finish(q, APHPOIS_ECANCEL);
    }
    if (by_slot.empty()) continue;
    // the replays run without api_mu (step takes it to rebuild and capture)
    const int rc = eng.step(1, &fin);
    if (rc != APHPOIS_OK) {
      // the device state is not trusted: fail this lane's requests and stop
      // the server (the other lanes fail theirs when they see stop)
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->fatal = eng.err();
      for (auto& kv : by_slot) {
        kv.second->res.status = rc;
        kv.second->res.t_done = now_s(srv->t0);
        kv.second->done = true;
      }
      by_slot.clear();
      for (auto& q : queue) {
        q->res.status = rc;
        q->done = true;
      }
      queue.clear();
      srv->stop = true;
      for (ServerLane& l : srv->lanes) l.cv_work.notify_all();
      srv->cv_done.notify_all();
      break;
    }
    std::lock_guard<std::mutex> al(srv->api_mu);
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

namespace {
thread_local std::string g_srv_err;
int srv_err(int code, const std::string& m) {
  g_srv_err = m;
  return code;
}

int lane_of(const aphpois_request* r) {
  const long long n = (long long)r->nx * r->ny * r->nz;
  return n <= kLaneSplitCells ? 0 : 1;
}

// Stops the workers, fails whatever is still admitted or queued with
// `status`, and waits until no thread is inside aphpois_server_wait.
void shutdown(aphpois_server* s, int status) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (ServerLane& l : s->lanes) l.cv_work.notify_all();
  }
  for (ServerLane& l : s->lanes)
    if (l.worker.joinable()) l.worker.join();
  std::unique_lock<std::mutex> lk(s->mu);
  for (ServerLane& l : s->lanes) {
    for (auto& q : l.queue) {
      q->res.status = status;
      q->done = true;
    }
    l.queue.clear();
    for (auto& kv : l.by_slot) {
      kv.second->res.status = status;
      kv.second->res.t_done = now_s(s->t0);
      kv.second->done = true;
    }
    l.by_slot.clear();
  }
  s->cv_done.notify_all();
  s->cv_quiet.wait(lk, [&] { return s->waiters == 0; });
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
  for (int i = 0; i < kLanes; ++i) {
    ServerLane& l = s->lanes[i];
    l.srv = s;
    l.index = i;
    l.eng.share_budget(&s->used);
    l.eng.share_api_lock(&s->api_mu);
    const int rc = l.eng.init(d->device, APHPOIS_PCG, d->precision, APHPOIS_RES_REL_TRUE,
                              d->tol, 0, d->maxiter, APHPOIS_NULL_AUTO, d->delta,
                              d->check_every, d->mem_budget,
                              d->max_active > 0 ? d->max_active : 1024, d->precond, d->mg_w);
    if (rc != APHPOIS_OK) {
      srv_err(rc, l.eng.err());
      delete s;
      return rc;
    }
  }
  for (ServerLane& l : s->lanes) l.worker = std::thread([&l] { l.loop(); });
  *out = s;
  return APHPOIS_OK;
}

int aphpois_server_submit(aphpois_server* s, const aphpois_request* r, long long* id) {
  if (!s || !r || !id) return srv_err(APHPOIS_EINVAL, "NULL argument");
  if (r->nx < 1 || r->ny < 1 || r->nz < 1 || !r->expr || r->nsteps < 1 || !(r->h3 > 0))
    return srv_err(APHPOIS_EINVAL, "bad request");
  auto q = std::make_shared<Req>();
  q->r = *r;
  q->lane = lane_of(r);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return srv_err(APHPOIS_ESTATE, "server stopped: " + s->fatal);
  q->id = s->next_id++;
  q->res.id = q->id;
  q->res.t_submit = now_s(s->t0);
  s->lanes[q->lane].queue.push_back(q);
  s->reqs[q->id] = q;
  *id = q->id;
  s->lanes[q->lane].cv_work.notify_one();
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
  ++s->waiters;
  bool ok = true;
  if (timeout_s < 0) {
    s->cv_done.wait(lk, pred);
  } else {
    ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
  }
  if (ok) {
    if (out) *out = q->res;
    s->reqs.erase(q->id);
  }
  // last touch of *s before the lock is released: destroy() frees the server
  // once no thread is left in here
  if (--s->waiters == 0) s->cv_quiet.notify_all();
  lk.unlock();
  if (!ok) return srv_err(APHPOIS_ESTATE, "timeout");
  return APHPOIS_OK;
}

int aphpois_server_cancel(aphpois_server* s, long long id) {
  if (!s) return srv_err(APHPOIS_EINVAL, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return srv_err(APHPOIS_EINVAL, "unknown request id");
  it->second->cancel = true;
  s->lanes[it->second->lane].cv_work.notify_one();
  return APHPOIS_OK;
}

int aphpois_server_get_stats(aphpois_server* s, aphpois_server_stats* st) {
  if (!s || !st) return srv_err(APHPOIS_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  st->queued = 0;
  st->device_bytes = 0;
  for (const ServerLane& l : s->lanes) {
    st->queued += (int)l.queue.size();
    st->device_bytes += l.dev_bytes.load();
  }
  st->active = s->resident_locked();
  st->completed = s->completed;
  st->time = now_s(s->t0);
  return APHPOIS_OK;
}

int aphpois_server_destroy(aphpois_server* s) {
  if (!s) return APHPOIS_OK;
  shutdown(s, APHPOIS_ECANCEL);
  delete s;
  return APHPOIS_OK;
}

}  // extern "C"
