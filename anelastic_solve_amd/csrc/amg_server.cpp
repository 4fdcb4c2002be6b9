// amg_server.cpp -- served Poisson solves with continuous batching.
//
// A worker thread owns one stream and advances every in-flight request by
// one V-cycle per round, with one launch per level step for all of them
// (Context::cycle_and_norms).  Each request has its own gate on the device
// (ItemGate): the round's last norm fold applies the stop rule to every
// request, and a request that stops has its kernels in later rounds return
// at once.  After a round, one copy brings every
// gate into a pinned mirror, followed by an event; the host waits for that
// event (not for the stream), replays the gates (Context::absorb_gate) and
// retires the finished requests (null-space removal, ghosts and the copy
// out, behind an event of their own).  A joining request's first defect
// initialises its gate on the device, so joining takes no host round trip.
// The code keeps kPipe rounds in flight; the stream is drained only when a
// member's hierarchy must be bound or evicted (the descriptor tables
// change), which happens once per member while the slots and the HBM budget
// hold every member.  A member keeps its warm-start pressure resident, and
// its hierarchy stays bound to a slot until the slot is needed by another
// member (least recently used).  Every per-request computation is
// independent of the rest of the batch (per-item kernels, per-item
// reductions), so a result is bitwise the same alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "amg.h"
#include "amg_solver.h"

using namespace amg;
using Clock = std::chrono::steady_clock;

namespace {

struct Member {
  amg_grid grid{};
  std::vector<double> metrics;       // owned copies of the grid arrays
  double* d_warm = nullptr;          // warm-start pressure, ghosted FP64
  size_t nghost = 0;
  Instance* inst = nullptr;          // bound hierarchy (may be evicted)
  Clock::time_point last_use{};
  bool busy = false;
};

struct Request {
  int id = 0;
  int member = 0;
  const double* d_rhs = nullptr;
  double* d_out = nullptr;
  Clock::time_point t_submit{}, t_done{};
  int iters = 0;
  double relres = 0.0;
  int status = AMG_OK;
  bool done = false;
};

}  // namespace

// Served rounds in flight (each with its own pinned copy of the gates).
// One: a round ahead would keep the device busy through the host's turn
// between rounds (tens of microseconds against rounds of milliseconds),
// but a request would then join and leave one round later; two and three
// rounds in flight were measured 5 and 10 % slower in solves/s at 1000
// requests/s on the served mix, with longer latencies.
constexpr int kPipe = 1;
// Rounds are launched directly, not captured as graphs: a capture in the
// worker thread makes a device-wide synchronisation in any other thread of
// the process fail (hipErrorStreamCaptureUnsupported, whatever the capture
// mode), and the caller cannot know when a round is being captured.

struct amg_server_s {
  Context ctx;
  ItemGate* h_mirror = nullptr;      // kPipe x kMaxInstances, pinned
  double budget = 0.0;
  size_t used = 0;
  int max_slots = 1, max_batch = 1;
  std::vector<Member> members;
  std::map<int, Request> reqs;
  std::deque<int> queue;             // waiting request ids
  std::vector<int> active;           // request ids in flight
  int next_id = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  int error = AMG_OK;
  std::thread worker;
  void run();
  int bind(Member& m);
};

// Binds a hierarchy to member m, evicting idle members' hierarchies (LRU)
// when the slot count or the HBM budget would be exceeded.
int amg_server_s::bind(Member& m) {
  if (m.inst) return AMG_OK;
  auto bound = [&]() {
    int n = 0;
    for (auto& x : members) n += x.inst != nullptr;
    return n;
  };
  // the hierarchy's size is known only after building it; estimate from the grid
  const double ncell = (double)m.grid.nx * m.grid.ny * m.grid.nz;
  const double w = ctx.cfg.precision == AMG_FP64 ? 8.0 : 4.0;
  const double est = ncell * (8.0 / 7.0) * w * 2.5 + ncell * 8.0 * (ctx.cfg.precision == AMG_FP64 ? 1.0 : 3.0);
  // warm-start pressures stay resident; a hierarchy that does not fit beside
  // them even with every other hierarchy evicted never will: the request fails
  double warm = 0.0;
  for (auto& x : members) warm += (double)x.nghost * 8.0;
  if (warm + est > budget) {
    set_error("amg_server: the member's hierarchy does not fit the HBM budget");
    return AMG_ENOMEM;
  }
  while (bound() >= max_slots || (double)used + est > budget) {
    Member* victim = nullptr;
    for (auto& x : members)
      if (x.inst && !x.busy && (!victim || x.last_use < victim->last_use)) victim = &x;
    if (!victim) return AMG_EBUSY;
    used -= victim->inst->device_bytes();
    ctx.remove_instance(victim->inst);
    victim->inst = nullptr;
  }
  Instance* I = nullptr;
  const int rc = ctx.add_instance(m.grid, &I);
  if (rc != AMG_OK) return rc;
  used += I->device_bytes();
  m.inst = I;
  return AMG_OK;
}

void amg_server_s::run() {
  struct Slot {
    int rid;
    Instance* inst;
    double* d_warm;
    const double* d_rhs;
    double* d_out;
    size_t nghost;
  };
  struct Round {
    hipEvent_t ev;
    int mirror;
    std::vector<int> rids;   // requests whose cycle the round ran
  };
  struct Leaving {
    hipEvent_t ev;
    std::vector<Slot> slots;
    int rc;
  };
  std::map<int, Slot> live;       // requests on the device, by id
  std::deque<Round> flight;
  std::deque<Leaving> leaving;
  std::vector<hipEvent_t> spare;
  int next_mirror = 0;
  int rc = AMG_OK;
  auto event = [&]() {
    hipEvent_t e = nullptr;
    if (!spare.empty()) {
      e = spare.back();
      spare.pop_back();
    } else if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      e = nullptr;
    }
    return e;
  };
  // requests whose copy-out completed (or every one after an error)
  auto complete = [&](bool wait) {
    while (!leaving.empty()) {
      Leaving& l = leaving.front();
      if (l.ev) {
        const hipError_t q = wait ? hipEventSynchronize(l.ev) : hipEventQuery(l.ev);
        if (q == hipErrorNotReady) return;
        if (q != hipSuccess && l.rc == AMG_OK) l.rc = AMG_EHIP;
        spare.push_back(l.ev);
      }
      const auto now = Clock::now();
      {
        std::lock_guard<std::mutex> g(mu);
        for (const Slot& x : l.slots) {
          Request& r = reqs[x.rid];
          Member& m = members[r.member];
          const Instance* I = x.inst;
          r.iters = I->converged ? I->iters : ctx.cfg.maxcy + 1;
          r.relres = I->relres;
          r.status = l.rc != AMG_OK ? l.rc
                     : (I->status != AMG_OK ? I->status : (I->converged ? AMG_OK : AMG_ENOTCONV));
          r.t_done = now;
          r.done = true;
          m.busy = false;
          m.last_use = now;
          active.erase(std::find(active.begin(), active.end(), x.rid));
        }
        if (l.rc != AMG_OK) error = l.rc;
      }
      cv_done.notify_all();
      leaving.pop_front();
    }
  };
  // the oldest round: wait for it, replay its gates, retire what finished
  auto land = [&]() {
    Round rd = std::move(flight.front());
    flight.pop_front();
    if (rc == AMG_OK && hipEventSynchronize(rd.ev) != hipSuccess) rc = AMG_EHIP;
    spare.push_back(rd.ev);
    std::vector<Slot> D;
    for (int rid : rd.rids) {
      auto it = live.find(rid);
      if (it == live.end()) continue;
      Instance* I = it->second.inst;
      if (rc == AMG_OK)
        rc = ctx.absorb_gate(I, h_mirror[(size_t)rd.mirror * kMaxInstances + I->gslot], nullptr);
      if (I->done || rc != AMG_OK) {
        D.push_back(it->second);
        live.erase(it);
      }
    }
    if (D.empty()) return;
    Leaving l{nullptr, D, rc};
    std::vector<Instance*> DI;
    for (const Slot& x : D) DI.push_back(x.inst);
    if (l.rc == AMG_OK) l.rc = ctx.finish(DI);
    for (const Slot& x : D)
      (void)hipMemcpyAsync(x.d_out, x.d_warm, x.nghost * 8, hipMemcpyDeviceToDevice, ctx.stream);
    l.ev = event();
    if (!l.ev || hipEventRecord(l.ev, ctx.stream) != hipSuccess) {
      if (l.ev) spare.push_back(l.ev);
      l.ev = nullptr;
      (void)hipStreamSynchronize(ctx.stream);
      if (l.rc == AMG_OK) l.rc = AMG_EHIP;
    }
    leaving.push_back(std::move(l));
  };
  auto drain = [&]() {
    while (!flight.empty()) land();
    complete(true);
  };

  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    if (flight.empty() && leaving.empty())
      cv_work.wait(lk, [&] { return stop || !queue.empty() || !live.empty(); });
    if (stop && live.empty() && flight.empty() && leaving.empty()) break;
    // admission, in arrival order; a member runs one request at a time
    // (its warm start is serial), so later requests of a busy member wait.
    // Binding changes the descriptor tables: only with nothing in flight.
    std::vector<int> joining;
    bool need_drain = false;
    for (size_t q = 0; q < queue.size() && (int)live.size() + (int)joining.size() < max_batch;) {
      const int rid = queue[q];
      Request& r = reqs[rid];
      Member& m = members[r.member];
      if (m.busy) { ++q; continue; }
      if (!m.inst && (!flight.empty() || !leaving.empty())) {
        need_drain = true;
        break;
      }
      const int brc = bind(m);
      if (brc == AMG_EBUSY) break;
      queue.erase(queue.begin() + q);
      if (brc != AMG_OK) {
        r.status = brc;
        r.done = true;
        r.t_done = Clock::now();
        cv_done.notify_all();
        continue;
      }
      m.busy = true;
      joining.push_back(rid);
    }
    std::vector<Slot> J;
    for (int rid : joining) {
      const Request& r = reqs[rid];
      const Member& m = members[r.member];
      J.push_back(Slot{rid, m.inst, m.d_warm, r.d_rhs, r.d_out, m.nghost});
      active.push_back(rid);
    }
    int nslot = 0;
    for (const Member& m : members)
      if (m.inst) nslot = std::max(nslot, m.inst->gslot + 1);
    lk.unlock();

    // join: RHS and warm start in, first defect, gates
    if (!J.empty() && rc == AMG_OK) {
      std::vector<Instance*> S;
      for (const Slot& x : J) {
        x.inst->d_phi = x.d_warm;
        x.inst->d_rhs = x.d_rhs;
        S.push_back(x.inst);
      }
      rc = ctx.prepare_async(S);
    }
    for (const Slot& x : J) live[x.rid] = x;
    // one round: one V-cycle of every live request (those that stopped on
    // the device return at once), then the gates to the host
    if (!live.empty() && rc == AMG_OK) {
      Round rd{event(), next_mirror, {}};
      next_mirror = (next_mirror + 1) % kPipe;
      std::vector<Instance*> C;
      for (auto& [rid, x] : live) {
        C.push_back(x.inst);
        rd.rids.push_back(rid);
      }
      rc = ctx.cycle_and_norms(C);
      if (rc == AMG_OK &&
          hipMemcpyAsync(h_mirror + (size_t)rd.mirror * kMaxInstances, ctx.d_igate,
                         sizeof(ItemGate) * nslot, hipMemcpyDeviceToHost, ctx.stream) != hipSuccess)
        rc = AMG_EHIP;
      if (rc == AMG_OK && (!rd.ev || hipEventRecord(rd.ev, ctx.stream) != hipSuccess)) rc = AMG_EHIP;
      if (rc == AMG_OK) flight.push_back(std::move(rd));
      else if (rd.ev) spare.push_back(rd.ev);
    }
    if (rc != AMG_OK) {
      // fail every live request
      (void)hipStreamSynchronize(ctx.stream);
      for (Round& r : flight) spare.push_back(r.ev);
      flight.clear();
      Leaving l{nullptr, {}, rc};
      for (auto& [rid, x] : live) l.slots.push_back(x);
      live.clear();
      leaving.push_back(std::move(l));
    }
    // keep kPipe - 1 rounds ahead of the host; with nothing left to enqueue,
    // land the rest; a member to bind needs the stream drained
    if (need_drain) {
      drain();
    } else {
      while (!flight.empty() && ((int)flight.size() >= kPipe || live.empty())) {
        // the copies out of retired requests precede the round in the stream
        if (!leaving.empty()) complete(true);
        land();
      }
      complete(live.empty() && flight.empty());
    }
    lk.lock();
  }
  for (hipEvent_t e : spare) (void)hipEventDestroy(e);
}

extern "C" {

int amg_server_create(const amg_config* cfg, double hbm_budget_bytes, int max_slots, int max_batch,
                      amg_server* out) {
  if (!cfg || !out || max_slots < 1 || max_batch < 1 || !(hbm_budget_bytes > 0.0)) {
    set_error("amg_server_create: bad argument");
    return AMG_EINVAL;
  }
  *out = nullptr;
  amg_server_s* s = new amg_server_s();
  amg_config c = *cfg;
  c.use_graph = 0;
  c.verbose = 0;
  const int rc = s->ctx.init(c);
  if (rc != AMG_OK) {
    delete s;
    return rc;
  }
  if (hipHostMalloc((void**)&s->h_mirror, sizeof(ItemGate) * kMaxInstances * kPipe,
                    hipHostMallocDefault) != hipSuccess) {
    s->h_mirror = nullptr;
    delete s;
    set_error("amg_server_create: pinned allocation failed");
    return AMG_ENOMEM;
  }
  s->budget = hbm_budget_bytes;
  s->max_slots = max_slots;
  s->max_batch = std::min(max_batch, kMaxItems);
  s->worker = std::thread([s] { s->run(); });
  *out = s;
  return AMG_OK;
}

int amg_server_destroy(amg_server s) {
  if (!s) return AMG_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->queue.clear();
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (auto& m : s->members)
    if (m.d_warm) (void)hipFree(m.d_warm);
  if (s->h_mirror) (void)hipHostFree(s->h_mirror);
  delete s;
  return AMG_OK;
}

int amg_server_add_member(amg_server s, const amg_grid* g) {
  if (!s || !g) {
    set_error("amg_server_add_member: NULL argument");
    return AMG_EINVAL;
  }
  Member m;
  m.grid = *g;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  // keep owned copies of the metric arrays
  const double* src[12] = {g->xe, g->xc, g->dxe, g->dxc, g->ye, g->yc, g->dye, g->dyc,
                           g->ze, g->zc, g->dze, g->dzc};
  const int len[12] = {nx + 2, nx + 2, nx + 2, nx + 2, ny + 2, ny + 2, ny + 2, ny + 2,
                       nz + 2, nz + 2, nz + 2, nz + 2};
  for (int q = 0; q < 12; ++q) {
    if (!src[q]) {
      set_error("amg_server_add_member: missing metric array");
      return AMG_EINVAL;
    }
    m.metrics.insert(m.metrics.end(), src[q], src[q] + len[q]);
  }
  if (g->rho) m.metrics.insert(m.metrics.end(), g->rho, g->rho + nz + 2);
  double* p = m.metrics.data();
  const double** dst[12] = {&m.grid.xe, &m.grid.xc, &m.grid.dxe, &m.grid.dxc, &m.grid.ye, &m.grid.yc,
                            &m.grid.dye, &m.grid.dyc, &m.grid.ze, &m.grid.zc, &m.grid.dze, &m.grid.dzc};
  for (int q = 0; q < 12; ++q) { *dst[q] = p; p += len[q]; }
  m.grid.rho = g->rho ? p : nullptr;
  m.nghost = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  std::lock_guard<std::mutex> lk(s->mu);
  if ((double)s->used + (double)m.nghost * 8 > s->budget) {
    set_error("amg_server_add_member: HBM budget exhausted");
    return AMG_ENOMEM;
  }
  if (hipMalloc((void**)&m.d_warm, m.nghost * 8) != hipSuccess) {
    set_error("amg_server_add_member: device allocation failed");
    return AMG_ENOMEM;
  }
  if (hipMemset(m.d_warm, 0, m.nghost * 8) != hipSuccess) {
    (void)hipFree(m.d_warm);
    set_error("amg_server_add_member: hipMemset failed");
    return AMG_EHIP;
  }
  s->used += m.nghost * 8;
  s->members.push_back(std::move(m));
  // fix the pointers after the move (vector storage moved with the member)
  Member& mm = s->members.back();
  p = mm.metrics.data();
  const double** dst2[12] = {&mm.grid.xe, &mm.grid.xc, &mm.grid.dxe, &mm.grid.dxc, &mm.grid.ye,
                             &mm.grid.yc, &mm.grid.dye, &mm.grid.dyc, &mm.grid.ze, &mm.grid.zc,
                             &mm.grid.dze, &mm.grid.dzc};
  for (int q = 0; q < 12; ++q) { *dst2[q] = p; p += len[q]; }
  mm.grid.rho = g->rho ? p : nullptr;
  return (int)s->members.size() - 1;
}

int amg_server_submit(amg_server s, int member, const double* d_rhs, double* d_out) {
  if (!s || !d_rhs || !d_out) {
    set_error("amg_server_submit: NULL argument");
    return AMG_EINVAL;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (member < 0 || member >= (int)s->members.size()) {
    set_error("amg_server_submit: unknown member");
    return AMG_EINVAL;
  }
  if (s->error != AMG_OK) return s->error;
  Request r;
  r.id = s->next_id++;
  r.member = member;
  r.d_rhs = d_rhs;
  r.d_out = d_out;
  r.t_submit = Clock::now();
  s->reqs[r.id] = r;
  s->queue.push_back(r.id);
  s->cv_work.notify_all();
  return r.id;
}

int amg_server_poll(amg_server s, int req, int* iters, double* relres, double* latency_ms) {
  if (!s) return AMG_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(req);
  if (it == s->reqs.end()) {
    set_error("amg_server_poll: unknown request");
    return AMG_EINVAL;
  }
  const Request& r = it->second;
  if (!r.done) return 0;
  if (iters) *iters = r.iters;
  if (relres) *relres = r.relres;
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
  return r.status == AMG_OK ? 1 : r.status;
}

int amg_server_drain(amg_server s) {
  if (!s) return AMG_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->queue.empty() && s->active.empty(); });
  return s->error;
}

}  // extern "C"
