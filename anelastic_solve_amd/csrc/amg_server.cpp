// amg_server.cpp -- served Poisson solves with continuous batching.
//
// A worker thread owns one stream and advances every in-flight request by
// one V-cycle per round, with one launch per level step for all of them
// (Context::cycle_and_norms).  Requests that arrive while others run are
// admitted at the next round boundary when a slot and HBM budget are free;
// finished requests leave at once.  Each ensemble member keeps its
// warm-start pressure resident, and its multigrid hierarchy stays bound to
// a slot until the slot is needed by another member (least recently used).
// Every per-request computation is independent of the rest of the batch
// (per-item kernels, per-item reductions), so a result is bitwise the same
// alone and in the mix.
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

struct amg_server_s {
  Context ctx;
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
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    cv_work.wait(lk, [&] { return stop || !queue.empty() || !active.empty(); });
    if (stop && active.empty()) break;
    // admission, in arrival order; a member runs one request at a time
    // (its warm start is serial), so later requests of a busy member wait
    std::vector<int> joining;
    for (size_t q = 0; q < queue.size() && (int)active.size() + (int)joining.size() < max_batch;) {
      const int rid = queue[q];
      Request& r = reqs[rid];
      Member& m = members[r.member];
      if (m.busy) { ++q; continue; }
      const int rc = bind(m);
      if (rc == AMG_EBUSY) break;
      queue.erase(queue.begin() + q);
      if (rc != AMG_OK) {
        r.status = rc;
        r.done = true;
        r.t_done = Clock::now();
        cv_done.notify_all();
        continue;
      }
      m.busy = true;
      joining.push_back(rid);
    }
    // snapshot what the device work needs (members/reqs may grow meanwhile)
    auto slot_of = [&](int rid) {
      const Request& r = reqs[rid];
      const Member& m = members[r.member];
      return Slot{rid, m.inst, m.d_warm, r.d_rhs, r.d_out, m.nghost};
    };
    std::vector<Slot> J, A;
    for (int rid : joining) {
      J.push_back(slot_of(rid));
      active.push_back(rid);
    }
    for (int rid : active) A.push_back(slot_of(rid));
    lk.unlock();

    // join: RHS and warm start in, initial residual
    std::vector<Instance*> S;
    for (const Slot& x : J) {
      x.inst->d_phi = x.d_warm;
      x.inst->d_rhs = x.d_rhs;
      S.push_back(x.inst);
    }
    int rc = S.empty() ? AMG_OK : ctx.prepare(S);
    // one round: one V-cycle of every in-flight request, one launch per level step
    std::vector<Instance*> C;
    for (const Slot& x : A)
      if (!x.inst->done) C.push_back(x.inst);
    if (rc == AMG_OK && !C.empty()) {
      rc = ctx.cycle_and_norms(C, false);
      if (rc == AMG_OK && hipStreamSynchronize(ctx.stream) != hipSuccess) rc = AMG_EHIP;
      if (rc == AMG_OK) ctx.post_norms(C);
    }
    // leave: finished requests write their solution (warm start kept resident)
    std::vector<Instance*> D;
    std::vector<Slot> DS;
    for (const Slot& x : A)
      if (x.inst->done || rc != AMG_OK) { D.push_back(x.inst); DS.push_back(x); }
    if (!D.empty()) {
      if (rc == AMG_OK) rc = ctx.finish(D);
      for (const Slot& x : DS)
        (void)hipMemcpyAsync(x.d_out, x.d_warm, x.nghost * 8, hipMemcpyDeviceToDevice, ctx.stream);
      if (hipStreamSynchronize(ctx.stream) != hipSuccess && rc == AMG_OK) rc = AMG_EHIP;
    }
    const auto now = Clock::now();
    lk.lock();
    for (const Slot& x : DS) {
      Request& r = reqs[x.rid];
      Member& m = members[r.member];
      Instance* I = x.inst;
      r.iters = I->converged ? I->iters : ctx.cfg.maxcy + 1;
      r.relres = I->relres;
      r.status = rc != AMG_OK ? rc
                 : (I->status != AMG_OK ? I->status : (I->converged ? AMG_OK : AMG_ENOTCONV));
      r.t_done = now;
      r.done = true;
      m.busy = false;
      m.last_use = now;
      active.erase(std::find(active.begin(), active.end(), x.rid));
    }
    if (rc != AMG_OK) error = rc;
    if (!DS.empty()) cv_done.notify_all();
  }
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
