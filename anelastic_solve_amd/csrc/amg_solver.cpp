// amg_solver.cpp -- device state, V-cycle step lists and the FP64 outer loop.
//
// The cycle order follows mgMK/mg_cycle.F90 (pre-smooth and restrict down
// to the coarsest level, solve it, then prolongation-add and post-smooth up)
// and the driver follows mgMK/mg_solver.F90 (loop of at most maxcy cycles
// with a convergence test after each, ghosts filled on return).  Every step
// is one launch over all instances of a set that take part in it.
#include "amg_solver.h"

#include "amg_direct.h"

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>

namespace amg {

namespace {
thread_local std::string g_last_error;

inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

// d_out/h_out hold four regions of kOutRegion doubles: [0] per-instance norms
// (4 per instance), [1] unused, [2] relmax changes (2 per instance),
// [3] weighted sums (1 per instance).
constexpr size_t kOutRegion = (size_t)kMaxInstances * 4;

// Launch fn(ll, item_offset, block_offset) over chunks of at most kMaxItems items.
void for_lists(const std::vector<Item>& items,
               const std::function<void(const LaunchList&, int, int)>& fn) {
  LaunchList ll;
  ll.nitems = 0;
  ll.nblocks = 0;
  ll.prefix[0] = 0;
  int item_off = 0, block_off = 0;
  for (const Item& it : items) {
    if (ll.nitems == kMaxItems) {
      fn(ll, item_off, block_off);
      item_off += ll.nitems;
      block_off += ll.nblocks;
      ll.nitems = 0;
      ll.nblocks = 0;
    }
    ll.items[ll.nitems] = it.idx;
    ll.nblocks += std::max(it.blocks, 0);
    ll.prefix[++ll.nitems] = ll.nblocks;
  }
  if (ll.nitems > 0) fn(ll, item_off, block_off);
}
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }

void* DevArena::alloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 8;
  if (hipMalloc(&p, n) != hipSuccess) return nullptr;
  ptrs.push_back(p);
  bytes += n;
  return p;
}

void DevArena::release() {
  for (void* p : ptrs) (void)hipFree(p);
  ptrs.clear();
  bytes = 0;
}

size_t partial_blocks(const Instance& I) {
  const HostLevel& F = I.H.lv[0];
  const long long n = F.ncell();
  long long b = march_blocks(F.nx / 2, F.ny, F.nz);
  b = std::max(b, (long long)cdiv(n, (long long)kBlock * kFinePer));
  b = std::max(b, (long long)cdiv(n, (long long)kBlock * kWsumPer));
  b = std::max(b, (long long)cdiv(n, kBlock));
  return (size_t)b + 1;
}

size_t cycle_bytes(const Instance& I, const std::vector<DevXfer>& X) {
  // X: the instance's own transfers, X[l] from level l to level l+1
  const double w = I.fp64 ? 8.0 : 4.0;
  double b = 0.0;
  const int L = (int)I.H.lv.size();
  // levels with smoothing and transfers: above the direct level, or every
  // level but the coarsest; a direct level reads f and writes u.  The
  // restriction reads u and b (2w per fine cell, w/8 written); the
  // prolongation reads and writes u (2w, w/8 read), or half of it when the
  // level's post-smoothing rewrites the other half (DevXfer::skip, 2:2:2);
  // the marching restriction skips b and the outside neighbours of the
  // cells whose residual the pre-smoothing solved (DevXfer::rskip: w/2
  // less).  Both apply to the V-cycle's own transfers, above the coarse tail.
  const int top = I.ld >= 0 ? I.ld : L - 1;
  for (int l = 0; l < top; ++l) {
    const DevXfer& x = X[(size_t)l];
    const bool cyc = l < I.lc();
    const double xfer = (cyc && x.rskip != kXferSkipNone ? 1.625 : 2.125) +
                        (cyc && x.skip != kXferSkipNone ? 1.125 : 2.125);
    b += w * (double)I.H.lv[l].ncell() * (3.0 * (I.pre[l] + I.post[l]) + xfer);
  }
  if (I.ld >= 0) b += 2.0 * w * (double)I.H.lv[I.ld].ncell();
  const double n0 = (double)I.H.lv[0].ncell();
  b += I.fp64 ? 16.0 * n0 : 32.0 * n0;
  return (size_t)b;
}

int Context::init(const amg_config& c) {
  cfg = c;
  // the relmax stop rule needs the solution change, which an FP64 solve
  // computes only when it reports relmax
  if (cfg.stop_rule == AMG_STOP_RELMAX && cfg.precision == AMG_FP64) cfg.report_relmax = 1;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    set_error("no HIP device visible");
    return AMG_ENOGPU;
  }
  AMG_HIP(hipGetDevice(&device));
  AMG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  AMG_HIP(hipHostMalloc((void**)&h_out, sizeof(double) * kOutRegion * 4, hipHostMallocDefault));
  AMG_HIP(hipMalloc((void**)&d_out, sizeof(double) * kOutRegion * 4));
  AMG_HIP(hipMalloc((void**)&d_igate, sizeof(ItemGate) * kMaxInstances));
  AMG_HIP(hipHostMalloc((void**)&h_igate, sizeof(ItemGate) * kMaxInstances, hipHostMallocDefault));
  for (auto& e : ev) AMG_HIP(hipEventCreate(&e));
  return AMG_OK;
}

void Context::drop_graphs() {
  for (auto& g : gchain)
    if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

Context::~Context() {
  if (stream) (void)hipStreamSynchronize(stream);
  drop_graphs();
  for (Instance* I : inst) delete I;
  inst.clear();
  if (dL) (void)hipFree(dL);
  if (dX) (void)hipFree(dX);
  if (dT) (void)hipFree(dT);
  if (d_partial) (void)hipFree(d_partial);
  if (d_out) (void)hipFree(d_out);
  if (h_out) (void)hipHostFree(h_out);
  if (d_igate) (void)hipFree(d_igate);
  if (h_igate) (void)hipHostFree(h_igate);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

// The device's state words (DevLevel::stop, warm, skipc, pcur: written by
// the gates) into the host's table, so that a re-upload of the table keeps
// them for solves in progress (a server binds and evicts hierarchies
// between the cycles of its running requests).  No work of the context may
// be in flight.
int Context::pull_state() {
  if (!dL || hL.empty()) return AMG_OK;
  std::vector<DevLevel> dev(hL.size());
  AMG_HIP(hipStreamSynchronize(stream));
  AMG_HIP(hipMemcpy(dev.data(), dL, hL.size() * sizeof(DevLevel), hipMemcpyDeviceToHost));
  for (size_t l = 0; l < hL.size(); ++l) {
    hL[l].stop = dev[l].stop;
    hL[l].warm = dev[l].warm;
    hL[l].skipc = dev[l].skipc;
    hL[l].pcur = dev[l].pcur;
    hL[l].bshift = dev[l].bshift;
  }
  return AMG_OK;
}

// Removes instance I and compacts the descriptor tables to the live
// instances' entries (re-indexed in instance order), so that a server that
// binds and evicts hierarchies for ever keeps tables of bounded size.  The
// device tables are re-uploaded, with the live instances' state words; no
// work of the context may be in flight.
void Context::remove_instance(Instance* I) {
  (void)pull_state();
  auto it = std::find(inst.begin(), inst.end(), I);
  if (it != inst.end()) inst.erase(it);
  delete I;
  std::vector<DevLevel> nL;
  std::vector<DevXfer> nX;
  std::vector<DevTail> nT;
  for (Instance* J : inst) {
    std::vector<int> olds(J->lev.begin(), J->lev.end());
    if (!J->fp64) {
      olds.push_back(J->fine[0]);
      olds.push_back(J->fine[1]);
    }
    auto remap = [&](int o) {
      for (size_t q = 0; q < olds.size(); ++q)
        if (olds[q] == o) return (int)nL.size() + (int)q;
      return -1;
    };
    const int xb = (int)nX.size();
    std::vector<int> oldx(J->xf.begin(), J->xf.end());
    auto remapx = [&](int o) {
      for (size_t q = 0; q < oldx.size(); ++q)
        if (oldx[q] == o) return xb + (int)q;
      return -1;
    };
    std::vector<DevLevel> mine;
    for (int o : olds) {
      DevLevel d = hL[o];
      if (d.link >= 0) d.link = remap(d.link);
      mine.push_back(d);
    }
    for (int o : oldx) {
      DevXfer x = hX[o];
      x.fine = remap(x.fine);
      x.coarse = remap(x.coarse);
      nX.push_back(x);
    }
    DevTail t = hT[J->tail];
    for (int d = 0; d < t.nlev; ++d) {
      t.lev[d] = remap(t.lev[d]);
      if (d + 1 < t.nlev) t.xf[d] = remapx(t.xf[d]);
    }
    for (int& l : J->lev) l = remap(l);
    J->fine[0] = remap(J->fine[0]);
    J->fine[1] = remap(J->fine[1]);
    for (int& x : J->xf) x = remapx(x);
    J->tail = (int)nT.size();
    nT.push_back(t);
    nL.insert(nL.end(), mine.begin(), mine.end());
  }
  hL.swap(nL);
  hX.swap(nX);
  hT.swap(nT);
  if (!inst.empty()) (void)upload_tables();
  else drop_graphs();
}

// The new instance's descriptors are collected in local tables and appended
// to hL/hX/hT only once everything is allocated and uploaded: an error on
// the way frees the instance (and its device memory) and leaves the
// context's tables as they were.
int check_grid(const amg_config& cfg, const amg_grid& g, Hierarchy& H) {
  std::string err;
  const int rc = build_hierarchy(cfg, g, H, err);
  if (rc != AMG_OK) {
    set_error(err);
    return rc;
  }
  if (cfg.smoother == AMG_SMOOTHER_LJ && g.nz > kZSeg * kZMaxSegs) {
    set_error("amg_create: the z-line smoother (LJ) supports nz <= " +
              std::to_string(kZSeg * kZMaxSegs) + " (got nz=" + std::to_string(g.nz) +
              "); use RBPGS or split the column");
    return AMG_EBADSIZE;
  }
  return AMG_OK;
}

int Context::add_instance(const amg_grid& g, Instance** out, Hierarchy* built) {
  {
    const int rc = pull_state();
    if (rc != AMG_OK) return rc;
  }
  std::unique_ptr<Instance> hold(new Instance());
  Instance* I = hold.get();
  if (built) {
    I->H = std::move(*built);
  } else {
    const int rc = check_grid(cfg, g, I->H);
    if (rc != AMG_OK) return rc;
  }
  I->id = inst.empty() ? 0 : inst.back()->id + 1;
  // the lowest gate slot no live instance holds
  {
    std::vector<char> used(kMaxInstances, 0);
    for (const Instance* J : inst) used[J->gslot] = 1;
    I->gslot = -1;
    for (int q = 0; q < kMaxInstances && I->gslot < 0; ++q)
      if (!used[q]) I->gslot = q;
    if (I->gslot < 0) {
      set_error("amg: too many instances on one context");
      return AMG_ENOMEM;
    }
  }
  I->nx = g.nx; I->ny = g.ny; I->nz = g.nz;
  I->fp64 = (cfg.precision == AMG_FP64);
  const int L = (int)I->H.lv.size();
  sweep_counts(cfg, L, I->pre, I->post);
  const size_t w = I->fp64 ? 8 : 4;
  auto fail = [&]() {
    set_error("device allocation failed");
    return AMG_ENOMEM;
  };
  // coefficient tables, volumes, density, folds
  std::vector<DevLevel> ents(L);
  for (int l = 0; l < L; ++l) {
    const HostLevel& H = I->H.lv[l];
    DevLevel& d = ents[l];
    std::memset(&d, 0, sizeof d);
    d.nx = H.nx; d.ny = H.ny; d.nz = H.nz; d.hx = H.nx / 2;
    auto magic = [](unsigned v) -> unsigned {
      return v <= 1 ? 0u : (unsigned)((0x100000000ull + v - 1) / v);
    };
    d.mhx = magic((unsigned)d.hx);
    d.mny = magic((unsigned)d.ny);
    d.gmhx = gm_magic((unsigned)d.hx, &d.gshx);
    d.gmny = gm_magic((unsigned)d.ny, &d.gsny);
    d.gmzp = gm_magic((unsigned)((d.nz + 15) & ~15), &d.gszp);
    d.px = I->H.bc[0][0] == AMG_BC_PERIODIC;
    d.py = I->H.bc[1][0] == AMG_BC_PERIODIC;
    d.pz = I->H.bc[2][0] == AMG_BC_PERIODIC;
    d.inst = I->id;
    d.gslot = I->gslot;
    d.ncell = (int)H.ncell();
    d.link = -1;
    d.bcx0 = I->H.bc[0][0]; d.bcx1 = I->H.bc[0][1];
    d.bcy0 = I->H.bc[1][0]; d.bcy1 = I->H.bc[1][1];
    d.bcz0 = I->H.bc[2][0]; d.bcz1 = I->H.bc[2][1];
    d.vsum = I->H.vol_sum;
    d.vrsum = I->H.vol_rho_sum;
    d.nullsp = I->H.singular ? 1 : 0;
    // the tables end with kZSeg zeros: the z-line sweep reads the rows of a
    // whole segment from its first row on, past nz in the last one
    const size_t nco = H.co.size() + kZSeg;
    double* c64 = (double*)I->mem.alloc(nco * 8);
    float* c32 = (float*)I->mem.alloc(nco * 4);
    double* vol = (double*)I->mem.alloc(H.vol.size() * 8);
    double* rho = (double*)I->mem.alloc(H.nz * 8);
    if (!c64 || !c32 || !vol || !rho) return fail();
    std::vector<double> co64(H.co.begin(), H.co.end());
    co64.resize(nco, 0.0);
    std::vector<float> co32(co64.begin(), co64.end());
    AMG_HIP(hipMemcpy(c64, co64.data(), nco * 8, hipMemcpyHostToDevice));
    AMG_HIP(hipMemcpy(c32, co32.data(), nco * 4, hipMemcpyHostToDevice));
    AMG_HIP(hipMemcpy(vol, H.vol.data(), H.vol.size() * 8, hipMemcpyHostToDevice));
    AMG_HIP(hipMemcpy(rho, H.rho.data() + 1, H.nz * 8, hipMemcpyHostToDevice));
    d.co64 = c64; d.co32 = c32; d.vol = vol; d.rho = rho;
    if (l == 0) {
      double* bf = (double*)I->mem.alloc(H.bcfold.size() * 8);
      if (!bf) return fail();
      AMG_HIP(hipMemcpy(bf, H.bcfold.data(), H.bcfold.size() * 8, hipMemcpyHostToDevice));
      d.bcfold = bf;
    }
    if (l == L - 1) {
      double* mv = (double*)I->mem.alloc(I->H.minvT.size() * 8);
      if (!mv) return fail();
      AMG_HIP(hipMemcpy(mv, I->H.minvT.data(), I->H.minvT.size() * 8, hipMemcpyHostToDevice));
      d.minvT = mv;
      if (d.ncell >= kSplitCells) {
        const size_t n = (size_t)d.ncell;
        std::vector<double> m(n * n);
        for (size_t r = 0; r < n; ++r)
          for (size_t c = 0; c < n; ++c) m[r * n + c] = I->H.minvT[c * n + r];
        double* mr = (double*)I->mem.alloc(n * n * 8);
        if (!mr) return fail();
        AMG_HIP(hipMemcpy(mr, m.data(), n * n * 8, hipMemcpyHostToDevice));
        d.minv = mr;
      }
    }
    // fields of the V-cycle precision
    d.u = I->mem.alloc((size_t)d.ncell * w);
    d.f = I->mem.alloc((size_t)d.ncell * w);
    d.scr = nullptr;
    if (!d.u || !d.f) return fail();
  }
  // FP64 finest entries
  DevLevel fA = ents[0], fB = ents[0];
  if (I->fp64) {
    fA.u2 = I->mem.alloc((size_t)fA.ncell * 8);   // previous iterate (relmax)
    if (!fA.u2) return fail();
    ents[0] = fA;
  } else {
    void* pA = I->mem.alloc((size_t)fA.ncell * 8);
    void* pB = I->mem.alloc((size_t)fA.ncell * 8);
    void* b = I->mem.alloc((size_t)fA.ncell * 8);
    if (!pA || !pB || !b) return fail();
    fA.u = pA; fA.u2 = pB; fA.f = b; fA.scr = nullptr;
    fB.u = pB; fB.u2 = pA; fB.f = b; fB.scr = nullptr;
  }
  // descriptor indices: this instance's entries follow the context's
  const int base = (int)hL.size(), xbase = (int)hX.size();
  std::vector<DevLevel> newL;
  std::vector<DevXfer> newX;
  for (int l = 0; l < L; ++l) {
    I->lev.push_back(base + l);
    newL.push_back(ents[l]);
  }
  if (I->fp64) {
    I->fine[0] = I->fine[1] = I->lev[0];
  } else {
    fA.link = fB.link = I->lev[0];
    I->fine[0] = base + (int)newL.size();
    newL.push_back(fA);
    I->fine[1] = base + (int)newL.size();
    newL.push_back(fB);
  }
  for (int l = 0; l + 1 < L; ++l) {
    DevXfer x;
    x.fine = I->lev[l];
    x.coarse = I->lev[l + 1];
    x.fx = I->H.lv[l].fx; x.fy = I->H.lv[l].fy; x.fz = I->H.lv[l].fz;
    // the V-cycle's post-smoothing of level l starts with colour 0, its
    // pre-smoothing ends with colour 1
    const bool line = cfg.smoother == AMG_SMOOTHER_LJ;
    x.skip = kXferSkipNone;
    if (I->post[l] >= 1 && x.fx == 2 && x.fy == 2 && x.fz == 2)
      x.skip = !line ? kXferSkipPoint : ents[l].pz ? kXferSkipLineWrap : kXferSkipLine;
    x.rskip = kXferSkipNone;
    if (I->pre[l] >= 1 && restrict222_ok(ents[l], ents[l + 1], x) && !(line && ents[l].pz))
      x.rskip = line ? kXferSkipLine : kXferSkipPoint;
    x.pmx = gm_magic((unsigned)prolong_duals_dir(ents[l + 1].nx, x.fx, ents[l].px), &x.psx);
    x.pmy = gm_magic((unsigned)prolong_duals_dir(ents[l + 1].ny, x.fy, ents[l].py), &x.psy);
    I->xf.push_back(xbase + (int)newX.size());
    newX.push_back(x);
  }
  DevTail tl;
  std::memset(&tl, 0, sizeof tl);
  // coarse tail: every level from the first one with <= kTailCells cells
  {
    // grow the tail upward from the coarsest level while it stays small,
    // short-columned and inside the LDS budget
    const size_t limit = tail_lds_limit();
    auto al = [](size_t b) { return (b + 15) & ~(size_t)15; };
    // LDS of one tail level: u, f, coefficient table, cell widths
    // the coarsest level's dense inverse is staged in LDS when it is small
    const size_t nco = (size_t)I->H.lv[L - 1].ncell();
    const size_t minv_bytes = nco * nco * w <= kTailMinvLds ? al(nco * nco * w) : 0;
    auto level_bytes = [&](const HostLevel& H, bool coarsest) {
      const size_t n1 = (size_t)H.nx + H.ny + H.nz;
      return 2 * al((size_t)H.ncell() * w) + al(3 * n1 * w) +
             (coarsest ? minv_bytes : al(n1 * 8));
    };
    int l0 = L - 1;
    size_t bytes = level_bytes(I->H.lv[L - 1], true);
    for (int l = L - 2; l >= 0; --l) {
      const HostLevel& H = I->H.lv[l];
      const size_t nb2 = bytes + level_bytes(H, false);
      if (H.ncell() > kTailCells || nb2 > limit) break;
      if (cfg.smoother == AMG_SMOOTHER_LJ && H.nz > (I->fp64 ? kTailNz64 : kTailNz)) break;
      l0 = l;
      bytes = nb2;
    }
    if (bytes > limit) {
      set_error("amg_create: the coarsest level does not fit in LDS");
      return AMG_EBADSIZE;
    }
    I->l0 = l0;
    tl.nlev = L - l0;
    const bool zero_first = l0 > 0 || !I->fp64;   // the first tail level starts from u = 0
    tl.copy_u_in = zero_first ? 0 : 1;
    tl.zline = cfg.smoother == AMG_SMOOTHER_LJ ? 1 : 0;
    size_t off = 0;
    auto take = [&](size_t b) {
      const size_t o = off;
      off += al(b);
      return (int)o;
    };
    for (int d = 0; d < tl.nlev; ++d) {
      const HostLevel& H = I->H.lv[l0 + d];
      const size_t n1 = (size_t)H.nx + H.ny + H.nz;
      tl.lev[d] = I->lev[l0 + d];
      if (l0 + d < L - 1) tl.xf[d] = I->xf[l0 + d];
      tl.lds_u[d] = take((size_t)H.ncell() * w);
      tl.lds_f[d] = take((size_t)H.ncell() * w);
      tl.lds_co[d] = take(3 * n1 * w);
      tl.lds_vol[d] = (d + 1 < tl.nlev) ? take(n1 * 8) : 0;
    }
    tl.lds_minv = minv_bytes ? take(minv_bytes) : -1;
    // the sub-V-cycle as a step list, in the order of mgMK/mg_cycle.F90
    // (pre-smoothing and restriction down, the coarsest solve, prolongation
    // and post-smoothing up); from a zero guess the first red pass sees
    // zeros everywhere and the first black pass zero old black values
    {
      std::vector<int> st;
      auto step = [&](int type, int d, int par, int zg) {
        st.push_back(type | (d << 3) | (par << 7) | (zg << 8));
      };
      for (int d = 0; d + 1 < tl.nlev; ++d) {
        const bool zf = d > 0 || zero_first;
        const int pre = I->pre[l0 + d];
        if (zf && pre == 0) step(kStepZero, d, 0, 0);
        for (int q = 0; q < pre; ++q) {
          step(kStepHalf, d, 0, (zf && q == 0) ? 3 : 0);
          step(kStepHalf, d, 1, (zf && q == 0) ? 2 : 0);
        }
        step(kStepRestrict, d, 0, 0);
      }
      step(kStepCoarse, tl.nlev - 1, 0, 0);
      for (int d = tl.nlev - 2; d >= 0; --d) {
        step(kStepProlong, d, 0, 0);
        for (int q = 0; q < I->post[l0 + d]; ++q) {
          step(kStepHalf, d, 0, 0);
          step(kStepHalf, d, 1, 0);
        }
      }
      if ((int)st.size() > kTailMaxSteps) {
        set_error("amg_create: too many smoothing sweeps on the coarse levels");
        return AMG_EINVAL;
      }
      tl.nsteps = (int)st.size();
      std::copy(st.begin(), st.end(), tl.steps);
    }
    int nzmax = 1;
    bool all222 = true;
    for (int d = 0; d < tl.nlev; ++d) {
      const HostLevel& H = I->H.lv[l0 + d];
      nzmax = std::max(nzmax, H.nz);
      if (d + 1 < tl.nlev && (H.fx != 2 || H.fy != 2 || H.fz != 2)) all222 = false;
    }
    I->tail_nzc = tail_nz_class(I->fp64, nzmax, cfg.smoother == AMG_SMOOTHER_LJ);
    I->tail_222 = all222;
    I->tail_split = tl.lds_minv < 0 && (int)nco >= kSplitCells;
    I->tail = (int)hT.size();
    I->tail_lds = off;
  }
  // separable direct solve of the first level below the finest, above the
  // coarse tail (the tail solves small levels in one launch), that the
  // direct solve takes: at most kDirectCells cells, nx and ny multiples of
  // 16, nz <= kDirectMaxNz, the operands within the LDS; the levels below it
  // are not cycled.  A mixed solve computes it in FP32 (its tables are
  // FP32), an FP64 one in FP64.
  if (cfg.coarse_direct != 0) {
    const bool f64 = I->fp64;
    const size_t es = f64 ? sizeof(double) : sizeof(float);
    for (int cand = 1; cand < I->l0; ++cand) {
      const HostLevel& H = I->H.lv[cand];
      if (H.ncell() > kDirectCells) continue;
      const size_t lds = direct_lds(H.nx, H.ny, H.nz, f64);
      DirectTables D;
      std::string derr;
      if (!(H.nx % 16 == 0 && H.ny % 16 == 0 && H.nz <= kDirectMaxNz && lds <= direct_lds_limit() &&
            direct_setup(H, I->H.bc, I->H.singular, &D, &derr)))
        continue;
      void* dm = I->mem.alloc(D.mats.size() * es);
      void* di = I->mem.alloc(D.inv.size() * es);
      void* ds = I->mem.alloc(2 * (size_t)H.ncell() * es);
      if (!dm || !di || !ds) return fail();
      if (f64) {
        AMG_HIP(hipMemcpy(dm, D.mats.data(), D.mats.size() * es, hipMemcpyHostToDevice));
        AMG_HIP(hipMemcpy(di, D.inv.data(), D.inv.size() * es, hipMemcpyHostToDevice));
      } else {
        const std::vector<float> m32(D.mats.begin(), D.mats.end()), i32(D.inv.begin(), D.inv.end());
        AMG_HIP(hipMemcpy(dm, m32.data(), m32.size() * es, hipMemcpyHostToDevice));
        AMG_HIP(hipMemcpy(di, i32.data(), i32.size() * es, hipMemcpyHostToDevice));
      }
      DevLevel& d = newL[cand];
      d.dmat = (double*)dm;
      d.dinv = (double*)di;
      d.dscr = (double*)ds;
      I->ld = cand;
      I->dlds = lds;
      break;
    }
  }
  // projection metrics
  {
    const HostLevel& F = I->H.lv[0];
    const int nx = F.nx, ny = F.ny, nz = F.nz;
    std::vector<double> m;
    m.insert(m.end(), F.rho.begin(), F.rho.end());
    auto inv = [&](const std::vector<double>& a) {
      for (double v : a) m.push_back(1.0 / v);
    };
    inv(F.ax.dxc); inv(F.ay.dxc); inv(F.az.dxc);
    inv(F.ax.dxe); inv(F.ay.dxe); inv(F.az.dxe);
    (void)nx; (void)ny; (void)nz;
    I->d_met = (double*)I->mem.alloc(m.size() * 8);
    if (!I->d_met) return fail();
    AMG_HIP(hipMemcpy(I->d_met, m.data(), m.size() * 8, hipMemcpyHostToDevice));
  }
  I->bytes_per_cycle = cycle_bytes(*I, newX);
  I->bytes_corr = (size_t)((I->fp64 ? 16.0 : 32.0) * (double)I->H.lv[0].ncell());
  I->bytes_vcycle = I->bytes_per_cycle - I->bytes_corr;
  // paired cycles: mixed precision, the residual stop rule, and a finest
  // level cycled by its own launches with a first sweep (the sweep that
  // reads the continued iterate)
  const int pairok = !I->fp64 && cfg.pair_cycles && cfg.stop_rule == AMG_STOP_RESIDUAL &&
                     I->lc() >= 1 && I->pre[0] > 0;
  for (DevLevel& d : newL) d.pairok = pairok;
  I->pairok = pairok != 0;
  hL.insert(hL.end(), newL.begin(), newL.end());
  hX.insert(hX.end(), newX.begin(), newX.end());
  hT.push_back(tl);
  inst.push_back(hold.release());
  const int rc2 = upload_tables();
  if (rc2 != AMG_OK) {
    const std::string e = last_error();
    remove_instance(I);
    set_error(e);
    return rc2;
  }
  *out = I;
  return AMG_OK;
}

int Context::upload_tables() {
  if (hL.size() > capL) {
    if (dL) AMG_HIP(hipFree(dL));
    capL = std::max(hL.size() * 2, (size_t)16);
    AMG_HIP(hipMalloc((void**)&dL, capL * sizeof(DevLevel)));
  }
  if (hX.size() > capX || !dX) {
    if (dX) AMG_HIP(hipFree(dX));
    capX = std::max(hX.size() * 2, (size_t)16);
    AMG_HIP(hipMalloc((void**)&dX, capX * sizeof(DevXfer)));
  }
  if (hT.size() > capT || !dT) {
    if (dT) AMG_HIP(hipFree(dT));
    capT = std::max(hT.size() * 2, (size_t)16);
    AMG_HIP(hipMalloc((void**)&dT, capT * sizeof(DevTail)));
  }
  AMG_HIP(hipMemcpy(dT, hT.data(), hT.size() * sizeof(DevTail), hipMemcpyHostToDevice));
  AMG_HIP(hipMemcpy(dL, hL.data(), hL.size() * sizeof(DevLevel), hipMemcpyHostToDevice));
  if (!hX.empty())
    AMG_HIP(hipMemcpy(dX, hX.data(), hX.size() * sizeof(DevXfer), hipMemcpyHostToDevice));
  // partial buffer: 4 values per block of every instance's largest
  // finest-level launch (the z-marching kernels, the change and the
  // weighted-sum folds)
  size_t need = 0;
  for (Instance* I : inst) need += partial_blocks(*I);
  need *= 4;
  if (need > cap_partial) {
    if (d_partial) AMG_HIP(hipFree(d_partial));
    cap_partial = need * 2;
    AMG_HIP(hipMalloc((void**)&d_partial, cap_partial * sizeof(double)));
  }
  drop_graphs();
  return AMG_OK;
}

// z-line launches: items grouped by their segment count (blockDim.y).
void Context::zline_launches(const std::vector<Item>& items, bool fp64, int parity, int zg,
                             int form) {
  std::vector<Item> bucket[kZMaxSegs + 1];
  for (const Item& it : items) {
    const int nz = hL[it.idx].nz;
    const int segs = std::min(kZMaxSegs, (nz + kZSeg - 1) / kZSeg);
    bucket[segs].push_back(it);
  }
  for (int sg = 1; sg <= kZMaxSegs; ++sg)
    if (!bucket[sg].empty())
      for_lists(bucket[sg], [&](const LaunchList& ll, int, int) {
        launch_zline(fp64, dL, ll, parity, zg, sg, stream, form);
      });
}

// Point half-sweeps of the items' levels: the levels large enough march in z
// (rb_march_planes), bucketed by slab length; the others take one thread
// per cell.  The items' block counts are recomputed here.
void Context::rb_launches(const std::vector<Item>& items, bool fp64, int parity, int zg) {
  std::vector<Item> bucket[9];
  for (const Item& it : items) {
    const DevLevel& d = hL[it.idx];
    const int K = rb_march_planes(d.hx, d.ny, d.nz);
    const int blocks = K ? cdiv((long long)d.hx * d.ny, kBlock) * ((d.nz + K - 1) / K)
                         : cdiv((long long)d.hx * d.ny * d.nz, kBlock);
    bucket[K].push_back({it.idx, blocks});
  }
  for (int K : {0, 2, 4, 8})
    if (!bucket[K].empty())
      for_lists(bucket[K], [&](const LaunchList& ll, int, int) {
        if (K) launch_rb_march(fp64, dL, ll, parity, zg, K, stream);
        else launch_rb_point(fp64, dL, ll, parity, zg, stream);
      });
}

// Direct solves: the plane and mode lists hold the same levels in the same
// order, launched in chunks of at most kMaxItems levels.
void Context::direct_launches(const std::vector<Item>& planes, const std::vector<Item>& modes,
                              size_t lds, bool fp64) {
  for (size_t q0 = 0; q0 < planes.size(); q0 += kMaxItems) {
    LaunchList lp, lm;
    lp.nitems = lm.nitems = 0;
    lp.nblocks = lm.nblocks = 0;
    lp.prefix[0] = lm.prefix[0] = 0;
    for (size_t q = q0; q < std::min(planes.size(), q0 + kMaxItems); ++q) {
      lp.items[lp.nitems] = planes[q].idx;
      lp.nblocks += planes[q].blocks;
      lp.prefix[++lp.nitems] = lp.nblocks;
      lm.items[lm.nitems] = modes[q].idx;
      lm.nblocks += modes[q].blocks;
      lm.prefix[++lm.nitems] = lm.nblocks;
    }
    launch_direct(fp64, dL, lp, lm, lds, stream);
  }
}

// ---------------------------------------------------------------------------
void Context::vcycle(const std::vector<Instance*>& S, bool fp64) {
  if (S.empty()) return;
  const bool zline = cfg.smoother == AMG_SMOOTHER_LJ;
  int maxL0 = 0;   // levels above the direct level / coarse tail, one launch per step
  for (Instance* I : S) maxL0 = std::max(maxL0, I->lc());
  const DevLevel* L = dL;
  auto half = [&](const std::vector<Item>& items, int parity, int zg) {
    if (zline) zline_launches(items, fp64, parity, zg);
    else rb_launches(items, fp64, parity, zg);
  };
  auto blocks_smooth = [&](const DevLevel& d) {
    return zline ? cdiv((long long)d.hx * d.ny, 64) : cdiv((long long)d.hx * d.ny * d.nz, kBlock);
  };
  bool any_pair = false;
  for (Instance* I : S) any_pair = any_pair || I->pairok;
  auto smooth = [&](int dpt, bool pre, bool zero_first) {
    int maxs = 0;
    for (Instance* I : S)
      if (dpt < I->lc()) maxs = std::max(maxs, pre ? I->pre[dpt] : I->post[dpt]);
    for (int s = 0; s < maxs; ++s) {
      std::vector<Item> items;
      for (Instance* I : S) {
        if (dpt >= I->lc()) continue;
        if (s >= (pre ? I->pre[dpt] : I->post[dpt])) continue;
        const int li = I->lev[dpt];
        items.push_back({li, blocks_smooth(hL[li])});
      }
      // from a zero initial guess: the red pass sees zero everywhere, the
      // black pass sees zero old black values (lagged periodic-z couplings);
      // the finest level of a mixed cycle that may continue the previous
      // cycle's iterate leaves that to the instance's state word
      const int dyn = (pre && dpt == 0 && !fp64 && any_pair) ? kZgDyn : 0;
      half(items, 0, (zero_first && s == 0) ? 3 | dyn : 0);
      half(items, 1, (zero_first && s == 0) ? 2 | dyn : 0);
    }
  };
  for (int dpt = 0; dpt < maxL0; ++dpt) {
    const bool zero_first = (dpt > 0) || !fp64;
    std::vector<Item> zero;
    for (Instance* I : S)
      if (dpt < I->lc() && zero_first && I->pre[dpt] == 0)
        zero.push_back({I->lev[dpt], cdiv(hL[I->lev[dpt]].ncell, kBlock)});
    for_lists(zero, [&](const LaunchList& ll, int, int) { launch_fill_zero(fp64, L, ll, stream); });
    smooth(dpt, true, zero_first);
    std::vector<Item> items, items222;
    for (Instance* I : S) {
      if (dpt >= I->lc()) continue;
      const int xi = I->xf[dpt];
      const DevLevel& C = hL[I->lev[dpt + 1]];
      if (restrict222_ok(hL[I->lev[dpt]], C, hX[xi])) items222.push_back({xi, restrict222_blocks(C)});
      else items.push_back({xi, cdiv(C.ncell, kBlock)});
    }
    // the residuals the pre-smoothing's last half-sweep solved are taken as
    // zero (DevXfer::rskip)
    for_lists(items222, [&](const LaunchList& ll, int, int) {
      launch_restrict222(fp64, L, dX, ll, stream, true);
    });
    for_lists(items, [&](const LaunchList& ll, int, int) {
      launch_restrict(fp64, L, dX, ll, stream);
    });
  }
  // the direct levels: transform, z solves, transform back
  {
    std::vector<Item> planes, modes;
    size_t lds = 0;
    for (Instance* I : S) {
      if (I->ld < 0) continue;
      const int li = I->lev[I->ld];
      planes.push_back({li, hL[li].nz * (hL[li].nx / kDirectCols)});
      modes.push_back({li, hL[li].ny * (hL[li].nx / kDirectCols)});
      lds = std::max(lds, I->dlds);
    }
    direct_launches(planes, modes, lds, fp64);
  }
  // the coarse tails, one launch per kernel variant; a split tail runs down,
  // then the coarse solve across the device, then up
  for (int nzc : {0, 8, 16, 32}) {
    for (int a222 = 0; a222 < 2; ++a222) {
      for (int split = 0; split < 2; ++split) {
        std::vector<Item> items, coarse;
        size_t lds = 0;
        for (Instance* I : S) {
          if (I->ld >= 0) continue;
          if (I->tail_nzc != nzc || (int)I->tail_222 != a222 || (int)I->tail_split != split) continue;
          items.push_back({I->tail, 1});
          lds = std::max(lds, I->tail_lds);
          const int lc = I->lev.back();
          coarse.push_back({lc, cdiv(hL[lc].ncell, kBlock / 64)});
        }
        if (items.empty()) continue;
        if (!split) {
          for_lists(items, [&](const LaunchList& ll, int, int) {
            launch_tail(fp64, nzc, a222 != 0, L, dX, dT, ll, lds, stream);
          });
          continue;
        }
        for_lists(items, [&](const LaunchList& ll, int, int) {
          launch_tail(fp64, nzc, a222 != 0, L, dX, dT, ll, lds, stream, 1);
        });
        for_lists(coarse, [&](const LaunchList& ll, int, int) {
          launch_coarse_wide(fp64, L, ll, stream);
        });
        for_lists(items, [&](const LaunchList& ll, int, int) {
          launch_tail(fp64, nzc, a222 != 0, L, dX, dT, ll, lds, stream, 2);
        });
      }
    }
  }
  for (int dpt = maxL0 - 1; dpt >= 0; --dpt) {
    std::vector<Item> items;
    for (Instance* I : S)
      if (dpt < I->lc())
        items.push_back({I->xf[dpt], cdiv(prolong_duals(hL[I->lev[dpt]], hL[I->lev[dpt + 1]],
                                                        hX[I->xf[dpt]]),
                                          kBlock)});
    // the cells the post-smoothing's first half-sweep rewrites are left
    // (DevXfer::skip)
    for_lists(items, [&](const LaunchList& ll, int, int) {
      launch_prolong(fp64, L, dX, ll, stream, true);
    });
    smooth(dpt, false, false);
  }
}

namespace {
std::vector<Item> fine_items(const Context& C, const std::vector<Instance*>& S, bool current,
                             long long per_block = kBlock) {
  std::vector<Item> items;
  for (Instance* I : S) {
    const int li = current ? I->fine[I->cur] : I->fine[0];
    items.push_back({li, cdiv(C.hL[li].ncell, per_block)});
  }
  return items;
}

// Items of the z-marching FP64 stencil kernels (residual, defect, correct+defect).
std::vector<Item> march_items(const Context& C, const std::vector<Instance*>& S, bool current) {
  std::vector<Item> items;
  for (Instance* I : S) {
    const int li = current ? I->fine[I->cur] : I->fine[0];
    const DevLevel& d = C.hL[li];
    items.push_back({li, (int)march_blocks(d.hx, d.ny, d.nz)});
  }
  return items;
}
}  // namespace

// Launches one cycle on every instance of S plus the norm kernels; the
// instances' gates (ItemGate) end the cycle on the device, recording the
// norms and deciding whether the next cycle's FP64 correction runs (paired
// cycles); the host reads them later (sync_gates).
int Context::cycle_and_norms(const std::vector<Instance*>& S) {
  if (S.empty()) return AMG_OK;
  const bool fp64 = S[0]->fp64;
  const bool relmax = fp64 && cfg.report_relmax;
  if (relmax) {
    auto items = fine_items(*this, S, true);
    for_lists(items, [&](const LaunchList& ll, int, int) { launch_copy_u(dL, ll, nullptr, stream); });
  }
  vcycle(S, fp64);
  // the FP64 finest entry 0: the kernels take the current pressure buffer
  // from the instance's state word pcur
  auto items = march_items(*this, S, false);
  const int nv = 4;
  // per-instance gates: item t of a list at offset ioff
  auto each_args = [&](int ioff) {
    return GateArgs{d_out + (size_t)ioff * nv, d_out + kOutRegion * 2 + (size_t)ioff * 2, dL,
                    (int)hL.size(), d_igate};
  };
  for_lists(items, [&](const LaunchList& ll, int ioff, int boff) {
    double* part = d_partial + (size_t)boff * nv;
    const GateArgs ge = each_args(ioff);
    const GateArgs* gr = relmax ? nullptr : &ge;
    if (fp64) {
      launch_residual_max(dL, ll, part, stream);
      launch_finalize(ll, part, 2, 0, d_out + (size_t)ioff * nv, nv, stream, gr);
    } else {
      launch_correct_defect32(dL, ll, part, stream);
      launch_finalize(ll, part, 3, 0, d_out + (size_t)ioff * nv, nv, stream, gr);
    }
  });
  if (relmax) {
    auto citems = fine_items(*this, S, true, (long long)kBlock * kFinePer);
    for_lists(citems, [&](const LaunchList& ll, int ioff, int boff) {
      double* part = d_partial + (size_t)boff * nv;
      const GateArgs ge = each_args(ioff);
      launch_change_max(dL, ll, part, stream);
      launch_finalize(ll, part, 2, 0, d_out + kOutRegion * 2 + (size_t)ioff * 2, 2, stream, &ge);
    });
  }
  return AMG_OK;
}

// The host's norm step for one correction closing m V-cycles, from h_out
// (and its relmax region): the stop rule of mgMK/mg_solver.F90:114-131.
void Context::post_norms(const std::vector<Instance*>& S, int m) {
  const int nv = 4;
  for (size_t q = 0; q < S.size(); ++q) {
    Instance* I = S[q];
    I->iters += m;
    const double* o = h_out + q * nv;
    const double rmax = o[0];
    double relmax = 0.0;
    if (I->fp64) {
      if (cfg.report_relmax) {
        const double* c = h_out + kOutRegion * 2 + q * 2;
        relmax = c[1] > 0.0 ? c[0] / c[1] : 0.0;
      } else {
        relmax = std::nan("");
      }
    } else {
      relmax = o[2] > 0.0 ? o[1] / o[2] : 0.0;
      I->cur ^= 1;
    }
    I->relres = I->bnorm > 0.0 ? rmax / I->bnorm : rmax;
    I->relmax = relmax;
    if (cfg.verbose) {
      if (std::isnan(relmax))
        std::fprintf(stderr, "       Iteration: %5d  relres= %12.5e\n", I->iters, I->relres);
      else
        std::fprintf(stderr, "       Iteration: %5d  RESIDUAL= %22.15e  relres= %12.5e\n", I->iters,
                     relmax, I->relres);
    }
    if (!std::isfinite(rmax)) {
      I->status = AMG_ENAN;
      I->done = true;
      continue;
    }
    const double crit = (cfg.stop_rule == AMG_STOP_RELMAX) ? relmax : I->relres;
    if (crit <= cfg.tol) {
      I->converged = true;
      I->done = true;
    } else if (I->iters >= cfg.maxcy) {
      I->done = true;
    }
  }
}

// Singular instances: rhs == true removes the incompatible part of b
// (b -= sum V b / sum V); otherwise removes the null-space part of the
// current pressure (p -= (sum V p / sum V rho) rho).
// sum V p of the current iterate of every instance of S (item order of S), or
// nullptr when no instance is singular
const double* Context::null_sums(const std::vector<Instance*>& S) {
  bool any = false;
  std::vector<Item> sums;
  for (Instance* I : S) {
    any = any || I->H.singular;
    const int li = I->fine[0];   // the device's pcur picks the current buffer
    sums.push_back({li, (int)march_blocks(hL[li].hx, hL[li].ny, hL[li].nz)});
  }
  if (!any) return nullptr;
  double* c = d_out + kOutRegion * 3;
  for_lists(sums, [&](const LaunchList& ll, int ioff, int boff) {
    launch_wsum(dL, ll, 0, d_partial + boff, stream);
    launch_finalize(ll, d_partial + boff, 1, 1, c + ioff, 1, stream);
  });
  return c;
}

// The state words of the instances of S back to zero (only_stopped: of those
// whose solve has stopped).
void Context::reset_gates(const std::vector<Instance*>& S, bool only_stopped) {
  std::vector<Item> mine;
  for (Instance* I : S) mine.push_back({I->fine[0], 1});
  for_lists(mine, [&](const LaunchList& ll, int, int) {
    launch_gate_reset(dL, (int)hL.size(), ll, stream, only_stopped);
  });
}

// The start of a solve for every instance of S, enqueued only: the state
// words cleared, the Fortran arrays in the chequer layout (b with the
// boundary fold) with sum V b, the first defect (mixed: also the FP32 defect
// of the first cycle; for a singular operator the incompatible part of b is
// removed first, in place, b -= sum V b / sum V), and the gates initialised
// from it on the device.
int Context::prepare_async(const std::vector<Instance*>& S) {
  const int nv = 4;
  PtrList rhs{}, phi{};
  std::vector<Item> items;
  for (Instance* I : S) {
    I->cur = 0;
    I->iters = 0;
    I->dev_cycles = 0;
    I->converged = false;
    I->done = false;
    I->status = AMG_OK;
    I->gate_seen = false;
    items.push_back({I->fine[0], cdiv(hL[I->fine[0]].ncell, (long long)kBlock * kWsumPer)});
  }
  reset_gates(S, false);
  double* c = d_out + kOutRegion * 3;
  for_lists(items, [&](const LaunchList& ll, int ioff, int boff) {
    for (int t = 0; t < ll.nitems; ++t) {
      rhs.p[t] = const_cast<double*>(S[ioff + t]->d_rhs);
      phi.p[t] = S[ioff + t]->d_phi;
    }
    launch_ingest(dL, ll, rhs, phi, d_partial + boff, stream);
    launch_finalize(ll, d_partial + boff, 1, 1, c + ioff, 1, stream);
  });
  const bool fp64 = S[0]->fp64;
  const auto fitems = march_items(*this, S, false);
  for_lists(fitems, [&](const LaunchList& ll, int ioff, int boff) {
    double* part = d_partial + (size_t)boff * nv;
    launch_first_defect(fp64, dL, ll, c + ioff, part, stream);
    launch_finalize(ll, part, 2, 0, d_out + (size_t)ioff * nv, nv, stream);
    launch_gate_init(ll, d_out + (size_t)ioff * nv, d_igate, dL, (int)hL.size(), cfg.maxcy,
                     cfg.stop_rule, cfg.tol, cfg.report_relmax, stream);
  });
  AMG_HIP(hipGetLastError());
  return AMG_OK;
}

int Context::gate_slots() const {
  int n = 0;
  for (const Instance* I : inst) n = std::max(n, I->gslot + 1);
  return n;
}

// Brings the gates of the instances of S to the host and replays them.
int Context::sync_gates(const std::vector<Instance*>& S, double* bytes) {
  AMG_HIP(hipMemcpyAsync(h_igate, d_igate, sizeof(ItemGate) * gate_slots(), hipMemcpyDeviceToHost,
                         stream));
  AMG_HIP(hipStreamSynchronize(stream));
  for (Instance* I : S) {
    const int rc = absorb_gate(I, h_igate[I->gslot], bytes);
    if (rc != AMG_OK) return rc;
  }
  return AMG_OK;
}


int Context::absorb_gate(Instance* I, const ItemGate& g, double* bytes) {
  if (!I->gate_seen) {
    // the first defect (the decisions of k_gate_init)
    I->gate_seen = true;
    I->bnorm = g.bnorm;
    I->res0 = g.res0;
    I->relres = g.res0;
    I->relmax = 0.0;
    if (!std::isfinite(g.bnorm) || !std::isfinite(g.res0)) {
      I->status = AMG_ENAN;
      I->done = true;
    } else if (cfg.stop_rule == AMG_STOP_RESIDUAL && (g.bnorm == 0.0 || g.res0 <= cfg.tol)) {
      I->converged = true;
      I->done = true;
    }
    I->units = 0;
  }
  if (g.count < I->units || g.count - I->units > kItemRing) {
    set_error("gated cycles: the host lost track of the device's corrections");
    return AMG_EHIP;
  }
  const std::vector<Instance*> A{I};
  for (int c = I->units; c < g.count; ++c) {
    if (I->done) {
      set_error("gated cycles: the device ran past the host's stop");
      return AMG_EHIP;
    }
    const double* h = g.hist[c % kItemRing];
    const int m = g.hcyc[c % kItemRing];
    if (m < 1 || m > 2) {
      set_error("gated cycles: a correction closing " + std::to_string(m) + " cycles");
      return AMG_EHIP;
    }
    for (int q = 0; q < 4; ++q) h_out[q] = h[q];
    h_out[kOutRegion * 2] = h[4];
    h_out[kOutRegion * 2 + 1] = h[5];
    post_norms(A, m);
    if (bytes) *bytes += (double)(m * I->bytes_vcycle + I->bytes_corr);
  }
  I->units = g.count;
  I->dev_cycles = g.cycles;
  if ((g.stop != 0) != I->done) {
    set_error("gated cycles: the device and the host stop at different cycles");
    return AMG_EHIP;
  }
  return AMG_OK;
}

int Context::finish(const std::vector<Instance*>& S) {
  // the first FP64 entry: the device's pcur picks the buffer of the current
  // pressure, so the copy out needs no host view of the solve
  std::vector<Item> cur;
  for (Instance* I : S) {
    const int li = I->fine[0];
    cur.push_back({li, cdiv(hL[li].ncell, kBlock)});
  }
  // null-space removal, fused into the copy out: p -= (sum V p / sum V rho) rho
  const double* c = null_sums(S);
  PtrList phi{};
  for_lists(cur, [&](const LaunchList& ll, int ioff, int) {
    for (int t = 0; t < ll.nitems; ++t) phi.p[t] = S[ioff + t]->d_phi;
    launch_to_fortran(dL, ll, phi, c ? c + ioff : nullptr, stream);
  });
  // ghost layers as mg_bdry(ngrid, phif, .true., .true.) (mgMK/mg_solver.F90:139)
  for (Instance* I : S) {
    const HostLevel& F = I->H.lv[0];
    const double edge[7] = {0.0, F.ay.dxc[F.ny], F.az.dxc[1], F.ay.dxc[1], F.az.dxc[F.nz],
                            F.ax.dxc[1], F.ax.dxc[F.nx]};
    GhostBC bc{};
    for (int face = 1; face <= 6; ++face) {
      const int t = cfg.tbc[face - 1];
      bc.type[face] = t == AMG_BC_NEUMANN ? 1 : t == AMG_BC_DIRICHLET ? 2 : 0;
      bc.v[face] = cfg.vbc[face - 1];
      bc.vd[face] = cfg.vbc[face - 1] * edge[face];
    }
    launch_ghost_fill(I->d_phi, F.nx, F.ny, F.nz, bc, stream);
  }
  return AMG_OK;
}

// The cycles of one instance in gated batches: a batch enqueues as many
// cycles as the instance's previous solve took (one when there is none)
// without a host round trip between them; each cycle ends with its gate
// (ItemGate), and once it stops, the rest of the batch's kernels return at
// once.  After the batch the host replays the recorded corrections.  A
// batch runs as graphs of 4, 2 and 1 consecutive cycles (the fewest graph
// launches: each graph boundary costs the device several microseconds),
// captured once: the launch lists do not depend on the solve's state, which
// the kernels read from the instance's state words.  When capture is
// unavailable the caller's loop takes over.
// A batch that should end the solve (it reaches the previous solve's cycle
// count) is followed, before the host looks, by the solve's tail: the copy
// out (which reads the device's state, not the host's) and the state reset
// of a stopped solve.  If the solve has not stopped after all, the reset
// leaves it alone and the copy out is made again at the end; *tail tells
// the caller whether it stands.
int Context::gated_cycles(Instance* I, double* bytes, bool* tail) {
  const std::vector<Instance*> A{I};
  *tail = false;
  while (!I->done) {
    int batch = I->last_iters > I->dev_cycles ? I->last_iters - I->dev_cycles : 1;
    batch = std::max(1, std::min(batch, std::min(kItemRing, cfg.maxcy - I->dev_cycles)));
    const bool last = I->last_iters > 0 && I->dev_cycles + batch >= I->last_iters;
    std::vector<int> plan;
    for (int left = batch; left > 0;) {
      const int q = left >= 4 ? 2 : left >= 2 ? 1 : 0;
      plan.push_back(q);
      left -= kChainLen[q];
    }
    for (int q : plan) {
      if (gchain[q]) continue;
      hipGraph_t g = nullptr;
      bool ok = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
      if (ok) {
        for (int c = 0; c < kChainLen[q]; ++c) cycle_and_norms(A);
        ok = hipStreamEndCapture(stream, &g) == hipSuccess && g;
      }
      if (ok) ok = hipGraphInstantiate(&gchain[q], g, nullptr, nullptr, 0) == hipSuccess;
      if (g) (void)hipGraphDestroy(g);
      if (!ok) {
        graph_ok = false;
        gchain[q] = nullptr;
        (void)hipGetLastError();
        return AMG_OK;
      }
    }
    for (int q : plan) AMG_HIP(hipGraphLaunch(gchain[q], stream));
    if (last) {
      AMG_HIP(hipEventRecord(ev[1], stream));
      const int rf = finish(A);
      if (rf != AMG_OK) return rf;
      reset_gates(A, true);
      AMG_HIP(hipEventRecord(ev[3], stream));
    }
    const int rc = sync_gates(A, bytes);
    if (rc != AMG_OK) return rc;
    *tail = last && I->done;
  }
  return AMG_OK;
}

int Context::solve(const std::vector<Instance*>& S, amg_stats* st) {
  if (S.empty()) return AMG_OK;
  AMG_HIP(hipEventRecord(ev[2], stream));
  // the first defect's gates are read with the first cycles' (a solve that
  // the first defect ends has its cycles return at once on the device)
  int rc = prepare_async(S);
  if (rc != AMG_OK) return rc;
  AMG_HIP(hipEventRecord(ev[0], stream));
  double bytes = 0.0;
  bool tail = false;
  if (S.size() == 1 && cfg.use_graph && graph_ok) {
    rc = gated_cycles(S[0], &bytes, &tail);
    if (rc != AMG_OK) return rc;
  }
  // one gated cycle of every running instance at a time
  for (int it = 0; it <= cfg.maxcy; ++it) {
    std::vector<Instance*> A;
    for (Instance* I : S)
      if (!I->done) A.push_back(I);
    if (A.empty()) break;
    rc = cycle_and_norms(A);
    if (rc == AMG_OK) rc = sync_gates(A, &bytes);
    if (rc != AMG_OK) return rc;
  }
  for (Instance* I : S)
    if (!I->done) {
      set_error("gated cycles: the device did not stop within maxcy cycles");
      return AMG_EHIP;
    }
  for (Instance* I : S) I->last_iters = I->iters;
  if (!tail) {
    AMG_HIP(hipEventRecord(ev[1], stream));
    rc = finish(S);
    if (rc != AMG_OK) return rc;
    // the state words back to running (kernels launched on the levels later,
    // e.g. run_kernel, must not find a stop)
    reset_gates(S, false);
    AMG_HIP(hipEventRecord(ev[3], stream));
    AMG_HIP(hipStreamSynchronize(stream));
  }
  AMG_HIP(hipGetLastError());
  int status = AMG_OK;
  for (Instance* I : S) {
    if (I->status != AMG_OK) status = I->status;
    else if (!I->converged) {
      I->status = AMG_ENOTCONV;
      if (status == AMG_OK) status = AMG_ENOTCONV;
    }
  }
  if (st) {
    float ms = 0.f;
    std::memset(st, 0, sizeof *st);
    if (hipEventElapsedTime(&ms, ev[2], ev[3]) == hipSuccess) st->ms_total = ms;
    if (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) st->ms_cycles = ms;
    st->bytes_cycles = bytes;
    Instance* I = S[0];
    st->iters = I->converged ? I->iters : cfg.maxcy + 1;
    st->converged = I->converged;
    st->relres = I->relres;
    st->relmax = I->relmax;
    st->res0 = I->res0;
    st->bnorm = I->bnorm;
    st->nlevels = (int)I->H.lv.size();
    for (int l = 0; l < st->nlevels && l < AMG_MAX_LEVELS; ++l) {
      st->level_n[l][0] = I->H.lv[l].nx;
      st->level_n[l][1] = I->H.lv[l].ny;
      st->level_n[l][2] = I->H.lv[l].nz;
    }
  }
  return status;
}

int Context::run_kernel(Instance* I, int which, int level, int reps, double* bytes) {
  const int L = (int)I->lev.size();
  if (level < 0 || level >= L || reps < 0) {
    set_error("amg_run_kernel: bad level or reps");
    return AMG_EINVAL;
  }
  const bool fp64 = I->fp64;
  const double w = fp64 ? 8.0 : 4.0;
  const DevLevel& d = hL[I->lev[level]];
  const bool zline = cfg.smoother == AMG_SMOOTHER_LJ;
  std::vector<Item> items;
  double b = 0.0;
  int form = 0;
  if (which >= 10 && which <= 13) {
    // the z-line half-sweep in a given form: 10/11 short, 12/13 lean (red/black)
    if (!zline || (fp64 && which < 12)) {
      set_error("amg_run_kernel: forms 10-13 are z-line sweeps; the short form is FP32 only");
      return AMG_EINVAL;
    }
    form = which < 12 ? 1 : 2;
    which = (which & 1) ? 4 : 0;
  }
  if (which == 0 || which == 4 || which == 6) {
    items.push_back({I->lev[level], zline ? cdiv((long long)d.hx * d.ny, 64)
                                          : cdiv((long long)d.hx * d.ny * d.nz, kBlock)});
    b = 1.5 * w * d.ncell;
  } else if (which == 1 || which == 2 || which == 15 || which == 16) {
    if (level >= L - 1) {
      set_error("amg_run_kernel: no coarser level");
      return AMG_EINVAL;
    }
    const DevLevel& c = hL[I->lev[level + 1]];
    const DevXfer& xf = hX[I->xf[level]];
    const bool restr = which == 1 || which == 16;
    const bool r222 = restr && restrict222_ok(d, c, xf);
    const int nc = restr ? c.ncell : prolong_duals(d, c, xf);
    items.push_back({I->xf[level], r222 ? restrict222_blocks(c) : cdiv(nc, kBlock)});
    double fine = 2.0;
    if (which == 15 && xf.skip != kXferSkipNone) fine = 1.0;
    if (which == 16 && r222 && xf.rskip != kXferSkipNone) fine = 1.5;
    b = (fine + 0.125) * w * d.ncell;
    if (r222) which = which == 16 ? 17 : 9;   // the z-march restriction (17: with rskip)
  } else if (which == 3) {
    const DevLevel& fd = hL[I->fine[I->cur]];
    items.push_back({I->fine[I->cur], (int)march_blocks(fd.hx, fd.ny, fd.nz)});
    b = fp64 ? 16.0 * d.ncell : 32.0 * d.ncell;
  } else if (which == 7 || which == 8) {
    if (!zline || d.nz > (fp64 ? kTailNz64 : kTailNz)) {
      set_error("amg_run_kernel: the short-column z-line sweep needs the LJ smoother and a short nz");
      return AMG_EINVAL;
    }
    items.push_back({I->lev[level], cdiv((long long)d.hx * d.ny, kBlock)});
    b = 1.5 * w * d.ncell;
  } else if (which == 14) {
    if (level != I->ld) {
      set_error("amg_run_kernel: the direct solve runs on the direct level");
      return AMG_EINVAL;
    }
    b = 2.0 * w * d.ncell;
    for (int r = 0; r < reps; ++r)
      direct_launches({{I->lev[level], d.nz * (d.nx / kDirectCols)}},
                      {{I->lev[level], d.ny * (d.nx / kDirectCols)}}, I->dlds, fp64);
    AMG_HIP(hipGetLastError());
    if (bytes) *bytes = b;
    return AMG_OK;
  } else if (which == 5) {
    if (level != L - 1) {
      set_error("amg_run_kernel: the coarse solve runs on the coarsest level");
      return AMG_EINVAL;
    }
    items.push_back({I->lev[level], cdiv(d.ncell, kBlock)});
    b = 2.0 * w * d.ncell;
  } else {
    set_error("amg_run_kernel: unknown kernel");
    return AMG_EINVAL;
  }
  for (int r = 0; r < reps; ++r) {
    for_lists(items, [&](const LaunchList& ll, int, int) {
      if (which == 0 || which == 4 || which == 6) {
        const int par = which == 4 ? 1 : 0, zg = which == 6 ? 3 : 0;
        if (zline) zline_launches(items, fp64, par, zg, form);
        else rb_launches(items, fp64, par, zg);
      } else if (which == 7 || which == 8) {
        launch_zline_col(fp64, dL, ll, which == 8 ? 1 : 0, 0, stream);
      } else if (which == 1 || which == 16) {
        launch_restrict(fp64, dL, dX, ll, stream);
      } else if (which == 9 || which == 17) {
        launch_restrict222(fp64, dL, dX, ll, stream, which == 17);
      } else if (which == 2 || which == 15) {
        launch_prolong(fp64, dL, dX, ll, stream, which == 15);
      } else if (which == 5) {
        launch_coarse(fp64, dL, ll, stream);
      } else if (fp64) {
        launch_residual_max(dL, ll, d_partial, stream);
      } else {
        launch_correct_defect32(dL, ll, d_partial, stream);
      }
    });
  }
  AMG_HIP(hipGetLastError());
  if (bytes) *bytes = b;
  return AMG_OK;
}

int Context::level_field(Instance* I, int level, int field, int dir, double* buf) {
  const int L = (int)I->lev.size();
  if (level < -1 || level >= L || (field != 0 && field != 1) || !buf) {
    set_error("amg_level_field: bad argument");
    return AMG_EINVAL;
  }
  const bool fine64 = level == -1;
  const DevLevel& d = hL[fine64 ? I->fine[I->cur] : I->lev[level]];
  const bool dbl = fine64 || I->fp64;
  const size_t n = (size_t)d.ncell;
  void* dev = field == 0 ? d.u : d.f;
  std::vector<double> h64;
  std::vector<float> h32;
  if (dbl) h64.resize(n); else h32.resize(n);
  void* hp = dbl ? (void*)h64.data() : (void*)h32.data();
  const size_t nb = n * (dbl ? 8 : 4);
  AMG_HIP(hipStreamSynchronize(stream));
  if (dir == 0) AMG_HIP(hipMemcpy(hp, dev, nb, hipMemcpyDeviceToHost));
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        const size_t nat = ((size_t)k * d.ny + j) * d.nx + i;
        const size_t ch = (size_t)chequer_index(i, j, k, d.nx, d.ny);
        if (dir == 0) buf[nat] = dbl ? h64[ch] : (double)h32[ch];
        else if (dbl) h64[ch] = buf[nat];
        else h32[ch] = (float)buf[nat];
      }
  if (dir == 1) AMG_HIP(hipMemcpy(dev, hp, nb, hipMemcpyHostToDevice));
  return AMG_OK;
}

}  // namespace amg
