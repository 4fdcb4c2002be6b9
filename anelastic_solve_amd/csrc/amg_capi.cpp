// amg_capi.cpp -- extern "C" entry points declared in include/amg.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

#include "amg.h"
#include "amg_flow.h"
#include "amg_handle.h"
#include "amg_solver.h"

using namespace amg;

struct amg_batch_s {
  Context ctx;
};

namespace {

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

// the finest grid's ghost-inclusive metrics and rho(z), for the flow kernels
FlowGrid flow_grid(const Instance* I) {
  const HostLevel& F = I->H.lv[0];
  FlowGrid G;
  G.nx = F.nx;
  G.ny = F.ny;
  G.nz = F.nz;
  const Axis* ax[3] = {&F.ax, &F.ay, &F.az};
  for (int d = 0; d < 3; ++d) {
    G.dxc[d] = ax[d]->dxc;
    G.dxe[d] = ax[d]->dxe;
  }
  G.rho = F.rho;
  return G;
}

// ---- INI reader with the reference's syntax (utils/readINI.F90) ------------
std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

using Ini = std::map<std::string, std::map<std::string, std::string>>;

bool read_ini(const char* path, Ini& ini) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line, block;
  while (std::getline(in, line)) {
    const size_t c = line.find_first_of("#;!");
    if (c != std::string::npos) line = line.substr(0, c);
    line = trim(line);
    if (line.empty()) continue;
    if (line.front() == '[') {
      const size_t e = line.find(']');
      block = lower(trim(line.substr(1, e == std::string::npos ? std::string::npos : e - 1)));
      continue;
    }
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    ini[block][lower(trim(line.substr(0, eq)))] = trim(line.substr(eq + 1));
  }
  return true;
}

// Fortran reals may use a D exponent (1.0d-4).
bool parse_real(std::string s, double& v) {
  for (char& c : s)
    if (c == 'd' || c == 'D') c = 'e';
  char* end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end && end != s.c_str() && trim(end).empty();
}

bool parse_int(const std::string& s, int& v) {
  char* end = nullptr;
  const long x = std::strtol(s.c_str(), &end, 10);
  if (!end || end == s.c_str() || !trim(end).empty()) return false;
  v = (int)x;
  return true;
}

const std::string* get(const Ini& ini, const char* block, const char* key) {
  auto b = ini.find(lower(block));
  if (b == ini.end()) return nullptr;
  auto k = b->second.find(lower(key));
  if (k == b->second.end()) return nullptr;
  return &k->second;
}

}  // namespace

void amg::metric_ptrs(const Instance* I, const double** rho, const double** rdxc,
                        const double** rdyc, const double** rdzc, const double** rdxe,
                        const double** rdye, const double** rdze) {
  const int nx = I->nx, ny = I->ny, nz = I->nz;
  const double* m = I->d_met;
  *rho = m; m += nz + 2;
  *rdxc = m; m += nx + 2;
  *rdyc = m; m += ny + 2;
  *rdzc = m; m += nz + 2;
  *rdxe = m; m += nx + 2;
  *rdye = m; m += ny + 2;
  *rdze = m;
}

// The divergence report into `partial` (divergence_blocks * 4 values, then
// the 4 results); copies the results to the host and fills the outputs.
int amg::divergence_report(amg_handle h, const double* U, const double* V, const double* W,
                             double* div, double* partial, double* avg, double* maxabs,
                             int loc[3]) {
  const double *rho, *rdxc, *rdyc, *rdzc, *rdxe, *rdye, *rdze;
  metric_ptrs(h->I, &rho, &rdxc, &rdyc, &rdzc, &rdxe, &rdye, &rdze);
  const int nx = h->I->nx, ny = h->I->ny, nz = h->I->nz;
  const int nb = divergence_blocks(nx, ny, nz);
  hipStream_t s = h->ctx.stream;
  launch_divergence(U, V, W, div, rdxc, rdyc, rdzc, nx, ny, nz, partial, s);
  launch_divergence_finalize(partial, nb, partial + 4 * (size_t)nb, s);
  AMG_HIP(hipGetLastError());
  double o[4];
  AMG_HIP(hipMemcpyAsync(o, partial + 4 * (size_t)nb, sizeof o, hipMemcpyDeviceToHost, s));
  AMG_HIP(hipStreamSynchronize(s));
  if (avg) *avg = o[0] / o[1];
  if (maxabs) *maxabs = o[2];
  if (loc) {
    const long long e = (long long)o[3];
    loc[0] = (int)(e % nx) + 2;
    loc[1] = (int)((e / nx) % ny) + 2;
    loc[2] = (int)(e / ((long long)nx * ny)) + 2;
  }
  return AMG_OK;
}

extern "C" {

void amg_config_default(amg_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof *c);
  c->smoother = AMG_SMOOTHER_LJ;
  c->maxcy = 100;
  c->tol = 1e-10;
  c->stop_rule = AMG_STOP_RESIDUAL;
  for (int i = 0; i < AMG_MAX_LEVELS; ++i) c->pre_relax[i] = c->post_relax[i] = 2;
  // periodic x/y, Neumann z (dp/dn = 0)
  c->tbc[0] = AMG_BC_PERIODIC; c->tbc[2] = AMG_BC_PERIODIC;
  c->tbc[4] = AMG_BC_PERIODIC; c->tbc[5] = AMG_BC_PERIODIC;
  c->tbc[1] = AMG_BC_NEUMANN;  c->tbc[3] = AMG_BC_NEUMANN;
  for (int i = 0; i < 6; ++i) c->vbc[i] = c->tbc[i] == AMG_BC_PERIODIC ? -1.0 : 0.0;
  c->precision = AMG_MIXED;
  c->report_relmax = 0;
  c->use_graph = 1;
  c->coarse_direct = 1;
  c->pair_cycles = 1;
}

// mgMK/mg_read_params.F90:44-214
int amg_read_params(const char* infile, amg_config* cfg) {
  if (!infile || !cfg) return fail(AMG_EINVAL, "amg_read_params: NULL argument");
  Ini ini;
  if (!read_ini(infile, ini)) return fail(AMG_EPARSE, std::string("cannot open ") + infile);
  amg_config c = *cfg;
  auto need = [&](const char* block, const char* key, const std::string*& out) {
    out = get(ini, block, key);
    if (!out) {
      set_error(std::string("missing key [") + block + "] " + key);
      return false;
    }
    return true;
  };
  const std::string* v = nullptr;
  if (!need("MGsolver", "Smoother", v)) return AMG_EPARSE;
  {
    const std::string s = trim(*v);
    if (s == "RBPGS") c.smoother = AMG_SMOOTHER_RBPGS;
    else if (s == "LJ") c.smoother = AMG_SMOOTHER_LJ;
    else return fail(AMG_EPARSE, "SMOOTHER: " + s + " UNKNOWN (this build has RBPGS and LJ)");
  }
  auto logical = [&](const char* key, int& dst) {
    const std::string* p = get(ini, "MGsolver", key);
    if (!p) return true;
    const std::string s = lower(trim(*p));
    if (s == "yes") dst = 1;
    else if (s == "no") dst = 0;
    else { set_error(std::string("ERROR: ") + key + " MUST BE yes/no READ: " + s); return false; }
    return true;
  };
  if (!logical("Strict", c.strict) || !logical("Verbose", c.verbose) ||
      !logical("CoarseDirect", c.coarse_direct) || !logical("PairCycles", c.pair_cycles))
    return AMG_EPARSE;
  const char* lk[3] = {"nxlevels", "nylevels", "nzlevels"};
  for (int d = 0; d < 3; ++d) {
    if (!need("MGsolver", lk[d], v)) return AMG_EPARSE;
    if (!parse_int(*v, c.nlevels[d]) || c.nlevels[d] < 1)
      return fail(AMG_EPARSE, std::string("bad ") + lk[d]);
  }
  if (!need("MGsolver", "maxcy", v) || !parse_int(*v, c.maxcy) || c.maxcy < 1)
    return fail(AMG_EPARSE, "bad or missing maxcy");
  if (!need("MGsolver", "tolmax", v) || !parse_real(*v, c.tol))
    return fail(AMG_EPARSE, "bad or missing tolmax");
  c.stop_rule = AMG_STOP_RELMAX;
  if (const std::string* t = get(ini, "MGsolver", "tolres")) {
    if (!parse_real(*t, c.tol)) return fail(AMG_EPARSE, "bad tolres");
    c.stop_rule = AMG_STOP_RESIDUAL;
  }
  if (const std::string* t = get(ini, "MGsolver", "Precision")) {
    const std::string s = lower(trim(*t));
    if (s == "fp64") c.precision = AMG_FP64;
    else if (s == "mixed") c.precision = AMG_MIXED;
    else return fail(AMG_EPARSE, "Precision must be fp64 or mixed");
  }
  const int ngrid = std::max(c.nlevels[0], std::max(c.nlevels[1], c.nlevels[2]));
  if (ngrid > AMG_MAX_LEVELS) return fail(AMG_EPARSE, "more than 16 levels");
  const char* rk[2] = {"PreRelax", "PstRelax"};
  for (int q = 0; q < 2; ++q) {
    if (!need("MGsolver", rk[q], v)) return AMG_EPARSE;
    std::stringstream ss(*v);
    std::string tok;
    int n = 0, vals[AMG_MAX_LEVELS + 1];
    while (std::getline(ss, tok, ',')) {
      int x;
      if (!parse_int(trim(tok), x) || n > AMG_MAX_LEVELS) return fail(AMG_EPARSE, std::string("bad ") + rk[q]);
      vals[n++] = x;
    }
    if (n != ngrid) {
      char buf[128];
      std::snprintf(buf, sizeof buf, "ERROR IN MGsolver/%s ngrid=%4d LEVELS INPUT=%4d", rk[q], ngrid, n);
      return fail(AMG_EPARSE, buf);
    }
    int* dst = q == 0 ? c.pre_relax : c.post_relax;
    for (int i = 0; i < AMG_MAX_LEVELS; ++i) dst[i] = i < n ? vals[i] : vals[n - 1];
  }
  // [BCs]: X1 -> faces 5,6; X2 -> 3,1; X3 -> 2,4 (1-based, mg_read_params.F90:105-214)
  const char* axn[3] = {"X1", "X2", "X3"};
  const int fmin[3] = {4, 2, 1}, fmax_[3] = {5, 0, 3};
  for (int d = 0; d < 3; ++d) {
    if (!need("BCs", axn[d], v)) return AMG_EPARSE;
    const std::string s = lower(trim(*v));
    if (s == "periodic" || s == "per") {
      c.tbc[fmin[d]] = c.tbc[fmax_[d]] = AMG_BC_PERIODIC;
      c.vbc[fmin[d]] = c.vbc[fmax_[d]] = -1.0;
    } else if (s == "read") {
      for (int side = 0; side < 2; ++side) {
        const std::string blk = std::string(axn[d]) + (side == 0 ? "min" : "max");
        const int f = side == 0 ? fmin[d] : fmax_[d];
        if (!need(blk.c_str(), "PTYPE", v)) return AMG_EPARSE;
        const std::string t = lower(trim(*v));
        if (t == "dirichlet" || t == "dir") c.tbc[f] = AMG_BC_DIRICHLET;
        else if (t == "neumann" || t == "neu") c.tbc[f] = AMG_BC_NEUMANN;
        else return fail(AMG_EPARSE, "NO VALID BC'S FOUND FOR P AT " + blk + ", FATAL! ");
        if (!need(blk.c_str(), "PVALUE", v) || !parse_real(*v, c.vbc[f]))
          return fail(AMG_EPARSE, "bad PVALUE in [" + blk + "]");
      }
    } else {
      return fail(AMG_EPARSE, std::string("bad BC type for ") + axn[d]);
    }
  }
  *cfg = c;
  return AMG_OK;
}

int amg_create(const amg_config* cfg, const amg_grid* grid, amg_handle* out) {
  if (!cfg || !grid || !out) return fail(AMG_EINVAL, "amg_create: NULL argument");
  *out = nullptr;
  if (cfg->smoother != AMG_SMOOTHER_RBPGS && cfg->smoother != AMG_SMOOTHER_LJ)
    return fail(AMG_EINVAL, "amg_create: unknown smoother");
  if (cfg->precision != AMG_FP64 && cfg->precision != AMG_MIXED)
    return fail(AMG_EINVAL, "amg_create: unknown precision");
  if (cfg->maxcy < 1 || !(cfg->tol >= 0.0)) return fail(AMG_EINVAL, "amg_create: bad maxcy/tol");
  // the grid is validated on the host before any device work
  Hierarchy H;
  int rc = check_grid(*cfg, *grid, H);
  if (rc != AMG_OK) return rc;
  amg_solver_s* h = new amg_solver_s();
  rc = h->ctx.init(*cfg);
  if (rc == AMG_OK) rc = h->ctx.add_instance(*grid, &h->I, &H);
  if (rc != AMG_OK) {
    delete h;
    return rc;
  }
  *out = h;
  return AMG_OK;
}

int amg_destroy(amg_handle h) {
  if (!h) return AMG_OK;
  if (h->d_phi_stage) (void)hipFree(h->d_phi_stage);
  if (h->d_rhs_stage) (void)hipFree(h->d_rhs_stage);
  if (h->d_div_partial) (void)hipFree(h->d_div_partial);
  if (h->d_cn) (void)hipFree(h->d_cn);
  if (h->h_cn) (void)hipHostFree(h->h_cn);
  delete h;
  return AMG_OK;
}

void* amg_stream(amg_handle h) { return h ? (void*)h->ctx.stream : nullptr; }

int amg_levels(amg_handle h, int* nlevels, int sizes[][3]) {
  if (!h || !nlevels) return fail(AMG_EINVAL, "amg_levels: NULL argument");
  const auto& lv = h->I->H.lv;
  *nlevels = (int)lv.size();
  if (sizes)
    for (size_t l = 0; l < lv.size(); ++l) {
      sizes[l][0] = lv[l].nx; sizes[l][1] = lv[l].ny; sizes[l][2] = lv[l].nz;
    }
  return AMG_OK;
}

int amg_direct_level(amg_handle h) { return h ? h->I->ld : -1; }

static int finish_solve(amg_handle h, int rc, const amg_stats& st, int* iter, double* relmax) {
  if (iter) *iter = st.iters;
  if (relmax) *relmax = st.relmax;
  if (rc == AMG_ENOTCONV) {
    // mgMK/mg_solver.F90:131-132
    std::fprintf(stderr, "WARNING: failed to achieve  convergence in  %6d cycles, RESIDUAL=%22.16f\n",
                 h->ctx.cfg.maxcy,
                 h->ctx.cfg.stop_rule == AMG_STOP_RELMAX ? st.relmax : st.relres);
  }
  return rc;
}

int amg_solve_device(amg_handle h, double* d_phif, const double* d_rhsf, int* iter, double* relmax,
                     amg_stats* st) {
  if (!h || !d_phif || !d_rhsf) return fail(AMG_EINVAL, "amg_solve_device: NULL argument");
  h->I->d_phi = d_phif;
  h->I->d_rhs = d_rhsf;
  amg_stats local;
  amg_stats* s = st ? st : &local;
  std::vector<Instance*> S{h->I};
  const int rc = h->ctx.solve(S, s);
  if (rc == AMG_EHIP) return rc;
  return finish_solve(h, rc, *s, iter, relmax);
}

int amg_solve_host(amg_handle h, double* phif, const double* rhsf, int* iter, double* relmax,
                   amg_stats* st) {
  if (!h || !phif || !rhsf) return fail(AMG_EINVAL, "amg_solve_host: NULL argument");
  const size_t n = (size_t)(h->I->nx + 2) * (h->I->ny + 2) * (h->I->nz + 2);
  if (!h->d_phi_stage) {
    AMG_HIP(hipMalloc((void**)&h->d_phi_stage, n * 8));
    AMG_HIP(hipMalloc((void**)&h->d_rhs_stage, n * 8));
  }
  hipStream_t s = h->ctx.stream;
  AMG_HIP(hipMemcpyAsync(h->d_phi_stage, phif, n * 8, hipMemcpyHostToDevice, s));
  AMG_HIP(hipMemcpyAsync(h->d_rhs_stage, rhsf, n * 8, hipMemcpyHostToDevice, s));
  const int rc = amg_solve_device(h, h->d_phi_stage, h->d_rhs_stage, iter, relmax, st);
  if (rc == AMG_EHIP || rc == AMG_EINVAL) return rc;
  AMG_HIP(hipMemcpyAsync(phif, h->d_phi_stage, n * 8, hipMemcpyDeviceToHost, s));
  AMG_HIP(hipStreamSynchronize(s));
  return rc;
}

// ---- projection ----------------------------------------------------------
int amg_psource_device(amg_handle h, const double* U, const double* V, const double* W, double* SRC,
                       double dt) {
  if (!h || !U || !V || !W || !SRC || dt == 0.0) return fail(AMG_EINVAL, "amg_psource_device: bad argument");
  const double *rho, *rdxc, *rdyc, *rdzc, *rdxe, *rdye, *rdze;
  metric_ptrs(h->I, &rho, &rdxc, &rdyc, &rdzc, &rdxe, &rdye, &rdze);
  launch_psource(U, V, W, SRC, rho, rdxc, rdyc, rdzc, h->I->nx, h->I->ny, h->I->nz, 1.0 / dt,
                 h->ctx.stream);
  AMG_HIP(hipGetLastError());
  AMG_HIP(hipStreamSynchronize(h->ctx.stream));
  return AMG_OK;
}

int amg_pgrad_device(amg_handle h, const double* p, double* U, double* V, double* W, double dt) {
  if (!h || !p || !U || !V || !W) return fail(AMG_EINVAL, "amg_pgrad_device: NULL argument");
  const double *rho, *rdxc, *rdyc, *rdzc, *rdxe, *rdye, *rdze;
  metric_ptrs(h->I, &rho, &rdxc, &rdyc, &rdzc, &rdxe, &rdye, &rdze);
  const int nx = h->I->nx, ny = h->I->ny, nz = h->I->nz;
  hipStream_t s = h->ctx.stream;
  const amg_config& cf = h->ctx.cfg;
  const int wallmask = (cf.tbc[5] == AMG_BC_NEUMANN ? 1 : 0) | (cf.tbc[0] == AMG_BC_NEUMANN ? 2 : 0) |
                       (cf.tbc[3] == AMG_BC_NEUMANN ? 4 : 0);
  launch_pgrad(p, U, V, W, rho, rdxe, rdye, rdze, nx, ny, nz, dt, wallmask, s);
  // periodic velocity ghosts (the periodic part of ghost(), dns/BC.F90)
  const int* bc = &h->I->H.bc[0][0];
  for (int d = 0; d < 3; ++d)
    if (bc[2 * d] == AMG_BC_PERIODIC)
      for (double* a : {U, V, W}) launch_ghost_wrap(a, nx, ny, nz, d, s);
  AMG_HIP(hipGetLastError());
  AMG_HIP(hipStreamSynchronize(s));
  return AMG_OK;
}

int amg_ghost_device(amg_handle h, double* var, const char* vtype, const amg_flowbc* bc) {
  if (!h || !var || !vtype || !bc) return fail(AMG_EINVAL, "amg_ghost_device: NULL argument");
  const FlowGrid G = flow_grid(h->I);
  GhostPlan P;
  std::string err;
  if (plan_ghost(G, vtype, *bc, &P, &err) != 0) return fail(AMG_EINVAL, "amg_ghost_device: " + err);
  launch_flow_ghost(var, G.nx, G.ny, G.nz, P, h->ctx.stream);
  AMG_HIP(hipGetLastError());
  AMG_HIP(hipStreamSynchronize(h->ctx.stream));
  return AMG_OK;
}

int amg_pgrad_ghost_device(amg_handle h, const double* p, double* U, double* V, double* W, double dt,
                           const amg_flowbc* bc) {
  if (!h || !p || !U || !V || !W || !bc) return fail(AMG_EINVAL, "amg_pgrad_ghost_device: NULL argument");
  const FlowGrid G = flow_grid(h->I);
  GhostPlan P[3];
  std::string err;
  const char* names[3] = {"utemp", "vtemp", "wtemp"};
  for (int c = 0; c < 3; ++c)
    if (plan_ghost(G, names[c], *bc, &P[c], &err) != 0)
      return fail(AMG_EINVAL, "amg_pgrad_ghost_device: " + err);
  const double *rho, *rdxc, *rdyc, *rdzc, *rdxe, *rdye, *rdze;
  metric_ptrs(h->I, &rho, &rdxc, &rdyc, &rdzc, &rdxe, &rdye, &rdze);
  const amg_config& cf = h->ctx.cfg;
  const int wallmask = (cf.tbc[5] == AMG_BC_NEUMANN ? 1 : 0) | (cf.tbc[0] == AMG_BC_NEUMANN ? 2 : 0) |
                       (cf.tbc[3] == AMG_BC_NEUMANN ? 4 : 0);
  hipStream_t s = h->ctx.stream;
  launch_pgrad(p, U, V, W, rho, rdxe, rdye, rdze, G.nx, G.ny, G.nz, dt, wallmask, s);
  double* f[3] = {U, V, W};
  for (int c = 0; c < 3; ++c) launch_flow_ghost(f[c], G.nx, G.ny, G.nz, P[c], s);
  AMG_HIP(hipGetLastError());
  AMG_HIP(hipStreamSynchronize(s));
  return AMG_OK;
}

int amg_cn_advance_device(amg_handle h, const double* veltemp1, double* veltemp2, double f1,
                          double f2, double f3, double rkdt, int var, const double* p,
                          const amg_flowbc* bc, double rRe, double rPr, double rSc) {
  if (!h || !veltemp1 || !veltemp2 || !bc) return fail(AMG_EINVAL, "amg_cn_advance_device: NULL argument");
  if (veltemp1 == veltemp2) return fail(AMG_EINVAL, "amg_cn_advance_device: veltemp1 and veltemp2 alias");
  static const char* ghost_type[9] = {nullptr, "u", "v", "w", nullptr, "rho", nullptr, nullptr, "scal1"};
  if (var < 1 || var > 8 || !ghost_type[var])
    return fail(AMG_EINVAL, "amg_cn_advance_device: ABORTING ADVANCE CN, VARID IS NOT VALID");
  const FlowGrid G = flow_grid(h->I);
  std::vector<CnRow> rows;
  std::string err;
  if (plan_cn(G, var, rkdt, rRe, rPr, rSc, *bc, &rows, &err) != 0)
    return fail(AMG_EINVAL, "amg_cn_advance_device: " + err);
  GhostPlan P;
  if (plan_ghost(G, ghost_type[var], *bc, &P, &err) != 0)
    return fail(AMG_EINVAL, "amg_cn_advance_device: " + err);
  hipStream_t s = h->ctx.stream;
  const size_t nb = sizeof(CnRow) * rows.size();
  if (!h->d_cn) {
    AMG_HIP(hipMalloc((void**)&h->d_cn, nb));
    AMG_HIP(hipHostMalloc((void**)&h->h_cn, nb, hipHostMallocDefault));
  }
  std::memcpy(h->h_cn, rows.data(), nb);
  AMG_HIP(hipMemcpyAsync(h->d_cn, h->h_cn, nb, hipMemcpyHostToDevice, s));
  const double *rho, *rdxc, *rdyc, *rdzc, *rdxe, *rdye, *rdze;
  metric_ptrs(h->I, &rho, &rdxc, &rdyc, &rdzc, &rdxe, &rdye, &rdze);
  launch_cn_column(veltemp1, veltemp2, p, h->d_cn, rdxe, rdye, G.nx, G.ny, G.nz, f1, f2, f3, rkdt, var,
                   s);
  launch_flow_ghost(veltemp2, G.nx, G.ny, G.nz, P, s);
  AMG_HIP(hipGetLastError());
  AMG_HIP(hipStreamSynchronize(s));
  return AMG_OK;
}

int amg_divergence_device(amg_handle h, const double* U, const double* V, const double* W,
                          double* div, double* avg, double* maxabs, int loc[3]) {
  if (!h || !U || !V || !W) return fail(AMG_EINVAL, "amg_divergence_device: NULL argument");
  const int nb = divergence_blocks(h->I->nx, h->I->ny, h->I->nz);
  if (!h->d_div_partial)
    AMG_HIP(hipMalloc((void**)&h->d_div_partial, sizeof(double) * (4 * (size_t)nb + 4)));
  return divergence_report(h, U, V, W, div, h->d_div_partial, avg, maxabs, loc);
}

// ---- batch ---------------------------------------------------------------
int amg_batch_create(int n, const amg_config* cfg, const amg_grid* grids, amg_batch* out) {
  if (n <= 0 || n > kMaxInstances || !cfg || !grids || !out)
    return fail(AMG_EINVAL, "amg_batch_create: bad argument");
  *out = nullptr;
  amg_batch_s* b = new amg_batch_s();
  amg_config c = *cfg;
  c.use_graph = 0;
  int rc = b->ctx.init(c);
  for (int i = 0; i < n && rc == AMG_OK; ++i) {
    Instance* I = nullptr;
    rc = b->ctx.add_instance(grids[i], &I);
  }
  if (rc != AMG_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return AMG_OK;
}

int amg_batch_destroy(amg_batch b) {
  delete b;
  return AMG_OK;
}

int amg_batch_solve_device(amg_batch b, double* const* d_phi, const double* const* d_rhs, int* iters,
                           double* relres, amg_stats* total) {
  if (!b || !d_phi || !d_rhs) return fail(AMG_EINVAL, "amg_batch_solve_device: NULL argument");
  auto& S = b->ctx.inst;
  for (size_t i = 0; i < S.size(); ++i) {
    if (!d_phi[i] || !d_rhs[i]) return fail(AMG_EINVAL, "amg_batch_solve_device: NULL field");
    S[i]->d_phi = d_phi[i];
    S[i]->d_rhs = d_rhs[i];
  }
  amg_stats st;
  const int rc = b->ctx.solve(S, &st);
  if (rc == AMG_EHIP) return rc;
  for (size_t i = 0; i < S.size(); ++i) {
    if (iters) iters[i] = S[i]->converged ? S[i]->iters : b->ctx.cfg.maxcy + 1;
    if (relres) relres[i] = S[i]->relres;
  }
  if (total) *total = st;
  return rc;
}

int amg_run_kernel(amg_handle h, int which, int level, int precision, int reps, double* bytes) {
  if (!h) return fail(AMG_EINVAL, "amg_run_kernel: NULL handle");
  if ((precision == AMG_FP64) != h->I->fp64)
    return fail(AMG_EINVAL, "amg_run_kernel: the handle was created with the other precision");
  return h->ctx.run_kernel(h->I, which, level, reps, bytes);
}

int amg_level_field(amg_handle h, int level, int field, int dir, double* buf) {
  if (!h) return fail(AMG_EINVAL, "amg_level_field: NULL handle");
  return h->ctx.level_field(h->I, level, field, dir, buf);
}

const char* amg_strerror(int code) {
  switch (code) {
    case AMG_OK: return "ok";
    case AMG_EINVAL: return "invalid argument";
    case AMG_EBADSIZE: return "grid sizes violate the level constraints";
    case AMG_EBADBC: return "bad boundary condition";
    case AMG_ENOMEM: return "out of device memory / HBM budget";
    case AMG_ENOTCONV: return "not converged in maxcy cycles";
    case AMG_ENAN: return "NaN or Inf encountered";
    case AMG_EHIP: return "HIP runtime error";
    case AMG_EUNSUPPORTED: return "unsupported";
    case AMG_EPARSE: return "INI parse error";
    case AMG_ENOGPU: return "no HIP device";
    case AMG_EBUSY: return "busy / not ready";
    default: return "unknown error";
  }
}

const char* amg_last_error(void) { return last_error().c_str(); }

}  // extern "C"
