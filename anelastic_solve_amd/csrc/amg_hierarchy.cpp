// amg_hierarchy.cpp -- level hierarchy of the multigrid pressure solve.
//
// Replaces the reference's per-solve coefficient rebuild (mgdpfpde/mgdphpde,
// mgMK/mg_pfpde.F90, mgMK/mg_phpde.F90: seven 3-D arrays per level) by
// 1-D tables built once: the stretched-metric stencil of mgMK/mg_coeff.F90
// with the density factors of mgMK/mg_pfpde.F90, re-discretised on every
// level from agglomerated metrics (instead of mgMK/mg_grids.F90).
#include "amg_hierarchy.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>

namespace amg {

Axis axis_from_faces(const std::vector<double>& F) {
  Axis a;
  const int n = (int)F.size() - 1;
  a.n = n;
  a.xe.assign(n + 2, 0.0);
  a.xc.assign(n + 2, 0.0);
  a.dxe.assign(n + 2, 0.0);
  a.dxc.assign(n + 2, 0.0);
  for (int i = 1; i <= n; ++i) a.dxc[i] = F[i] - F[i - 1];
  a.dxc[0] = a.dxc[1];
  a.dxc[n + 1] = a.dxc[n];
  for (int i = 0; i <= n; ++i) a.xe[i] = F[i];
  a.xe[n + 1] = F[n] + a.dxc[n + 1];
  for (int i = 1; i <= n; ++i) a.xc[i] = 0.5 * (F[i] + F[i - 1]);
  a.xc[n + 1] = F[n] + 0.5 * a.dxc[n + 1];
  for (int i = 1; i <= n; ++i) a.dxe[i] = a.xc[i + 1] - a.xc[i];
  // ghost fix-ups of utils/loadGrid.F90:79-90
  a.dxe[0] = a.dxe[1];
  a.dxe[n + 1] = a.dxe[n];
  a.xc[0] = a.xc[1] - a.dxe[0];
  return a;
}

Axis axis_from_arrays(int n, const double* xe, const double* xc, const double* dxe,
                      const double* dxc) {
  Axis a;
  a.n = n;
  a.xe.assign(xe, xe + n + 2);
  a.xc.assign(xc, xc + n + 2);
  a.dxe.assign(dxe, dxe + n + 2);
  a.dxc.assign(dxc, dxc + n + 2);
  return a;
}

Axis coarsen_axis(const Axis& a, int factor, bool wall_lo, bool wall_hi) {
  if (factor == 1) return a;
  std::vector<double> F;
  for (int i = 0; i <= a.n; i += 2) F.push_back(a.xe[i]);
  Axis c = axis_from_faces(F);
  // At a wall the ghost centre stays where the finest level has it: a
  // Dirichlet value sits at that point, and a coarse correction must vanish
  // at the same place.  A coarse ghost half a coarse cell out would move the
  // wall outward at every coarsening; with several z-only coarsenings the
  // coarse-grid correction then no longer matches the fine problem and the
  // V-cycle diverges.  (Neumann walls carry no flux, so the distance is
  // unused there.)
  const int n = c.n;
  if (wall_lo) {
    c.xc[0] = a.xc[0];
    c.dxe[0] = c.xc[1] - c.xc[0];
  }
  if (wall_hi) {
    c.xc[n + 1] = a.xc[a.n + 1];
    c.dxe[n] = c.xc[n + 1] - c.xc[n];
  }
  return c;
}

namespace {

// Horizontal coefficients, mgMK/mg_coeff.F90:78-91 (+ Neumann zeroing :119-145).
void axis_coefs(const Axis& a, int bc0, int bc1, double* wm, double* wp, double* d,
                std::vector<double>& am, std::vector<double>& ap) {
  const int n = a.n;
  am.assign(n, 0.0);
  ap.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    am[i] = 1.0 / (a.dxc[i + 1] * a.dxe[i]);
    ap[i] = 1.0 / (a.dxc[i + 1] * a.dxe[i + 1]);
    wm[i] = am[i];
    wp[i] = ap[i];
    d[i] = -(am[i] + ap[i]);
  }
  if (bc0 != AMG_BC_PERIODIC) {
    wm[0] = 0.0;
    if (bc0 == AMG_BC_NEUMANN) d[0] += am[0];
  }
  if (bc1 != AMG_BC_PERIODIC) {
    wp[n - 1] = 0.0;
    if (bc1 == AMG_BC_NEUMANN) d[n - 1] += ap[n - 1];
  }
}

// Vertical coefficients of D(rho G(p/rho)): dns/NSsolver.F90:2212-2214 with
// :2076-2078; density factors as mgMK/mg_pfpde.F90:49-57.
void z_coefs(const Axis& a, const std::vector<double>& r, int bc0, int bc1, double* wm,
             double* wp, double* d, std::vector<double>& am, std::vector<double>& ap) {
  const int n = a.n;
  am.assign(n, 0.0);
  ap.assign(n, 0.0);
  std::vector<double> gm(n), gp(n);
  for (int k = 0; k < n; ++k) {
    const double rk = r[k + 1], rm = r[k], rp = r[k + 2];
    gm[k] = 1.0 / (a.dxc[k + 1] * a.dxe[k]) * 0.5 * (rk + rm);
    gp[k] = 1.0 / (a.dxc[k + 1] * a.dxe[k + 1]) * 0.5 * (rk + rp);
    am[k] = gm[k] / rm;
    ap[k] = gp[k] / rp;
    wm[k] = am[k];
    wp[k] = ap[k];
    d[k] = -(gm[k] / rk + gp[k] / rk);
  }
  if (bc0 != AMG_BC_PERIODIC) {
    wm[0] = 0.0;
    if (bc0 == AMG_BC_NEUMANN) d[0] += gm[0] / r[1];
  }
  if (bc1 != AMG_BC_PERIODIC) {
    wp[n - 1] = 0.0;
    if (bc1 == AMG_BC_NEUMANN) d[n - 1] += gp[n - 1] / r[n];
  }
}

std::vector<double> coarsen_rho(const std::vector<double>& r, const Axis& az, bool per, int f) {
  if (f == 1) return r;
  const int n = az.n, nc = n / 2;
  std::vector<double> out(nc + 2);
  for (int K = 1; K <= nc; ++K) {
    const double wa = az.dxc[2 * K - 1], wb = az.dxc[2 * K];
    out[K] = (wa * r[2 * K - 1] + wb * r[2 * K]) / (wa + wb);
  }
  if (per) { out[0] = out[nc]; out[nc + 1] = out[1]; }
  else { out[0] = out[1]; out[nc + 1] = out[nc]; }
  return out;
}

void fill_level(HostLevel& L, const int bc[3][2]) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  L.co.assign(3 * (nx + ny + nz), 0.0);
  double* cx = L.co.data();
  double* cy = cx + 3 * nx;
  double* cz = cy + 3 * ny;
  axis_coefs(L.ax, bc[0][0], bc[0][1], cx, cx + nx, cx + 2 * nx, L.amx, L.apx);
  axis_coefs(L.ay, bc[1][0], bc[1][1], cy, cy + ny, cy + 2 * ny, L.amy, L.apy);
  z_coefs(L.az, L.rho, bc[2][0], bc[2][1], cz, cz + nz, cz + 2 * nz, L.amz, L.apz);
  L.vol.assign(nx + ny + nz, 0.0);
  for (int i = 0; i < nx; ++i) L.vol[i] = L.ax.dxc[i + 1];
  for (int j = 0; j < ny; ++j) L.vol[nx + j] = L.ay.dxc[j + 1];
  for (int k = 0; k < nz; ++k) L.vol[nx + ny + k] = L.az.dxc[k + 1];
}

// Dense matrix of the level operator in chequer order (row-major, n x n).
std::vector<double> dense_operator(const HostLevel& L, const int bc[3][2]) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const long long n = L.ncell();
  std::vector<double> A((size_t)(n * n), 0.0);
  const double* cx = L.co.data();
  const double* cy = cx + 3 * nx;
  const double* cz = cy + 3 * ny;
  const bool px = bc[0][0] == AMG_BC_PERIODIC, py = bc[1][0] == AMG_BC_PERIODIC,
             pz = bc[2][0] == AMG_BC_PERIODIC;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long r = chequer_index(i, j, k, nx, ny);
        auto add = [&](int ii, int jj, int kk, double w) {
          if (w == 0.0) return;
          ii = (ii + nx) % nx; jj = (jj + ny) % ny; kk = (kk + nz) % nz;
          A[(size_t)(r * n + chequer_index(ii, jj, kk, nx, ny))] += w;
        };
        add(i, j, k, cz[2 * nz + k] + cy[2 * ny + j] + cx[2 * nx + i]);
        if (px || i > 0) add(i - 1, j, k, cx[i]);
        if (px || i < nx - 1) add(i + 1, j, k, cx[nx + i]);
        if (py || j > 0) add(i, j - 1, k, cy[j]);
        if (py || j < ny - 1) add(i, j + 1, k, cy[ny + j]);
        if (pz || k > 0) add(i, j, k - 1, cz[k]);
        if (pz || k < nz - 1) add(i, j, k + 1, cz[nz + k]);
      }
  return A;
}

// Gauss-Jordan inverse with partial pivoting (FP64), in place.
bool invert(std::vector<double>& M, int n) {
  std::vector<double> I((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    double best = std::fabs(M[(size_t)c * n + c]);
    for (int r = c + 1; r < n; ++r) {
      const double v = std::fabs(M[(size_t)r * n + c]);
      if (v > best) { best = v; p = r; }
    }
    if (best == 0.0 || !std::isfinite(best)) return false;
    if (p != c) {
      for (int q = 0; q < n; ++q) {
        std::swap(M[(size_t)c * n + q], M[(size_t)p * n + q]);
        std::swap(I[(size_t)c * n + q], I[(size_t)p * n + q]);
      }
    }
    const double inv = 1.0 / M[(size_t)c * n + c];
    for (int q = 0; q < n; ++q) {
      M[(size_t)c * n + q] *= inv;
      I[(size_t)c * n + q] *= inv;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double fct = M[(size_t)r * n + c];
      if (fct == 0.0) continue;
      for (int q = 0; q < n; ++q) {
        M[(size_t)r * n + q] -= fct * M[(size_t)c * n + q];
        I[(size_t)r * n + q] -= fct * I[(size_t)c * n + q];
      }
    }
  }
  M.swap(I);
  return true;
}

}  // namespace

void sweep_counts(const amg_config& cfg, int nlev, std::vector<int>& pre, std::vector<int>& post) {
  // The reference's PreRelax/PstRelax are coarsest-first (mgMK/mg_read_params.F90:67-80):
  // finest-first level l reads entry nlev-1-l.
  pre.assign(nlev, 2);
  post.assign(nlev, 2);
  for (int l = 0; l < nlev; ++l) {
    const int rc = nlev - 1 - l;
    if (rc < AMG_MAX_LEVELS) {
      pre[l] = cfg.pre_relax[rc] < 0 ? 0 : cfg.pre_relax[rc];
      post[l] = cfg.post_relax[rc] < 0 ? 0 : cfg.post_relax[rc];
    }
  }
}

int build_hierarchy(const amg_config& cfg, const amg_grid& g, Hierarchy& h, std::string& err) {
  char buf[256];
  if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0 || !g.xe || !g.xc || !g.dxe || !g.dxc || !g.ye ||
      !g.yc || !g.dye || !g.dyc || !g.ze || !g.zc || !g.dze || !g.dzc) {
    err = "amg_create: grid sizes must be positive and all metric arrays given";
    return AMG_EINVAL;
  }
  for (int f = 0; f < 6; ++f)
    if (cfg.tbc[f] < 0 || cfg.tbc[f] > 2) {
      std::snprintf(buf, sizeof buf, "MG INITIALIZATION FAILED BAD BOUNDARY CONDITION CHECK %d", f + 1);
      err = buf;
      return AMG_EBADBC;
    }
  // face map (mgMK/mg_init.F90:265-305): X1 = (5,6), X2 = (3,1), X3 = (2,4)
  h.bc[0][0] = cfg.tbc[4]; h.bc[0][1] = cfg.tbc[5];
  h.bc[1][0] = cfg.tbc[2]; h.bc[1][1] = cfg.tbc[0];
  h.bc[2][0] = cfg.tbc[1]; h.bc[2][1] = cfg.tbc[3];
  for (int d = 0; d < 3; ++d)
    if ((h.bc[d][0] == AMG_BC_PERIODIC) != (h.bc[d][1] == AMG_BC_PERIODIC)) {
      err = "amg_create: a direction must be periodic on both faces or on neither";
      return AMG_EBADBC;
    }
  h.singular = true;
  for (int f = 0; f < 6; ++f)
    if (cfg.tbc[f] == AMG_BC_DIRICHLET) h.singular = false;

  // transition factors
  std::vector<std::array<int, 3>> fac;
  int nx = g.nx, ny = g.ny, nz = g.nz;
  const bool autolev = cfg.nlevels[0] <= 0 && cfg.nlevels[1] <= 0 && cfg.nlevels[2] <= 0;
  if (autolev) {
    while (nx % 4 == 0 && ny % 2 == 0 && nz % 2 == 0 && std::min(nx, std::min(ny, nz)) > 4) {
      fac.push_back({2, 2, 2});
      nx /= 2; ny /= 2; nz /= 2;
    }
  } else {
    int lv[3];
    for (int d = 0; d < 3; ++d) lv[d] = std::max(1, cfg.nlevels[d]);
    const int ngrid = std::max(lv[0], std::max(lv[1], lv[2]));
    if (ngrid > AMG_MAX_LEVELS) {
      err = "amg_create: more than 16 levels";
      return AMG_EBADSIZE;
    }
    // mgMK/mg_check_size.F90:71-98: n must equal nCp * 2**(levels-1)
    const int n3[3] = {g.nx, g.ny, g.nz};
    for (int d = 0; d < 3; ++d)
      if (n3[d] % (1 << (lv[d] - 1)) != 0) {
        std::snprintf(buf, sizeof buf,
                      "ERROR: mg_check_size: n%d=%d is not a multiple of 2**(levels-1)=%d",
                      d + 1, n3[d], 1 << (lv[d] - 1));
        err = buf;
        return AMG_EBADSIZE;
      }
    for (int t = 0; t < ngrid - 1; ++t)
      fac.push_back({t < lv[0] - 1 ? 2 : 1, t < lv[1] - 1 ? 2 : 1, t < lv[2] - 1 ? 2 : 1});
  }
  if ((int)fac.size() + 1 > AMG_MAX_LEVELS) {
    err = "amg_create: more than 16 levels";
    return AMG_EBADSIZE;
  }

  // levels
  h.lv.clear();
  HostLevel L0;
  L0.nx = g.nx; L0.ny = g.ny; L0.nz = g.nz;
  L0.ax = axis_from_arrays(g.nx, g.xe, g.xc, g.dxe, g.dxc);
  L0.ay = axis_from_arrays(g.ny, g.ye, g.yc, g.dye, g.dyc);
  L0.az = axis_from_arrays(g.nz, g.ze, g.zc, g.dze, g.dzc);
  L0.rho.assign(g.nz + 2, 1.0);
  if (g.rho) L0.rho.assign(g.rho, g.rho + g.nz + 2);
  for (double r : L0.rho)
    if (!(r > 0.0) || !std::isfinite(r)) {
      err = "amg_create: rho must be positive and finite";
      return AMG_EINVAL;
    }
  h.lv.push_back(L0);
  for (size_t t = 0; t < fac.size(); ++t) {
    const HostLevel& F = h.lv.back();
    HostLevel C;
    C.nx = F.nx / fac[t][0]; C.ny = F.ny / fac[t][1]; C.nz = F.nz / fac[t][2];
    const bool wall[3] = {h.bc[0][0] != AMG_BC_PERIODIC, h.bc[1][0] != AMG_BC_PERIODIC,
                          h.bc[2][0] != AMG_BC_PERIODIC};
    C.ax = coarsen_axis(F.ax, fac[t][0], wall[0], wall[0]);
    C.ay = coarsen_axis(F.ay, fac[t][1], wall[1], wall[1]);
    C.az = coarsen_axis(F.az, fac[t][2], wall[2], wall[2]);
    C.rho = coarsen_rho(F.rho, F.az, h.bc[2][0] == AMG_BC_PERIODIC, fac[t][2]);
    h.lv.back().fx = fac[t][0]; h.lv.back().fy = fac[t][1]; h.lv.back().fz = fac[t][2];
    h.lv.push_back(C);
  }
  const int nl = (int)h.lv.size();
  for (int l = 0; l < nl; ++l) {
    HostLevel& L = h.lv[l];
    const bool coarsest = (l == nl - 1);
    const int n3[3] = {L.nx, L.ny, L.nz};
    if (L.nx % 2 != 0 || L.nx < 2) {
      std::snprintf(buf, sizeof buf, "amg_create: level %d has nx=%d; nx must be even on every level", l, L.nx);
      err = buf;
      return AMG_EBADSIZE;
    }
    if (!coarsest) {
      for (int d = 1; d < 3; ++d)
        if (h.bc[d][0] == AMG_BC_PERIODIC && (n3[d] % 2 != 0 || n3[d] < 2)) {
          std::snprintf(buf, sizeof buf,
                        "amg_create: level %d: periodic direction %d needs an even size (got %d)", l,
                        d + 1, n3[d]);
          err = buf;
          return AMG_EBADSIZE;
        }
      const int f3[3] = {L.fx, L.fy, L.fz};
      for (int d = 0; d < 3; ++d)
        if (f3[d] == 2 && n3[d] % 2 != 0) {
          err = "amg_create: coarsening an odd size";
          return AMG_EBADSIZE;
        }
    }
    fill_level(L, h.bc);
  }
  const HostLevel& C = h.lv.back();
  const long long nc = C.ncell();
  if (nc > 1024) {
    std::snprintf(buf, sizeof buf,
                  "amg_create: coarsest level %dx%dx%d has %lld cells; at most 1024 are solved "
                  "directly (use more levels)", C.nx, C.ny, C.nz, nc);
    err = buf;
    return AMG_EBADSIZE;
  }

  // boundary-value folds on the finest level (mgMK/mg_bdry.F90:103-190 with phibc = vbc)
  {
    HostLevel& F = h.lv[0];
    F.bcfold.assign(F.nx + F.ny + F.nz, 0.0);
    double* bx = F.bcfold.data();
    double* by = bx + F.nx;
    double* bz = by + F.ny;
    const int* t = cfg.tbc;
    const double* v = cfg.vbc;
    auto fold = [](int type, double am_or_ap, double val, double width, double sign) {
      if (type == AMG_BC_NEUMANN) return sign * am_or_ap * val * width;
      if (type == AMG_BC_DIRICHLET) return -am_or_ap * val;
      return 0.0;
    };
    bx[0] += fold(t[4], F.amx[0], v[4], F.ax.dxc[1], +1.0);
    bx[F.nx - 1] += fold(t[5], F.apx[F.nx - 1], v[5], F.ax.dxc[F.nx], -1.0);
    by[0] += fold(t[2], F.amy[0], v[2], F.ay.dxc[1], +1.0);
    by[F.ny - 1] += fold(t[0], F.apy[F.ny - 1], v[0], F.ay.dxc[F.ny], -1.0);
    bz[0] += fold(t[1], F.amz[0], v[1], F.az.dxc[1], +1.0);
    bz[F.nz - 1] += fold(t[3], F.apz[F.nz - 1], v[3], F.az.dxc[F.nz], -1.0);
    h.vol_sum = 0.0;
    h.vol_rho_sum = 0.0;
    double sx = 0, sy = 0, sz = 0, szr = 0;
    for (int i = 0; i < F.nx; ++i) sx += F.vol[i];
    for (int j = 0; j < F.ny; ++j) sy += F.vol[F.nx + j];
    for (int k = 0; k < F.nz; ++k) {
      sz += F.vol[F.nx + F.ny + k];
      szr += F.vol[F.nx + F.ny + k] * F.rho[k + 1];
    }
    h.vol_sum = sx * sy * sz;
    h.vol_rho_sum = sx * sy * szr;
  }

  // coarsest inverse (bordered when singular)
  {
    const int n = (int)nc;
    std::vector<double> A = dense_operator(C, h.bc);
    int m = n;
    std::vector<double> M;
    if (h.singular) {
      m = n + 1;
      M.assign((size_t)m * m, 0.0);
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) M[(size_t)r * m + c] = A[(size_t)r * n + c];
      for (int k = 0; k < C.nz; ++k)
        for (int j = 0; j < C.ny; ++j)
          for (int i = 0; i < C.nx; ++i) {
            const long long r = chequer_index(i, j, k, C.nx, C.ny);
            M[(size_t)r * m + n] = 1.0;
            M[(size_t)n * m + r] = C.vol[i] * C.vol[C.nx + j] * C.vol[C.nx + C.ny + k];
          }
    } else {
      M = A;
    }
    if (!invert(M, m)) {
      err = "amg_create: coarsest operator is singular";
      return AMG_EBADSIZE;
    }
    h.minvT.assign((size_t)n * n, 0.0);
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) h.minvT[(size_t)c * n + r] = M[(size_t)r * m + c];
  }
  return AMG_OK;
}

}  // namespace amg
