// amg_direct.cpp -- host set-up of the separable direct solve of a coarse level.
//
// A level's operator is a sum of 1-D operators, A = Ax (x) I (x) I +
// I (x) Ay (x) I + I (x) I (x) Az, each a tridiagonal (periodic: cyclic)
// matrix of the level's coefficient table.  Ax = Dx^-1 Sx with Dx the cell
// widths and Sx symmetric (the fluxes 1/dxe between neighbours), so
// Dx^1/2 Ax Dx^-1/2 = Q diag(lambda) Q^T and Ax = Vx diag(lambda) Vx^-1 with
// Vx = Dx^-1/2 Q, Vx^-1 = Q^T Dx^1/2; the same for y, and for z with the
// weights dz / rho (Az = Dz^-1 S R^-1, S symmetric).  In the basis of the
// three eigenvector sets A is diagonal, lambda_x + lambda_y + lambda_z, so
// u = (Vx (x) Vy (x) Vz) diag(1 / lambda) (Vx^-1 (x) Vy^-1 (x) Vz^-1) f:
// three dense transforms each way, which the kernels run on the matrix
// cores (k_direct_* in amg_kernels.hip).
#include <cmath>
#include <string>
#include <vector>

#include "amg_direct.h"

namespace amg {

namespace {

// Cyclic Jacobi eigen-decomposition of the symmetric n x n matrix a
// (row-major, destroyed): a = q diag(w) q^T, q's columns the eigenvectors.
bool jacobi_eigen(std::vector<double>& a, int n, std::vector<double>& w, std::vector<double>& q) {
  q.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) q[(size_t)i * n + i] = 1.0;
  auto A = [&](int r, int c) -> double& { return a[(size_t)r * n + c]; };
  double scale = 0.0;
  for (double v : a) scale = std::fmax(scale, std::fabs(v));
  if (!(scale > 0.0) || !std::isfinite(scale)) return false;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int r = 0; r < n; ++r)
      for (int c = r + 1; c < n; ++c) off = std::fmax(off, std::fabs(A(r, c)));
    if (off <= 1e-14 * scale) {
      w.resize(n);
      for (int i = 0; i < n; ++i) w[i] = A(i, i);
      return true;
    }
    for (int p = 0; p < n; ++p)
      for (int r = p + 1; r < n; ++r) {
        const double apr = A(p, r);
        if (std::fabs(apr) <= 1e-300) continue;
        // rotation that zeroes a(p, r) (Golub and Van Loan, Matrix
        // Computations, sec. 8.5: the symmetric Schur decomposition)
        const double tau = (A(r, r) - A(p, p)) / (2.0 * apr);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
        for (int k = 0; k < n; ++k) {
          const double akp = A(k, p), akr = A(k, r);
          A(k, p) = cs * akp - sn * akr;
          A(k, r) = sn * akp + cs * akr;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A(p, k), ark = A(r, k);
          A(p, k) = cs * apk - sn * ark;
          A(r, k) = sn * apk + cs * ark;
        }
        for (int k = 0; k < n; ++k) {
          const double qkp = q[(size_t)k * n + p], qkr = q[(size_t)k * n + r];
          q[(size_t)k * n + p] = cs * qkp - sn * qkr;
          q[(size_t)k * n + r] = sn * qkp + cs * qkr;
        }
      }
  }
  return false;
}

// The 1-D operator of one direction (coefficient table [wm | wp | d] of
// length n each; periodic wraps, walls carry no coupling) and the weights
// w that make W^1/2 A W^-1/2 symmetric (x, y: the cell widths; z: the cell
// widths over rho, since Az = Dz^-1 S R^-1 with S symmetric); fills the
// forward transform Vinv (row-major: x^ = Vinv x), the inverse transform V
// and the eigenvalues.
bool axis_eigen(int n, const double* wm, const double* wp, const double* d, const double* dxc,
                bool per, std::vector<double>& Vinv, std::vector<double>& V,
                std::vector<double>& lam, std::string* err) {
  std::vector<double> A((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    A[(size_t)i * n + i] += d[i];
    if (per || i > 0) A[(size_t)i * n + (i - 1 + n) % n] += wm[i];
    if (per || i < n - 1) A[(size_t)i * n + (i + 1) % n] += wp[i];
  }
  // B = D^1/2 A D^-1/2 must be symmetric
  std::vector<double> B((size_t)n * n);
  double amax = 0.0;
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) {
      B[(size_t)r * n + c] = std::sqrt(dxc[r]) * A[(size_t)r * n + c] / std::sqrt(dxc[c]);
      amax = std::fmax(amax, std::fabs(B[(size_t)r * n + c]));
    }
  for (int r = 0; r < n; ++r)
    for (int c = r + 1; c < n; ++c) {
      const double x = B[(size_t)r * n + c], y = B[(size_t)c * n + r];
      if (std::fabs(x - y) > 1e-12 * amax) {
        if (err) *err = "direct solve: the 1-D operator is not symmetrizable";
        return false;
      }
      B[(size_t)r * n + c] = B[(size_t)c * n + r] = 0.5 * (x + y);
    }
  std::vector<double> Q;
  if (!jacobi_eigen(B, n, lam, Q)) {
    if (err) *err = "direct solve: eigen-decomposition did not converge";
    return false;
  }
  Vinv.assign((size_t)n * n, 0.0);
  V.assign((size_t)n * n, 0.0);
  for (int e = 0; e < n; ++e)
    for (int i = 0; i < n; ++i) {
      const double qie = Q[(size_t)i * n + e];
      Vinv[(size_t)e * n + i] = qie * std::sqrt(dxc[i]);   // (Q^T D^1/2)[e][i]
      V[(size_t)i * n + e] = qie / std::sqrt(dxc[i]);      // (D^-1/2 Q)[i][e]
    }
  return true;
}

int argmin_abs(const std::vector<double>& v) {
  int b = 0;
  for (int i = 1; i < (int)v.size(); ++i)
    if (std::fabs(v[i]) < std::fabs(v[b])) b = i;
  return b;
}

}  // namespace

bool direct_setup(const HostLevel& L, const int bc[3][2], bool singular, DirectTables* out,
                  std::string* err) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  if (bc[2][0] == AMG_BC_PERIODIC) {
    if (err) *err = "direct solve: periodic z";
    return false;
  }
  const int nzp = (nz + 15) & ~15;
  const double* cx = L.co.data();
  const double* cy = cx + 3 * nx;
  const double* cz = cy + 3 * ny;
  std::vector<double> wz(nz);
  for (int k = 0; k < nz; ++k) wz[k] = L.vol[nx + ny + k] / L.rho[k + 1];
  std::vector<double> Vxi, Vx, lx, Vyi, Vy, ly, Vzi, Vz, lz;
  if (!axis_eigen(nx, cx, cx + nx, cx + 2 * nx, L.vol.data(), bc[0][0] == AMG_BC_PERIODIC, Vxi, Vx,
                  lx, err) ||
      !axis_eigen(ny, cy, cy + ny, cy + 2 * ny, L.vol.data() + nx, bc[1][0] == AMG_BC_PERIODIC, Vyi,
                  Vy, ly, err) ||
      !axis_eigen(nz, cz, cz + nz, cz + 2 * nz, wz.data(), false, Vzi, Vz, lz, err))
    return false;
  int sx0 = -1, sy0 = -1, sz0 = -1;
  if (singular) {
    // a singular operator (no Dirichlet face): one zero eigenvalue per
    // direction (constants in x and y, rho in z); that mode is dropped
    // (its coefficient of the solution is zero: sum V x = 0, the
    // null-space convention of the whole solver)
    sx0 = argmin_abs(lx);
    sy0 = argmin_abs(ly);
    sz0 = argmin_abs(lz);
    double mx = 0.0, my = 0.0, mz = 0.0;
    for (double v : lx) mx = std::fmax(mx, std::fabs(v));
    for (double v : ly) my = std::fmax(my, std::fabs(v));
    for (double v : lz) mz = std::fmax(mz, std::fabs(v));
    if (std::fabs(lx[sx0]) > 1e-10 * mx || std::fabs(ly[sy0]) > 1e-10 * my ||
        std::fabs(lz[sz0]) > 1e-10 * mz) {
      if (err) *err = "direct solve: no zero eigenvalue in a singular operator";
      return false;
    }
  }
  // [Vx^-T | Vx^T | Vy^-1 | Vy | Vz^-1 | Vz], the x matrices transposed so
  // that a plane's transform reads them along rows; the z matrices padded
  // to nzp = nz rounded up to 16 with zeros
  auto& m = out->mats;
  m.clear();
  for (int i = 0; i < nx; ++i)
    for (int e = 0; e < nx; ++e) m.push_back(Vxi[(size_t)e * nx + i]);
  for (int e = 0; e < nx; ++e)
    for (int i = 0; i < nx; ++i) m.push_back(Vx[(size_t)i * nx + e]);
  m.insert(m.end(), Vyi.begin(), Vyi.end());
  m.insert(m.end(), Vy.begin(), Vy.end());
  for (int r = 0; r < nzp; ++r)
    for (int c = 0; c < nzp; ++c) m.push_back(r < nz && c < nz ? Vzi[(size_t)r * nz + c] : 0.0);
  for (int r = 0; r < nzp; ++r)
    for (int c = 0; c < nzp; ++c) m.push_back(r < nz && c < nz ? Vz[(size_t)r * nz + c] : 0.0);
  // 1 / (lx_e + ly_g + lz_k) per eigenmode, laid out [k][g][e]; 0 for the
  // dropped null mode
  out->inv.assign((size_t)nx * ny * nz, 0.0);
  for (int k = 0; k < nz; ++k)
    for (int g = 0; g < ny; ++g)
      for (int e = 0; e < nx; ++e) {
        const bool null = singular && e == sx0 && g == sy0 && k == sz0;
        out->inv[((size_t)k * ny + g) * nx + e] = null ? 0.0 : 1.0 / ((lx[e] + ly[g]) + lz[k]);
      }
  out->nzp = nzp;
  return true;
}

}  // namespace amg
