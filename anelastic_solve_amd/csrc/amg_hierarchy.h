// amg_hierarchy.h -- host-side level hierarchy: metrics, 1-D coefficient
// tables, coarse density, boundary folds and the coarsest-level inverse.
#pragma once
#include <string>
#include <vector>

#include "amg.h"

namespace amg {

// One direction's metrics, 0-based and ghost-inclusive (length n+2), with
// the reference's definitions (include/modules.F90:17-39): xe = upper face,
// xc = centre, dxc = cell width, dxe = centre distance.
struct Axis {
  int n = 0;
  std::vector<double> xe, xc, dxe, dxc;
};

Axis axis_from_faces(const std::vector<double>& faces);
Axis axis_from_arrays(int n, const double* xe, const double* xc, const double* dxe,
                      const double* dxc);
Axis coarsen_axis(const Axis& a, int factor, bool wall_lo, bool wall_hi);

struct HostLevel {
  int nx = 0, ny = 0, nz = 0;
  int fx = 1, fy = 1, fz = 1;          // factors to the next coarser level
  Axis ax, ay, az;
  std::vector<double> rho;             // rho(z) with ghosts, nz+2
  std::vector<double> co;              // [wxm wxp dx | wym wyp dy | wzm wzp dz]
  std::vector<double> amx, apx, amy, apy, amz, apz;   // raw face coefficients
  std::vector<double> vol;             // [dxc | dyc | dzc] interior
  std::vector<double> bcfold;          // finest only: [x | y | z] RHS shifts
  long long ncell() const { return (long long)nx * ny * nz; }
};

struct Hierarchy {
  std::vector<HostLevel> lv;           // finest first
  bool singular = false;
  int bc[3][2] = {{0, 0}, {0, 0}, {0, 0}};   // per axis (lo, hi)
  std::vector<double> minvT;           // coarsest dense inverse, transposed, chequer order
  double vol_sum = 0.0;                // sum of fine cell volumes
  double vol_rho_sum = 0.0;            // sum of fine V*rho
};

// Builds the hierarchy; returns AMG_OK or an error code with a message.
int build_hierarchy(const amg_config& cfg, const amg_grid& g, Hierarchy& h, std::string& err);

// Chequer index of (i,j,k) on a level with nx (even) x ny x nz cells.
inline long long chequer_index(int i, int j, int k, int nx, int ny) {
  const int hx = nx / 2;
  return ((long long)(k * ny + j) * 2 + ((i + j) & 1)) * hx + (i >> 1);
}

// Per-level sweep counts (finest first) from the reference-order arrays
// (coarsest first), or a uniform count when the arrays do not match.
void sweep_counts(const amg_config& cfg, int nlev, std::vector<int>& pre, std::vector<int>& post);

}  // namespace amg
