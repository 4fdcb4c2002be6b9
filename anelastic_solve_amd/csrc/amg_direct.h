// amg_direct.h -- host set-up of the separable direct solve of a coarse level.
#pragma once
#include <string>
#include <vector>

#include "amg_hierarchy.h"

namespace amg {

struct DirectTables {
  std::vector<double> mats;   // [Vx^-T | Vx^T | Vy^-1 | Vy | Vz^-1 | Vz] (row-major; z padded)
  std::vector<double> inv;    // 1 / (lx + ly + lz) per eigenmode [k][g][e], 0 for the null mode
  int nzp = 0;                // nz rounded up to a multiple of 16
};

// Eigen-decomposition of the level's 1-D operators in x, y and z
// (symmetrised with the cell widths, and rho in z); false with a message
// when the level cannot take the direct solve (periodic z, a table that is
// not symmetrisable, no convergence).
bool direct_setup(const HostLevel& L, const int bc[3][2], bool singular, DirectTables* out,
                  std::string* err);

}  // namespace amg
