// amg_handle.h -- the single-solver handle behind amg_handle (include/amg.h)
// and the helpers that the entry points of several source files share.
#pragma once
#include "amg.h"
#include "amg_flow.h"
#include "amg_solver.h"

struct amg_solver_s {
  amg::Context ctx;
  amg::Instance* I = nullptr;
  double* d_phi_stage = nullptr;   // staging for the host entry point
  double* d_rhs_stage = nullptr;
  double* d_div_partial = nullptr; // divergence partials (amg_divergence_device)
  amg::CnRow* d_cn = nullptr;      // rows of the vertical CN system (nz+2)
  amg::CnRow* h_cn = nullptr;      //   and their pinned host staging
};

namespace amg {

// Device pointers of the finest grid's ghost-inclusive projection metrics:
// rho(z) and the reciprocal cell widths and centre distances.
void metric_ptrs(const Instance* I, const double** rho, const double** rdxc, const double** rdyc,
                 const double** rdzc, const double** rdxe, const double** rdye,
                 const double** rdze);

// The divergence report of dns/Routines.F90:123-173 with its block
// partials and results in `partial` (4 * divergence_blocks + 4 doubles on
// the handle's device); synchronous.
int divergence_report(amg_handle h, const double* U, const double* V, const double* W,
                      double* div, double* partial, double* avg, double* maxabs, int loc[3]);

}  // namespace amg
