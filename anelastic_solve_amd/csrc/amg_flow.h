// Time-stepper boundary handling (ghost) and the implicit vertical viscous
// step (var_advance_CN_PL): host planning and device launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "amg.h"

namespace amg {

// Ghost-inclusive 1-D metrics of the finest grid (length n+2 per direction,
// direction 0/1/2 = x/y/z) and rho(z).
struct FlowGrid {
  int nx = 0, ny = 0, nz = 0;
  std::vector<double> dxc[3], dxe[3];
  std::vector<double> rho;
};

// One face's bound_update rule with its products value*metric precomputed.
struct FaceRule {
  int btype;
  int rt;
  double value;
  double prod;      // value * metric at the ghost-side index
  double prod_rt;   // value * metric one further in (rtype == 1 rules)
};

// ghost(var, vtype): the periodic wrap of every ghost, then per direction
// nrules face rules (0 when the direction has nothing to apply).
struct GhostPlan {
  int bound;              // bound_update knows the type (else wraps only)
  FaceRule r[3][2];
  int nrules[3];
};

// One row k of the vertical CN system (every column alike): the eliminated
// sub-diagonal ap and pivot bp, the super-diagonal c, a constant right-hand
// side (kind 1) or the physics row's viscous factors c1, m1, c2, m2, m3.
struct CnRow {
  double ap, bp, c;
  double rbp, cbp;  // 1/bp and -c/bp (the segmented solve's products)
  double cval;
  double c1, m1, c2, m2, m3;
  double pz;        // 1/dze(k): the predicted-pressure metric of w
  int kind;
};

int plan_ghost(const FlowGrid& G, const char* vtype, const amg_flowbc& bc, GhostPlan* out,
               std::string* err);
void launch_flow_ghost(double* a, int nx, int ny, int nz, const GhostPlan& P, hipStream_t s);

int plan_cn(const FlowGrid& G, int var, double rkdt, double rRe, double rPr, double rSc,
            const amg_flowbc& bc, std::vector<CnRow>* rows, std::string* err);
void launch_cn_column(const double* v1, double* v2, const double* p, const CnRow* T,
                      const double* rdxe, const double* rdye, int nx, int ny, int nz, double f1,
                      double f2, double f3, double rkdt, int var, hipStream_t s);

}  // namespace amg
