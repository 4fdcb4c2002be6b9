// Time-stepper boundary handling and the implicit vertical viscous step:
//   ghost(var, vtype)            dns/BC.F90:1-187 (+ bound_update :334-584), serial
//   var_advance_CN_PL(...)       dns/NSsolver.F90:3054-3465, DNS branch: the Thomas
//                                solve of mgMK/mg_relax.F90:541-595 on every interior
//                                column, in segments (nz <= 256) or serially
// Fields are the reference's ghosted FP64 arrays, x fastest:
//   e(i,j,k) = (k*(ny+2) + j)*(nx+2) + i,   0 <= i <= nx+1 (ghosts 0 and nx+1).
// Floating-point contraction is off in this file: every product and sum is
// rounded as written, in the order of the Fortran source, so the results
// equal a plain IEEE evaluation of the reference's expressions (the
// segmented line solve: to rounding).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "amg_flow.h"

#pragma clang fp contract(off)

namespace amg {
namespace {

constexpr int kFlowBlock = 256;

__device__ __forceinline__ size_t fidx(int i, int j, int k, int nx, int ny) {
  return ((size_t)k * (ny + 2) + j) * (size_t)(nx + 2) + i;
}

// Every ghost first takes its periodic image (dns/BC.F90:51-61: x, then y,
// then z over the full extents).  Composed, the three wraps map a ghost cell
// to the interior cell (w(i), w(j), w(k)) with w(0) = n, w(n+1) = 1, so each
// ghost is one copy from an interior cell and the copies are independent.
// Ghost cells are enumerated as the two z planes (full), the two y planes
// (interior k) and the two x planes (interior j, k).
__global__ __launch_bounds__(kFlowBlock) void k_flow_wrap(double* __restrict__ a, int nx, int ny,
                                                          int nz) {
  const long long t = (long long)blockIdx.x * kFlowBlock + threadIdx.x;
  const long long Z = 2LL * (nx + 2) * (ny + 2), Y = 2LL * (nx + 2) * nz, X = 2LL * ny * nz;
  int i, j, k;
  if (t < Z) {
    const long long pl = (nx + 2) * (long long)(ny + 2);
    k = t < pl ? 0 : nz + 1;
    const long long r = t % pl;
    j = (int)(r / (nx + 2));
    i = (int)(r % (nx + 2));
  } else if (t < Z + Y) {
    const long long u = t - Z, pl = (nx + 2) * (long long)nz;
    j = u < pl ? 0 : ny + 1;
    const long long r = u % pl;
    k = 1 + (int)(r / (nx + 2));
    i = (int)(r % (nx + 2));
  } else if (t < Z + Y + X) {
    const long long u = t - Z - Y, pl = (long long)ny * nz;
    i = u < pl ? 0 : nx + 1;
    const long long r = u % pl;
    k = 1 + (int)(r / ny);
    j = 1 + (int)(r % ny);
  } else {
    return;
  }
  const int si = i == 0 ? nx : (i == nx + 1 ? 1 : i);
  const int sj = j == 0 ? ny : (j == ny + 1 ? 1 : j);
  const int sk = k == 0 ? nz : (k == nz + 1 ? 1 : k);
  a[fidx(i, j, k, nx, ny)] = a[fidx(si, sj, sk, nx, ny)];
}

// One direction's face rules (bound_update) on every line of that direction,
// over the full extents of the other two (ghosts included); face 1 then face 2
// on the same line, as the reference calls them.  Products value*metric are
// precomputed on the host (FaceRule::prod, prod_rt).
__device__ __forceinline__ void face_rule(double* __restrict__ a, size_t base, long long st, int n,
                                          const FaceRule& R, int face, int d) {
  const int s = 1, e = n;
  auto at = [&](int q) -> double& { return a[base + (size_t)((long long)q * st)]; };
  switch (R.btype) {
    case 1:
      if (face == 1) at(s - 1) = R.value;
      else at(e + 1) = R.value;
      break;
    case 2:
      if (face == 1) {
        if (R.rt) at(s) = at(s + 1) - R.prod_rt;
        at(s - 1) = at(s) - R.prod;
      } else {
        if (R.rt) at(e) = at(e - 1) + R.prod_rt;
        at(e + 1) = at(e) + R.prod;
      }
      break;
    case 3:
      at(s - 1) = at(e);
      at(e + 1) = at(s);
      break;
    case 4:
      if (face == 1) at(s - 1) = at(s) - R.prod;
      else at(e + 1) = at(e) + R.prod;
      break;
    case 6:
      if (d == 2) {
        if (face == 1) at(s - 1) = 2.0 * R.value - 1.0 * at(s);
        else at(e + 1) = 2.0 * R.value - 1.0 * at(e);
      }
      break;
    case 7:
      if (d == 2) {
        if (face == 1) {
          at(s) = 0.0;
          at(s - 1) = -1.0 * at(s + 1);
        } else {
          at(e) = 0.0;
          at(e + 1) = -1.0 * at(e - 1);
        }
      }
      break;
    case 8:
      if (d == 0) {
        if (face == 1) at(s - 1) = 2.0 * at(s) - at(s + 1);
        else at(e + 1) = 2.0 * at(e) - at(e - 1);
      }
      break;
    default:
      break;
  }
}

__global__ __launch_bounds__(kFlowBlock) void k_flow_face(double* __restrict__ a, int nx, int ny,
                                                          int nz, int d, FaceRule r1, FaceRule r2,
                                                          int nrules) {
  const long long t = (long long)blockIdx.x * kFlowBlock + threadIdx.x;
  const int ex = nx + 2, ey = ny + 2;
  size_t base;
  long long st;
  int n;
  if (d == 0) {          // lines along x, one per (j, k)
    if (t >= (long long)ey * (nz + 2)) return;
    base = (size_t)t * ex;
    st = 1;
    n = nx;
  } else if (d == 1) {   // lines along y, one per (i, k)
    if (t >= (long long)ex * (nz + 2)) return;
    const long long k = t / ex, i = t % ex;
    base = (size_t)(k * ey) * ex + (size_t)i;
    st = ex;
    n = ny;
  } else {               // lines along z, one per (i, j)
    if (t >= (long long)ex * ey) return;
    base = (size_t)t;
    st = (long long)ex * ey;
    n = nz;
  }
  face_rule(a, base, st, n, r1, 1, d);
  if (nrules > 1) face_rule(a, base, st, n, r2, 2, d);
}

// var_advance_CN_PL on one interior column (i, j) per lane (nz > 256;
// k_cn_seg below takes nz <= 256): the right-hand side of every physics row
// from the old values (f1*v2 + f2*v1*f3, the predicted
// pressure term when p is given, the explicit half of the vertical viscous
// term), the constant rows (walls, and rows a ghost rule zeroed), then the
// mgMK Thomas with the column-independent eliminated rows ap/bp: forward
// G_k = r_k + ap_k G_{k-1}, written into v2 as it goes (the old v2 values
// the next rows need are held in registers), and backward
// x_k = (G_k - c_k x_{k+1}) / bp_k.  Table row k (CnRow) is wave-uniform.
// Both passes run in chunks of rows (16 or 32 forward, 32 backward) whose
// loads (which do not depend on the recurrence) are issued before the
// chunk's dependent steps, so a thread keeps a chunk of loads in flight
// instead of one row.

// PV selects the predicted-pressure term at compile time: 0 none, 1 the x
// neighbour (u), 2 the y neighbour (v), 3 the z neighbour (w), so the march
// has no run-time branch on the variable.
template <int PV>
__global__ __launch_bounds__(kFlowBlock) void k_cn_column(const double* __restrict__ v1,
                                                          double* __restrict__ v2,
                                                          const double* __restrict__ p,
                                                          const CnRow* __restrict__ T,
                                                          const double* __restrict__ rdxe,
                                                          const double* __restrict__ rdye,
                                                          int nx, int ny, int nz, double f1,
                                                          double f2, double f3, double rkdt,
                                                          double h) {
  const int t = blockIdx.x * kFlowBlock + threadIdx.x;
  if (t >= nx * ny) return;
  const int i = 1 + t % nx, j = 1 + t / nx;
  const size_t st = (size_t)(nx + 2) * (ny + 2);
  const size_t base = (size_t)j * (nx + 2) + i;
  const int n = nz + 2;
  // the pressure neighbour is in the plane for u and v, one plane up for w
  // (clamped at the top row, which is a constant row); its metric is
  // per thread for u and v, per row (the table) for w
  constexpr bool pcv = PV != 0;
  const size_t poff = PV == 1 ? 1 : (size_t)(nx + 2);
  auto pnb = [&](int k) -> size_t {
    return PV == 3 ? base + (size_t)min(k + 1, n - 1) * st : base + (size_t)k * st + poff;
  };
  const double dmc = PV == 1 ? rdxe[i] : (PV == 2 ? rdye[j] : 0.0);
  // old values of rows k-1 and k; row 0 is a constant row, so vm is first
  // read at k = 1 (row 0)
  double vm = 0.0, vc = v2[base];
  double G = 0.0;
  // Every row evaluates the physics expression and a constant row then
  // selects its value, so no load is used only under a branch (the compiler
  // would move such loads down to their use and undo the prefetch).
  auto step = [&](int k, double vp, double a1, double pa, double pb) {
    const CnRow R = T[k];
    double r = f1 * vc + f2 * a1 * f3;
    if constexpr (pcv) r = r - rkdt * (pb - pa) * (PV == 3 ? R.pz : dmc);
    const double tv = R.c1 * (vp - vc) * R.m1 - R.c2 * (vc - vm) * R.m2;
    r = r + h * tv * R.m3;
    r = R.kind ? R.cval : r;
    G = k == 0 ? r : r + R.ap * G;
    v2[base + (size_t)k * st] = G;
    vm = vc;
    vc = vp;
  };
  // Forward, then backward, in self-contained chunks: all loads of a chunk
  // are issued together, then its dependent steps run.  With one column
  // per thread a 256^2 grid gives one wave per SIMD, so the chunks are made
  // long (deep prefetch) rather than double-buffered: a buffer live across
  // the loop's back edge would make the compiler wait for every load there.
  constexpr int KF = PV ? 16 : 32;
  constexpr int KB = 32;
  const int ncf = n / KF;
  for (int c = 0; c < ncf; ++c) {
    double vp[KF], a1[KF], pa[KF], pb[KF];
#pragma unroll
    for (int q = 0; q < KF; ++q) {
      const int kk = c * KF + q;
      vp[q] = v2[base + (size_t)min(kk + 1, n - 1) * st];   // the top row's is unused
      a1[q] = v1[base + (size_t)kk * st];
      if constexpr (pcv) {
        pa[q] = p[base + (size_t)kk * st];
        pb[q] = p[pnb(kk)];
      } else {
        pa[q] = pb[q] = 0.0;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < KF; ++q) step(c * KF + q, vp[q], a1[q], pa[q], pb[q]);
  }
  for (int k = ncf * KF; k < n; ++k) {
    const double vp = v2[base + (size_t)min(k + 1, n - 1) * st];
    const double a1 = v1[base + (size_t)k * st];
    const double pa = pcv ? p[base + (size_t)k * st] : 0.0;
    const double pb = pcv ? p[pnb(k)] : 0.0;
    step(k, vp, a1, pa, pb);
  }
  // backward: the top row, the rows above the full chunks, then the chunks
  // downward (chunk b holds rows rb - q with rb = n - 2 - r0 - b*KB)
  double x = G / T[n - 1].bp;
  v2[base + (size_t)(n - 1) * st] = x;
  const int r0 = (n - 1) % KB;            // rows n-2 .. n-1-r0 go first, one by one
  for (int k = n - 2; k > n - 2 - r0; --k) {
    const size_t e = base + (size_t)k * st;
    x = (v2[e] - T[k].c * x) / T[k].bp;
    v2[e] = x;
  }
  const int ncb = (n - 1) / KB;
  for (int b = 0; b < ncb; ++b) {
    const int rb = n - 2 - r0 - b * KB;
    double g[KB];
#pragma unroll
    for (int q = 0; q < KB; ++q) g[q] = v2[base + (size_t)(rb - q) * st];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < KB; ++q) {
      const int kk = rb - q;
      x = (g[q] - T[kk].c * x) / T[kk].bp;
      v2[base + (size_t)kk * st] = x;
    }
  }
}

// The same step with every column split into P segments of S interior rows
// (at most kCnSeg rows and kCnSeg segments: nz <= 256), one wave per segment
// and 64 columns per workgroup.  The recurrences' coefficients are
// the same in every column, so their values are linear in the value carried
// into a segment:
//   forward   G_k = G'_k + (prod_{i=r0..k} ap_i) G_{r0-1},        G' from G_{r0-1} = 0,
//   backward  x_k = x'_k + (prod_{i=k..r1-1} (-c_i/bp_i)) x_{r1},  x' from x_{r1} = 0.
// Each wave solves its segment locally with the serial kernel's operations
// (the division by the pivot as a product with its reciprocal, from the
// table), the segments' end values are chained through LDS (a scan over at
// most 16 segments), and each row adds its carry term: a few roundings more
// per row than the serial solve, whose result this equals to rounding.
// The column is read once and written once (24 B/cell against the serial
// kernel's 40), and a 256^2 grid has 16 waves per column group instead of
// one per SIMD.  Rows 0 and nz+1 (the wall rows) are taken by the first and
// the last wave with the serial kernel's expressions.  Every load comes
// before the first barrier and every store after the last, so the column is
// rewritten in place.
constexpr int kCnSeg = 16;     // rows per wave and waves per column group, at most
constexpr int kCnChunk = 4;    // rows whose loads are in flight together
template <int PV>
__global__ __launch_bounds__(1024) void k_cn_seg(const double* __restrict__ v1,
                                                 double* __restrict__ v2,
                                                 const double* __restrict__ p,
                                                 const CnRow* __restrict__ T,
                                                 const double* __restrict__ rdxe,
                                                 const double* __restrict__ rdye, int nx, int ny,
                                                 int nz, int S, double f1, double f2, double f3,
                                                 double rkdt, double h) {
  __shared__ double gend[kCnSeg][64], xfirst[kCnSeg][64], g0[64], xtop[64];
  __shared__ double pend[kCnSeg], qfirst[kCnSeg];
  const int lane = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.y), P = blockDim.y;
  const int t = blockIdx.x * 64 + lane;
  const bool live = t < nx * ny;
  const int tc = live ? t : nx * ny - 1;              // a spare lane repeats the last column
  const int i = 1 + tc % nx, j = 1 + tc / nx;
  const size_t st = (size_t)(nx + 2) * (ny + 2);
  const size_t base = (size_t)j * (nx + 2) + i;
  const int n = nz + 2;
  constexpr bool pcv = PV != 0;
  const size_t poff = PV == 1 ? 1 : (size_t)(nx + 2);
  auto at = [&](int k) { return base + (size_t)k * st; };
  auto pnb = [&](int k) -> size_t {
    return PV == 3 ? at(min(k + 1, n - 1)) : at(k) + poff;
  };
  const double dmc = PV == 1 ? rdxe[i] : (PV == 2 ? rdye[j] : 0.0);
  // the serial kernel's right-hand side of row k
  auto rhs = [&](int k, double vm, double vc, double vp, double a1, double pa, double pb) {
    const CnRow R = T[k];
    double r = f1 * vc + f2 * a1 * f3;
    if constexpr (pcv) r = r - rkdt * (pb - pa) * (PV == 3 ? R.pz : dmc);
    const double tv = R.c1 * (vp - vc) * R.m1 - R.c2 * (vc - vm) * R.m2;
    r = r + h * tv * R.m3;
    return R.kind ? R.cval : r;
  };
  const int r0 = 1 + w * S, r1 = min(r0 + S, nz + 1);
  // the wall rows' inputs, with the segment's (all before the first barrier)
  double G0 = 0.0, rtop = 0.0;
  if (w == 0)
    G0 = rhs(0, 0.0, v2[at(0)], v2[at(1)], v1[at(0)], pcv ? p[at(0)] : 0.0,
             pcv ? p[pnb(0)] : 0.0);
  if (w == P - 1)
    rtop = rhs(n - 1, v2[at(n - 2)], v2[at(n - 1)], v2[at(n - 1)], v1[at(n - 1)],
               pcv ? p[at(n - 1)] : 0.0, pcv ? p[pnb(n - 1)] : 0.0);
  // forward, locally: the loads of kCnChunk rows, then their steps
  double G[kCnSeg];
  double g = 0.0, pr = 1.0;
#pragma unroll
  for (int c0 = 0; c0 < kCnSeg; c0 += kCnChunk) {
    if (c0 < S) {
      double wv[kCnChunk + 2], a1[kCnChunk], pa[kCnChunk], pb[kCnChunk];
#pragma unroll
      for (int q = 0; q < kCnChunk + 2; ++q) wv[q] = v2[at(min(r0 - 1 + c0 + q, n - 1))];
#pragma unroll
      for (int q = 0; q < kCnChunk; ++q) {
        const int k = min(r0 + c0 + q, n - 1);
        a1[q] = v1[at(k)];
        pa[q] = pcv ? p[at(k)] : 0.0;
        pb[q] = pcv ? p[pnb(k)] : 0.0;
      }
      // the right-hand sides (independent of the recurrence), then its steps
      double rr[kCnChunk], apq[kCnChunk];
#pragma unroll
      for (int q = 0; q < kCnChunk; ++q) {
        const int k = min(r0 + c0 + q, n - 1);
        rr[q] = rhs(k, wv[q], wv[q + 1], wv[q + 2], a1[q], pa[q], pb[q]);
        apq[q] = T[k].ap;
      }
#pragma unroll
      for (int q = 0; q < kCnChunk; ++q) {
        if (r0 + c0 + q < r1) {
          g = rr[q] + apq[q] * g;
          pr = apq[q] * pr;
        }
        G[c0 + q] = g;
      }
    }
  }
  gend[w][lane] = g;
  if (w == 0) g0[lane] = G0;
  if (lane == 0) pend[w] = pr;
  __syncthreads();
  // G_{r0-1}, the value carried into this segment, and the segment's G
  double C = g0[lane];
  for (int q = 0; q < w; ++q) C = gend[q][lane] + pend[q] * C;
  pr = 1.0;
#pragma unroll
  for (int q = 0; q < kCnSeg; ++q) {
    const int k = min(r0 + q, n - 1);
    const double ap = T[k].ap;
    if (r0 + q < r1) {
      pr = ap * pr;
      G[q] = G[q] + pr * C;
    }
  }
  // the top wall row: G_{nz+1} from G_nz, x_{nz+1}
  if (w == P - 1) {
    double Gn = 0.0;
#pragma unroll
    for (int q = 0; q < kCnSeg; ++q)
      if (r0 + q == r1 - 1) Gn = G[q];
    xtop[lane] = (rtop + T[n - 1].ap * Gn) * T[n - 1].rbp;
  }
  // backward, locally (the rows' coefficients read before the chain)
  double cq[kCnSeg], rq[kCnSeg], bq[kCnSeg];
#pragma unroll
  for (int q = 0; q < kCnSeg; ++q) {
    const int k = min(r0 + q, n - 1);
    cq[q] = T[k].c;
    rq[q] = T[k].rbp;
    bq[q] = T[k].cbp;
  }
  double x = 0.0, qr = 1.0;
#pragma unroll
  for (int q = kCnSeg - 1; q >= 0; --q) {
    if (r0 + q < r1) {
      x = (G[q] - cq[q] * x) * rq[q];
      qr = bq[q] * qr;
      G[q] = x;
    }
  }
  xfirst[w][lane] = x;
  if (lane == 0) qfirst[w] = qr;
  __syncthreads();
  // x_{r1}, the value carried into this segment from above, and the stores
  double X = xtop[lane];
  for (int q = P - 1; q > w; --q) X = xfirst[q][lane] + qfirst[q] * X;
  double qk = 1.0, x1 = 0.0;
#pragma unroll
  for (int q = kCnSeg - 1; q >= 0; --q) {
    const int k = r0 + q;
    if (k < r1) {
      qk = bq[q] * qk;
      const double xk = G[q] + qk * X;
      if (live) v2[at(k)] = xk;
      x1 = xk;
    }
  }
  if (live && w == 0) v2[at(0)] = (G0 - T[0].c * x1) * T[0].rbp;
  if (live && w == P - 1) v2[at(n - 1)] = xtop[lane];
}

unsigned cdiv(long long a, int b) { return (unsigned)((a + b - 1) / b); }

// ghost()'s variable classes (dns/BC.F90:66-95) and bound_update's metrics
// (:384-447): l, rt, and per direction 0 = centre distance dxe, 1 = width dxc.
struct VClass {
  const char* name;
  int l, rt, bound;       // bound: bound_update knows the type
  int mx, my, mz, of1, of2;
};
const VClass kClasses[] = {
    {"u", 1, 0, 1, 1, 0, 0, 0, 1},       {"uc", 1, 0, 1, 0, 0, 0, -1, 0},
    {"v", 2, 0, 1, 0, 1, 0, 0, 1},       {"vc", 2, 0, 1, 0, 0, 0, -1, 0},
    {"w", 3, 0, 1, 0, 0, 1, 0, 1},       {"wc", 3, 0, 1, 0, 0, 0, -1, 0},
    {"p", 4, 0, 1, 0, 0, 0, -1, 0},      {"r", 5, 1, 0, 0, 0, 0, 0, 0},
    {"rho", 5, 1, 1, 0, 0, 0, -1, 0},    {"rf", 5, 1, 1, 0, 0, 0, -1, 0},
    {"cfluc", 6, 0, 1, 0, 0, 0, -1, 0},  {"fluc", 6, 0, 1, 0, 0, 0, -1, 0},
    {"rp", 6, 1, 1, 0, 0, 0, -1, 0},     {"psource", 7, 0, 1, 0, 0, 0, -1, 0},
    {"utemp", 7, 0, 1, 1, 0, 0, 0, 1},   {"vtemp", 7, 0, 1, 0, 1, 0, 0, 1},
    {"wtemp", 7, 0, 1, 0, 0, 1, 0, 1},   {"rtemp", 7, 0, 1, 0, 0, 0, -1, 0},
    {"s1temp", 7, 0, 1, 0, 0, 0, -1, 0}, {"scal1", 8, 1, 1, 0, 0, 0, -1, 0},
    {"scal1temp", 8, 0, 1, 0, 0, 0, -1, 0},
};

}  // namespace

int plan_ghost(const FlowGrid& G, const char* vtype, const amg_flowbc& bc, GhostPlan* out,
               std::string* err) {
  const VClass* c = nullptr;
  for (const VClass& v : kClasses)
    if (vtype && std::strcmp(v.name, vtype) == 0) c = &v;
  if (!c) {
    *err = std::string("BC TYPE NOT AVAILABLE: ") + (vtype ? vtype : "(null)");
    return -1;
  }
  GhostPlan P{};
  P.bound = c->bound;
  const int n[3] = {G.nx, G.ny, G.nz};
  const int mk[3] = {c->mx, c->my, c->mz};
  for (int d = 0; d < 3; ++d) {
    const std::vector<double>& dt = mk[d] ? G.dxc[d] : G.dxe[d];
    const int l = c->l;
    auto TB = [&](int f) { return bc.tb[d + 3 * (f - 1) + 6 * (l - 1)]; };
    auto VB = [&](int f) { return bc.vb[d + 3 * (f - 1) + 6 * (l - 1)]; };
    const int s = 1, e = n[d];
    int nr = 0;
    for (int m = 1; m <= 2; ++m) {
      // face 1's type decides for both faces (dns/BC.F90:100-176)
      if (!(TB(1) != 3 || m == 1)) continue;
      FaceRule R{};
      R.btype = TB(m);
      R.rt = c->rt;
      R.value = VB(m);
      const int i1 = m == 1 ? s + c->of1 : e + c->of2;
      const int i2 = m == 1 ? s + 1 + c->of1 : e - 1 + c->of2;
      if (R.btype == 2 || R.btype == 4) {
        if (i1 < 0 || i1 > n[d] + 1 || i2 < 0 || i2 > n[d] + 1) {
          *err = "ghost: metric index out of range";
          return -1;
        }
        R.prod = R.value * dt[i1];
        R.prod_rt = R.value * dt[i2];
      }
      (m == 1 ? P.r[d][0] : P.r[d][1]) = R;
      nr = m;
    }
    P.nrules[d] = nr;
  }
  *out = P;
  return 0;
}

void launch_flow_ghost(double* a, int nx, int ny, int nz, const GhostPlan& P, hipStream_t s) {
  const long long ng = 2LL * (nx + 2) * (ny + 2) + 2LL * (nx + 2) * nz + 2LL * ny * nz;
  hipLaunchKernelGGL(k_flow_wrap, dim3(cdiv(ng, kFlowBlock)), dim3(kFlowBlock), 0, s, a, nx, ny, nz);
  if (!P.bound) return;
  const long long lines[3] = {(long long)(ny + 2) * (nz + 2), (long long)(nx + 2) * (nz + 2),
                              (long long)(nx + 2) * (ny + 2)};
  for (int d = 0; d < 3; ++d) {
    if (P.nrules[d] == 0) continue;
    hipLaunchKernelGGL(k_flow_face, dim3(cdiv(lines[d], kFlowBlock)), dim3(kFlowBlock), 0, s, a, nx,
                       ny, nz, d, P.r[d][0], P.r[d][1], P.nrules[d]);
  }
}

int plan_cn(const FlowGrid& G, int var, double rkdt, double rRe, double rPr, double rSc,
            const amg_flowbc& bc, std::vector<CnRow>* rows, std::string* err) {
  const int nz = G.nz;
  const std::vector<double>& r = G.rho;
  const std::vector<double>& dzc = G.dxc[2];
  const std::vector<double>& dze = G.dxe[2];
  if (var != 1 && var != 2 && var != 3 && var != 5 && var != 8) {
    *err = "ABORTING ADVANCE CN, VARID IS NOT VALID";
    return -1;
  }
  std::vector<double> rdzc(nz + 2), rdze(nz + 2);
  for (int k = 0; k < nz + 2; ++k) {
    rdzc[k] = 1.0 / dzc[k];
    rdze[k] = 1.0 / dze[k];
  }
  // rows of the system, every column alike (dns/NSsolver.F90:3136-3216)
  std::vector<double> sub(nz + 2, 0.0), diag(nz + 2, 0.0), sup(nz + 2, 0.0);
  std::vector<CnRow> R(nz + 2);
  for (int k = 1; k <= nz; ++k) {
    double tmp1, tmp2;
    if (var == 1 || var == 2 || var == 5) {
      const double raver1 = 0.5 * (r[k + 1] + r[k]);
      const double raver2 = 0.5 * (r[k] + r[k - 1]);
      const double c = var == 5 ? rRe * rPr : rRe;
      tmp1 = c * raver1 / r[k];
      tmp2 = c * raver2 / r[k];
    } else if (var == 8) {
      tmp1 = tmp2 = rRe * rSc;
    } else {
      tmp1 = rRe * r[k] / (0.5 * (r[k + 1] + r[k]));
      tmp2 = rRe * r[k + 1] / (0.5 * (r[k + 1] + r[k]));
    }
    CnRow& w = R[k];
    w.c1 = tmp2;
    w.c2 = tmp1;
    w.pz = rdze[k];
    if (var == 3) {
      sub[k] = -0.5 * rkdt * tmp1 * rdze[k] * rdzc[k];
      sup[k] = -0.5 * rkdt * tmp2 * rdze[k] * rdzc[k + 1];
      diag[k] = 1.0 + 0.5 * rkdt * (tmp1 * rdzc[k] + tmp2 * rdzc[k + 1]) * rdze[k];
      w.m1 = rdzc[k + 1];
      w.m2 = rdzc[k];
      w.m3 = rdze[k];
    } else {
      sub[k] = -0.5 * rkdt * tmp1 * rdze[k - 1] * rdzc[k];
      sup[k] = -0.5 * rkdt * tmp2 * rdze[k] * rdzc[k];
      diag[k] = 1.0 + 0.5 * rkdt * (tmp1 * rdze[k - 1] + tmp2 * rdze[k]) * rdzc[k];
      w.m1 = rdze[k];
      w.m2 = rdze[k - 1];
      w.m3 = rdzc[k];
    }
  }
  // ghost(sub/diag/sup/rhs, 'cfluc') (:3331-3335): on interior columns only
  // the z rules reach rows that survive the wall rows below: type 7 zeroes
  // row sz (ez) of all four arrays.  Rows 0 and nz+1 are always rewritten
  // by the wall rows (every accepted type sets them).
  auto TB = [&](int f, int v) { return bc.tb[2 + 3 * (f - 1) + 6 * (v - 1)]; };
  auto VB = [&](int f, int v) { return bc.vb[2 + 3 * (f - 1) + 6 * (v - 1)]; };
  R[0].kind = R[nz + 1].kind = 1;
  for (int m = 1; m <= 2; ++m) {
    if (!(TB(1, 6) != 3 || m == 1)) continue;
    if (TB(m, 6) == 7) {
      const int k = m == 1 ? 1 : nz;
      sub[k] = diag[k] = sup[k] = 0.0;
      R[k].kind = 1;
      R[k].cval = 0.0;
    }
  }
  // the wall rows (:3337-3428)
  auto row = [&](int k, double a, double b, double c, double v) {
    sub[k] = a;
    diag[k] = b;
    sup[k] = c;
    R[k].kind = 1;
    R[k].cval = v;
  };
  const int s = 1, e = nz;
  int tb = TB(1, var);
  double vb = VB(1, var);
  if (tb == 1) {
    row(s - 1, 0.0, 1.0, 0.0, vb);
  } else if (tb == 2) {
    if (var == 3) {
      row(s - 1, 0.0, -1.0, 1.0, dzc[s - 1] * vb);
    } else if (var == 5 || var == 8) {
      row(s - 1, 0.0, -1.0, 1.0, dze[s - 1] * vb);
      row(s, 0.0, -1.0, 1.0, dze[s] * vb);
    } else {
      row(s - 1, 0.0, -1.0, 1.0, dze[s - 1] * vb);
    }
  } else if (tb == 6) {
    row(s - 1, 0.0, 1.0, 1.0, 2.0 * vb);
  } else {
    *err = "Incorrect BC for bottom for variable " + std::to_string(var) + " in var_advance_CN";
    return -1;
  }
  tb = TB(2, var);
  vb = VB(2, var);
  if (tb == 1) {
    row(e + 1, 0.0, 1.0, 0.0, vb);
  } else if (tb == 2) {
    if (var == 3) {
      row(e + 1, -1.0, 1.0, 0.0, dzc[e + 1] * vb);
    } else if (var == 5 || var == 8) {
      row(e + 1, -1.0, 1.0, 0.0, dze[e + 1] * vb);
      row(e, -1.0, 1.0, 0.0, dze[e] * vb);
    } else {
      row(e + 1, -1.0, 1.0, 0.0, dze[e + 1] * vb);
    }
  } else if (tb == 6) {
    row(e + 1, 1.0, 1.0, 0.0, 2.0 * vb);
  } else if (tb == 7) {
    row(e + 1, 0.0, 1.0, 0.0, 0.0);
    row(e, 0.0, 1.0, 0.0, 0.0);
  } else {
    *err = "Incorrect BC for top for variable " + std::to_string(var) + " in var_advance_CN";
    return -1;
  }
  // the matrix half of Thomas (mgMK/mg_relax.F90:577-581), the same for
  // every column
  std::vector<double> a = sub, b = diag;
  for (int j = 0; j < nz + 1; ++j) {
    if (b[j] == 0.0) {
      *err = "var_advance_CN: zero pivot in the vertical system";
      return -1;
    }
    a[j + 1] = -a[j + 1] / b[j];
    b[j + 1] = b[j + 1] + a[j + 1] * sup[j];
  }
  if (b[nz + 1] == 0.0) {
    *err = "var_advance_CN: zero pivot in the vertical system";
    return -1;
  }
  for (int k = 0; k < nz + 2; ++k) {
    R[k].ap = a[k];
    R[k].bp = b[k];
    R[k].c = sup[k];
    R[k].rbp = 1.0 / b[k];
    R[k].cbp = -(sup[k] / b[k]);
  }
  *rows = std::move(R);
  return 0;
}

void launch_cn_column(const double* v1, double* v2, const double* p, const CnRow* T,
                      const double* rdxe, const double* rdye, int nx, int ny, int nz, double f1,
                      double f2, double f3, double rkdt, int var, hipStream_t s) {
  const int pv = (p != nullptr && var <= 3) ? var : 0;
  const double h = rkdt * 0.5;
  if (nz <= kCnSeg * kCnSeg) {
    // segments of S <= 16 rows, at most 16 of them
    const int S = (nz + kCnSeg - 1) / kCnSeg, P = (nz + S - 1) / S;
    const dim3 g(cdiv((long long)nx * ny, 64)), b(64, P);
#define CS(V) hipLaunchKernelGGL(k_cn_seg<V>, g, b, 0, s, v1, v2, p, T, rdxe, rdye, nx, ny, nz, S, f1, f2, f3, rkdt, h)
    if (pv == 1) CS(1);
    else if (pv == 2) CS(2);
    else if (pv == 3) CS(3);
    else CS(0);
#undef CS
    return;
  }
  const dim3 g(cdiv((long long)nx * ny, kFlowBlock)), b(kFlowBlock);
  if (pv == 1) hipLaunchKernelGGL(k_cn_column<1>, g, b, 0, s, v1, v2, p, T, rdxe, rdye, nx, ny, nz, f1, f2, f3, rkdt, h);
  else if (pv == 2) hipLaunchKernelGGL(k_cn_column<2>, g, b, 0, s, v1, v2, p, T, rdxe, rdye, nx, ny, nz, f1, f2, f3, rkdt, h);
  else if (pv == 3) hipLaunchKernelGGL(k_cn_column<3>, g, b, 0, s, v1, v2, p, T, rdxe, rdye, nx, ny, nz, f1, f2, f3, rkdt, h);
  else hipLaunchKernelGGL(k_cn_column<0>, g, b, 0, s, v1, v2, p, T, rdxe, rdye, nx, ny, nz, f1, f2, f3, rkdt, h);
}

}  // namespace amg
