// amg_internal.h -- shared device descriptors and kernel launchers.
//
// Field layout in HBM ("chequer" layout).  A level with nx*ny*nz interior
// cells (nx even) stores each x-row (j,k) as two contiguous half-rows: the
// cells of column colour c = (i+j)&1, ordered by m = i>>1:
//
//     index(i,j,k) = ((k*ny + j)*2 + ((i+j)&1))*hx + (i>>1),   hx = nx/2.
//
// Both red-black smoothers then stream contiguous memory: a point-RB colour
// in plane k is one column colour ((i+j+k)&1 == p  <=>  (i+j)&1 == (p+k)&1),
// a z-line colour is one column colour in every plane, and all six
// neighbours of a cell are at the same m (y, z) or at m-1+r, m+r (x, with
// r = i&1) of a half-row of known colour.  No ghost cells are stored:
// periodic directions wrap the index, walls carry zero coupling in the 1-D
// coefficient tables.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace amg {

// Coefficient table of one level (precision T): 3*nx + 3*ny + 3*nz values,
//   [wxm | wxp | dx] [wym | wyp | dy] [wzm | wzp | dz]
// with (A u)_ijk = wxm_i u_{i-1} + wxp_i u_{i+1} + ... + (dx_i+dy_j+dz_k) u_ijk.
struct DevLevel {
  int nx, ny, nz, hx;
  int px, py, pz;            // periodic flags
  int inst;                  // instance id (batched/served runs)
  int ncell;                 // nx*ny*nz
  int link;                  // FP64 finest entry: index of the FP32 finest level (mixed)
  int nullsp;                // the operator is singular (periodic/Neumann): constants in the null space
  // the instance's solve state, written by the device (the gate at the end
  // of every cycle) on every level of the instance:
  int stop;                  // the stop rule held: the instance's kernels return at once
  int warm;                  // this cycle continues the previous one's correction (the second
                             //   of a pair): the finest level's first sweep reads the iterate
  int skipc;                 // this cycle is the first of a pair: no FP64 correction after it
  int pcur;                  // FP64 pressure buffers of a mixed solve: 1 when u2 holds the current
  int pairok;                // the instance may pair cycles (mixed, residual stop rule)
  int gslot;                 // the instance's slot in the per-instance gate array
  unsigned mhx, mny;         // ceil(2^32/hx), ceil(2^32/ny) (0 for a divisor of 1): e/hx = umulhi(e, mhx)
                             // exactly while e*hx < 2^32 (the coarse tail's small levels)
  unsigned gmhx, gmny;       // division of any 32-bit index by hx, ny: multiplier and
  int gshx, gsny;            // shifts (s1 | s2 << 8) of gm_magic
  unsigned gmzp;             // the same for nz rounded up to a multiple of 16 (the direct
  int gszp;                  // solve's padded z extent)
  int bcx0, bcx1, bcy0, bcy1, bcz0, bcz1;   // boundary codes (prolongation ghosts)
  double vsum;               // finest: sum of cell volumes
  double vrsum;              // finest: sum of V*rho
  double bshift;             // mixed FP64 finest: the compatibility shift sum V b / sum V, which
                             // the outer step subtracts from b as it loads it (a state word:
                             // written by the first defect)
  const double* co64;        // coefficient table, FP64
  const float* co32;         // coefficient table, FP32
  const double* vol;         // [dxc(nx) | dyc(ny) | dzc(nz)] interior widths
  const double* rho;         // rho(z) interior, nz values
  const double* bcfold;      // [x(nx) | y(ny) | z(nz)] boundary-value RHS shifts (finest only)
  const double* minvT;       // coarsest level: transposed dense (pseudo-)inverse, ncell^2
  const double* minv;        // the same, row-major, when the coarse solve is its own kernel
  void* u;                   // solution / correction (chequer layout)
  void* u2;                  // FP64 finest: ping-pong partner of u
  void* f;                   // right-hand side (chequer layout)
  void* scr;                 // z-line scratch: nz * hx * ny values
  // separable direct solve of a coarse level (amg_direct.cpp, k_direct_*)
  // the direct level's tables, in its compute type (FP64 for an FP64
  // V-cycle, FP32 -- stored behind these pointers -- for a mixed solve's):
  const double* dmat;        // [Vx^-T | Vx^T | Vy^-1 | Vy | Vz^-1 | Vz] (z padded to 16)
  const double* dinv;        // 1 / (lx + ly + lz) per eigenmode, [k][g][e]
  double* dscr;              // transformed right-hand side, then solution: 2 * nz*ny*nx (mode fastest)
};

// One kernel launch over several items (levels of different instances),
// passed by value in the kernel arguments (no device upload, so a served
// batch can change every round).  Block b belongs to the item t with
// prefix[t] <= b < prefix[t+1].
constexpr int kMaxItems = 128;
struct LaunchList {
  int nitems;
  int nblocks;
  int items[kMaxItems];        // item -> DevLevel index (or transfer index)
  int prefix[kMaxItems + 1];   // block prefix
};

// The gate of a solve (one per instance, in one array indexed by
// DevLevel::gslot, so that the host reads every gate of a batch with one
// copy).  The host enqueues cycles without waiting for their norms: the last
// norm fold of every cycle runs the gate, which records the norms of each
// FP64 correction in a ring (entry count % kItemRing), applies the stop
// rule of the host's norm step to them and writes the instance's state
// words (DevLevel::stop, warm, skipc, pcur); once it stops, every later
// kernel of the instance returns at once.  The host replays the ring
// (Context::absorb_gate), at most kItemRing corrections behind.
//
// Paired cycles (mixed precision, residual stop rule): after a correction
// whose residual r (and r' at the previous one, m V-cycles earlier)
// predicts that one more cycle cannot reach the tolerance, r * rho > tol
// with rho = (r / r')^(1/m), the next two V-cycles share one correction:
// the first skips it, the second continues from the first's iterate.  The
// per-cycle rate of the FP32 V-cycle is unchanged by it, so the stop falls
// at the same cycle with fewer FP64 passes.  The rate is estimated from the
// solve's own history only, so a solve gives the same bits whatever ran
// before it or beside it.
constexpr int kItemRing = 16;
// zero-guess bits of a smoothing launch: the instance's state word decides
// whether the sweep starts from zero (kZgDyn) or continues (kZgWarm)
constexpr int kZgDyn = 4;
constexpr int kZgWarm = 8;
constexpr double kPairRho0 = 1e-3;   // the first correction: pair when res0 * kPairRho0 > tol
struct ItemGate {
  int stop;                    // the stop rule held (or the first defect ended the solve)
  int count;                   // corrections recorded in this solve
  int cycles;                  // V-cycles run in this solve
  int maxcy, rule, fp64, report, pair;
  double bnorm, res0, tol;
  double rprev;                // relative residual at the last correction (res0 before the first)
  double hist[kItemRing][6];   // per correction: the 4 norms of its fold, the 2 of relmax (FP64)
  int hcyc[kItemRing];         // per correction: the V-cycles it closes (1 or 2)
};

// Whether the next correction should close two V-cycles (the rule above;
// multiplications only, so the host's replay decides alike).
__host__ __device__ inline bool pair_next(double r, double rp, int m, double tol) {
  if (!(rp > 0.0)) return false;
  const double q = r / rp;
  return m == 1 ? r * q > tol : (r * r) * q > tol * tol;
}

// Per-item caller arrays (reference-layout fields), by value.
struct PtrList {
  double* p[kMaxItems];
};

// Cells the V-cycle's prolongation may leave unchanged (DevXfer::skip):
// those that the fine level's first post-smoothing half-sweep (colour 0)
// rewrites from the other colour and b alone.  The z-line sweep updates the
// cells with (i+j) even, but with periodic z it also reads its own old
// values at rows 0 and nz-1 (the lagged wrap); the point sweep updates the
// cells with (i+j+k) even.
// Cells whose residual the V-cycle's restriction takes as zero
// (DevXfer::rskip, Line or Point): those of colour 1, whose equations the
// last half-sweep of the fine level's pre-smoothing has just solved (the
// z-line sweep's lagged periodic-z wrap leaves rows 0 and nz-1 unsolved, so
// a periodic-z level keeps every residual).
enum { kXferSkipNone = 0, kXferSkipLine = 1, kXferSkipLineWrap = 2, kXferSkipPoint = 3 };

// Fine->coarse pair for transfers.
struct DevXfer {
  int fine, coarse;
  int fx, fy, fz;
  int skip;          // kXferSkip*: set for a 2:2:2 transfer whose fine level has a
                     // post-smoothing sweep
  int rskip;         // kXferSkip*: set for a marching restriction (restrict222_ok)
                     // whose fine level has a pre-smoothing sweep
  unsigned pmx, pmy; // prolongation: division of a dual index by the duals per
  int psx, psy;      // x row and per y column (gm_magic multipliers and shifts)
};

constexpr int kBlock = 256;

// Prolongation runs one thread per dual cell: the (up to) 2x2x2 fine cells
// between neighbouring coarse centres.  Per direction a factor-2 transfer
// has nc duals when periodic, nc+1 otherwise (the half-duals at the walls);
// a factor-1 direction has one dual per cell.
__host__ __device__ inline int prolong_duals_dir(int nc, int fac, int per) {
  return fac == 1 ? nc : (per ? nc : nc + 1);
}
__host__ __device__ inline int prolong_duals(const DevLevel& F, const DevLevel& C, const DevXfer& x) {
  return prolong_duals_dir(C.nx, x.fx, F.px) * prolong_duals_dir(C.ny, x.fy, F.py) *
         prolong_duals_dir(C.nz, x.fz, F.pz);
}

// Coarse tail: the levels of an instance at and below the first level with
// at most kTailCells cells (and nz <= kTailNz for the z-line smoother) are
// cycled inside one workgroup of 64*kTailWaves
// threads (one launch per V-cycle instead of ~10 per level).
constexpr int kTailCells = 32768;
constexpr int kTailNz = 32;      // z-line tail levels: whole columns in registers (FP32)
constexpr int kTailNz64 = 16;    // the same for FP64 (register budget)
constexpr int kTailWaves = 8;
constexpr int kTailWavesPoint = 16;   // the point-smoother tail (no line-solve registers)
constexpr int kTailMaxLevels = 16;
// Steps of the tail's sub-V-cycle, precomputed by the host in cycle order:
// code = type | level << 3 | parity << 7 | zero_guess << 8.
enum { kStepZero = 0, kStepHalf = 1, kStepRestrict = 2, kStepCoarse = 3, kStepProlong = 4 };
constexpr int kTailMaxSteps = 1024;
constexpr size_t kTailMinvLds = 16384;   // largest coarsest inverse staged in LDS (bytes)
// A coarsest level of at least kSplitCells cells (an inverse of 512 KiB in
// FP64) is solved by its own kernel across the device: the tail runs in two
// phases around it (one workgroup reading the whole inverse is bound by its
// memory latency, about 65 GB/s).
constexpr int kSplitCells = 256;
struct DevTail {
  int nlev;                       // tail levels, the last is the coarsest
  int copy_u_in;                  // the first tail level's u is read (not zero)
  int zline;                      // smoother: 1 z-line, 0 point red-black
  int nsteps;
  int steps[kTailMaxSteps];
  // LDS byte offsets (16-byte aligned) per tail level: u and f (level
  // precision), the coefficient table (3*(nx+ny+nz) values of the level
  // precision) and the cell widths (nx+ny+nz doubles; not for the coarsest)
  int lds_u[kTailMaxLevels];
  int lds_f[kTailMaxLevels];
  int lds_co[kTailMaxLevels];
  int lds_vol[kTailMaxLevels];
  int lds_minv;                   // the coarsest inverse staged in LDS (level precision), or -1
  int lev[kTailMaxLevels];        // DevLevel index per tail level
  int xf[kTailMaxLevels];         // DevXfer index of tail level d -> d+1
};

void launch_fill_zero(bool fp64, const DevLevel* L, const LaunchList& ll, hipStream_t s);

// ---- V-cycle kernels (amg_kernels.hip); fp64 selects the level precision --
void launch_rb_point(bool fp64, const DevLevel* L, const LaunchList& ll, int parity,
                     int zero_guess, hipStream_t s);
// The point half-sweep marching in z over slabs of 'planes' (8, 4 or 2)
// planes: cdiv(hx*ny, kBlock) * cdiv(nz, planes) blocks per item.
void launch_rb_march(bool fp64, const DevLevel* L, const LaunchList& ll, int parity,
                     int zero_guess, int planes, hipStream_t s);
// Slab length of the marching point sweep for a level (0: one thread per
// cell): the longest of 8, 4, 2 planes that keeps 2^18 threads in a launch.
inline int rb_march_planes(int hx, int ny, int nz) {
  for (int K : {8, 4, 2})
    if (nz >= K && (long long)hx * ny * ((nz + K - 1) / K) >= (1 << 18)) return K;
  return 0;
}
// z-line sweeps run 64 columns per workgroup and 'segs' = ceil(nz_max/kZSeg)
// waves (segments of kZSeg cells) per column group; every item of one
// launch must have nz <= segs*kZSeg.
constexpr int kZSeg = 16;
// rows of a segment whose right-hand-side loads are issued together
#ifndef AMG_ZCHUNK32
#define AMG_ZCHUNK32 16
#endif
#ifndef AMG_ZCHUNK64
#define AMG_ZCHUNK64 8
#endif
constexpr int kZChunk32 = AMG_ZCHUNK32;
constexpr int kZChunk64 = AMG_ZCHUNK64;
constexpr int kZMaxSegs = 16;
// form: 0 the launch's choice (FP32: the lean form when the launch has more
// blocks than the device has CUs), 1 the short form (FP32 only), 2 the lean
// form.  Both forms give the same bits.
void launch_zline(bool fp64, const DevLevel* L, const LaunchList& ll, int parity,
                  int zero_guess, int segs, hipStream_t s, int form = 0);
void launch_restrict(bool fp64, const DevLevel* L, const DevXfer* X, const LaunchList& ll,
                     hipStream_t s);
// Residual + restriction marching in z, for a transfer with factor 2 in every
// direction and fine sizes twice the coarse ones, on coarse levels of at
// least kRestrictMinCells cells (restrict222_ok).  A thread walks
// restrict222_k coarse planes: up to 8, fewer where that would leave fewer
// than kRestrictThreads threads (a smaller level is latency-bound).  Blocks
// per item = restrict222_blocks(coarse).  Same result as launch_restrict.
#ifndef AMG_RESTRICT_MIN_CELLS
#define AMG_RESTRICT_MIN_CELLS (1 << 18)
#endif
constexpr long long kRestrictMinCells = AMG_RESTRICT_MIN_CELLS;
constexpr long long kRestrictThreads = 1 << 18;
__host__ __device__ inline int restrict222_k(const DevLevel& C) {
  const long long cells = (long long)C.nx * C.ny * C.nz;
  int k = 8;
  while (k > 1 && cells / k < kRestrictThreads) k >>= 1;
  return k;
}
inline bool restrict222_ok(const DevLevel& F, const DevLevel& C, const DevXfer& x) {
  return x.fx == 2 && x.fy == 2 && x.fz == 2 && F.nx == 2 * C.nx && F.ny == 2 * C.ny &&
         F.nz == 2 * C.nz && F.nx == 2 * F.hx && C.nx == 2 * C.hx &&
         (long long)C.nx * C.ny * C.nz >= kRestrictMinCells;
}
inline int restrict222_blocks(const DevLevel& C) {
  const int kr = restrict222_k(C);
  return (int)(((long long)C.nx * C.ny + kBlock - 1) / kBlock) * ((C.nz + kr - 1) / kr);
}
void launch_restrict222(bool fp64, const DevLevel* L, const DevXfer* X, const LaunchList& ll,
                        hipStream_t s, bool use_skip = false);
void launch_prolong(bool fp64, const DevLevel* L, const DevXfer* X, const LaunchList& ll,
                    hipStream_t s, bool use_skip = false);
void launch_coarse(bool fp64, const DevLevel* L, const LaunchList& ll, hipStream_t s);
// Separable direct solve of the items' levels (u = A^-1 f), for levels of
// at most kDirectCells cells with nx, ny multiples of 16 and nz <= 64,
// computed in FP64 for an FP64 V-cycle and in FP32 for the FP32 V-cycle of
// a mixed solve.  Every pass splits its output into slabs of kDirectCols
// columns, one workgroup each (a level of 64^3 fills the device with 256):
// 'planes' has nz*nx/kDirectCols blocks per item, 'slabs' nx*ny/kDirectCols;
// lds = the largest direct_lds() of the items, at most direct_lds_limit().
// (A larger FP32 level fits the LDS -- 128^3 does -- but its passes take
// longer than the levels they would replace.)
constexpr long long kDirectCells = 1 << 18;
constexpr int kDirectCols = 16;
constexpr int kDirectMaxNz = 64;
size_t direct_lds(int nx, int ny, int nz, bool fp64);
size_t direct_lds_limit();
void launch_direct(bool fp64, const DevLevel* L, const LaunchList& planes, const LaunchList& slabs,
                   size_t lds, hipStream_t s);
// the coarse tail's short-column z-line sweep as a standalone launch (tests)
void launch_zline_col(bool fp64, const DevLevel* L, const LaunchList& ll, int parity,
                      int zero_guess, hipStream_t s);
// one workgroup per item (items index DevTail entries), lds_bytes of
// dynamic LDS (the largest tail of the items); tail_lds_limit() is the most
// a launch may request on this device.
// variant: the kernel's column length class (tail_nz_class) and whether
// every tail transfer coarsens 2:1 in all three directions.
int tail_nz_class(bool fp64, int nz_max, bool zline);
// phase 0: the whole sub-cycle; 1: down to the coarsest level, then every
// tail level's u and f are saved to their global arrays; 2: those restored,
// up from the coarsest level (whose u the coarse kernel wrote).
void launch_tail(bool fp64, int nz_class, bool all222, const DevLevel* L, const DevXfer* X,
                 const DevTail* T, const LaunchList& ll, size_t lds_bytes, hipStream_t s,
                 int phase = 0);
// u = M f on coarsest levels (row-major inverse), one wave per row.
void launch_coarse_wide(bool fp64, const DevLevel* L, const LaunchList& ll, hipStream_t s);
size_t tail_lds_limit();

// ---- FP64 finest-level kernels (items index FP64 finest entries) -------------
// Each block writes nvals partial maxima to partial[block*nvals + v]; the
// stencil kernels march in z (blocks per item: march_blocks), the change
// kernel handles kFinePer cells per thread (cdiv(ncell, kBlock*kFinePer));
// launch_finalize folds them per item into out[item*ostride + v].
// residual_max:     {max|b - A u|, max|b|}
// change_max:       {max|u - u2|, max|u|}        (u2 holds the previous iterate)
// correct_defect32: u2 = u + e (e = FP32 u of 'link'); r = b' - A u2 -> FP32 f
//                   of 'link', with b' = b - bshift; {max|r|, max|e|, max|u2|}
constexpr int kFinePer = 8;
// The FP64 finest-level stencil kernels (residual, defect, correct+defect)
// march in z over slabs of kMarchK planes: blocks per item =
// cdiv(2*hx*ny, kBlock) * cdiv(nz, kMarchK), one partial per block.
#ifndef AMG_MARCHK
#define AMG_MARCHK 16
#endif
constexpr int kMarchK = AMG_MARCHK;
inline long long march_blocks(int hx, int ny, int nz) {
  return ((2LL * hx * ny + kBlock - 1) / kBlock) * ((nz + kMarchK - 1) / kMarchK);
}
void launch_residual_max(const DevLevel* L, const LaunchList& ll, double* partial, hipStream_t s);
void launch_change_max(const DevLevel* L, const LaunchList& ll, double* partial, hipStream_t s);
void launch_correct_defect32(const DevLevel* L, const LaunchList& ll, double* partial,
                             hipStream_t s);
void launch_zero(double* a, int n, hipStream_t s);
// the ingest folds kWsumPer cells per thread: blocks per item = cdiv(ncell, kBlock*kWsumPer)
constexpr int kWsumPer = 16;

// Division by an invariant d >= 1 of any 32-bit n (Granlund and Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, fig.
// 4.1): with l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// t = mulhi(m, n) and q = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0).
// Returns m; *sh packs the two shifts as s1 | s2 << 8.
inline unsigned gm_magic(unsigned d, int* sh) {
  int l = 0;
  while ((1ull << l) < d) ++l;
  const int s1 = l < 1 ? l : 1, s2 = l > 1 ? l - 1 : 0;
  *sh = s1 | (s2 << 8);
  return (unsigned)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
}
// The end of a gated cycle, run by the cycle's last norm fold (ig nullptr:
// no gate): item t's norms are out + t*4 (and outc + t*2, the 2 relmax
// values of an FP64 solve that reports them), its gate ig[Lt[item].gslot].
struct GateArgs {
  const double* out;
  const double* outc;
  DevLevel* Lt;
  int nL;
  ItemGate* ig;
};
// Clears the state words of the instances of the items of ll.
// only_stopped: just the instances whose solve has stopped (a reset that may
// be enqueued before the host knows whether it has).
void launch_gate_reset(DevLevel* L, int nL, const LaunchList& ll, hipStream_t s,
                       bool only_stopped = false);
// Initialises the gates of the items of ll (FP64 finest levels) from the
// first defect's norms out + t*4 = {max|r|, max|b|, ..}; an item whose
// solve ends there (NaN, or the residual rule already holds) is stopped.
void launch_gate_init(const LaunchList& ll, const double* out, ItemGate* ig, DevLevel* L, int nL,
                      int maxcy, int rule, double tol, int report, hipStream_t s);
void launch_finalize(const LaunchList& ll, const double* partial, int nvals, int op_sum,
                     double* out, int ostride, hipStream_t s, const GateArgs* gate = nullptr);

// Conversions between the reference's ghosted Fortran arrays (one pointer
// per item in 'ptr') and the FP64 finest level.
// from_fortran: dst = (which==0 ? u : f) = src interior (+ bcfold if add_fold)
// wsum:         {sum V x} (x = u or f), or {sum V rho} when which == 2;
//               march_blocks per item (z slabs of kMarchK planes)
// to_fortran:   dst interior = u, or u - (c[item]/vrsum) rho(z) for singular
//               items when c is given (c: the sums of launch_wsum)
// ingest:       f = rhs interior + bcfold, u = phi interior, and one partial
//               of sum V*f per block (kWsumPer cells per thread)
// first_defect: the first defect of a solve (march_blocks per item): FP64
//               {max|b'-Au|, max|b'|} with b' = f - c[item]/vsum written back
//               in place for singular items (c: the sums of launch_ingest);
//               mixed the same norms and the FP32 defect b' - Au, with f kept
//               and the shift recorded in the item's bshift
// copy_u:       u2 = u
void launch_from_fortran(const DevLevel* L, const LaunchList& ll, const PtrList& ptr,
                         int which, int add_fold, hipStream_t s);
void launch_wsum(const DevLevel* L, const LaunchList& ll, int which, double* partial,
                 hipStream_t s);
void launch_to_fortran(const DevLevel* L, const LaunchList& ll, const PtrList& ptr,
                       const double* c, hipStream_t s);
void launch_ingest(const DevLevel* L, const LaunchList& ll, const PtrList& rhs, const PtrList& phi,
                   double* partial, hipStream_t s);
void launch_first_defect(bool fp64, const DevLevel* L, const LaunchList& ll, const double* c,
                         double* partial, hipStream_t s);
void launch_copy_u(const DevLevel* L, const LaunchList& ll, const int* active, hipStream_t s);

// Ghost layers of a ghosted FP64 array (the serial path of mgMK/mg_bdry.F90):
// the periodic wrap of one direction over full planes, or every ghost cell
// at once (the wraps, then the faces' rules of bc, in one pass).
struct GhostBC {
  int type[7];      // faces 1..6: 0 periodic (no rule), 1 Neumann, 2 Dirichlet
  double v[7];      // the face value
  double vd[7];     // v * dxc of the face's edge cell (the Neumann offset)
};
void launch_ghost_wrap(double* a, int nx, int ny, int nz, int dir, hipStream_t s);
void launch_ghost_fill(double* a, int nx, int ny, int nz, const GhostBC& bc, hipStream_t s);

// Projection step (reference layout, FP64).
void launch_psource(const double* U, const double* V, const double* W, double* SRC,
                    const double* rho, const double* rdxc, const double* rdyc,
                    const double* rdzc, int nx, int ny, int nz, double rdt, hipStream_t s);
// wallmask bit d (x=1, y=2, z=4): the upper face of direction d is a
// pressure-Neumann wall whose normal velocity is left to the velocity BC.
void launch_pgrad(const double* p, double* U, double* V, double* W, const double* rho,
                  const double* rdxe, const double* rdye, const double* rdze, int nx, int ny,
                  int nz, double dt, int wallmask, hipStream_t s);
// kDivPer cells per thread: blocks = cdiv(cells, kBlock*kDivPer), each
// writes 4 partials; the finalize writes 4 values.
constexpr int kDivPer = 16;
int divergence_blocks(int nx, int ny, int nz);
void launch_divergence(const double* U, const double* V, const double* W, double* div,
                       const double* rdxc, const double* rdyc, const double* rdzc, int nx,
                       int ny, int nz, double* partial, hipStream_t s);
void launch_divergence_finalize(const double* partial, int nblocks, double* out, hipStream_t s);

}  // namespace amg
