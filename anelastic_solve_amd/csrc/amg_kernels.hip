// amg_kernels.hip -- gfx950 kernels of the multigrid pressure solve.
//
// Every V-cycle kernel runs over a work list: one launch covers the same
// step of several instances/levels (LaunchList in amg_internal.h), so a
// single solve and a batch of uneven solves share the same code.  All
// kernels are HBM-streaming stencils (about 0.5 flop/byte): no MFMA.
//
// Field pointers carry their address space: global (1) in the standalone
// kernels, LDS (3) inside the coarse tail.  Generic (flat) accesses would
// make every load wait for both the vector-memory and the LDS counters and
// serialise the loads of a line solve; the element routines are therefore
// templates over the pointer types and the kernels cast the level
// descriptors' pointers before use.  Chequer indices are 32-bit.
#include "amg.h"
#include "amg_internal.h"

#include <algorithm>
#include <cmath>

namespace amg {
namespace {

template <int AS, class T> using sp = __attribute__((address_space(AS))) T*;
constexpr int kG = 1;   // global memory
constexpr int kS = 3;   // LDS

template <class T> __device__ __forceinline__ sp<kG, T> gp(T* p) { return (sp<kG, T>)p; }
template <class T> __device__ __forceinline__ sp<kG, T> gfield(void* p) {
  return (sp<kG, T>)static_cast<T*>(p);
}
template <class T> __device__ __forceinline__ sp<kG, const T> gfield(const void* p) {
  return (sp<kG, const T>)static_cast<const T*>(p);
}

__device__ __forceinline__ int find_item(const int* prefix, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// A gated cycle has stopped (DevLevel::stop): the kernel returns at once.
__device__ __forceinline__ bool level_off(const DevLevel& L) { return L.stop != 0; }

// Zero-guess flags of a smoothing launch (bit 0: the other colour is zero;
// bit 1: the own colour's old values are zero).  With kZgDyn, the finest
// level's first sweep of a mixed cycle, they hold only when the cycle does
// not continue the previous one's iterate (DevLevel::warm); kZgWarm marks
// the launch that does that sweep when the cycle continues it.
__device__ __forceinline__ int zero_flags(const DevLevel& L, int zero_guess) {
  if (zero_guess & kZgWarm) return 0;
  if (!(zero_guess & kZgDyn)) return zero_guess;
  return L.warm ? 0 : (zero_guess & 3);
}

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2);
// remap so that every XCD walks a contiguous range of tiles and neighbouring
// rows share an L2.  A bijection on [0, n).
__device__ __forceinline__ int xcd_swizzle(int b, int n) {
  const int q = n >> 3, rem = n & 7, x = b & 7, idx = b >> 3;
  return x * q + min(x, rem) + idx;
}

// Reciprocal for the line solves: the hardware rcp (1 ulp) in FP32, a
// correctly rounded division in FP64.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }

// a / d in the point smoother: times the hardware reciprocal in FP32 (an
// IEEE division is about a dozen instructions), a division in FP64.
__device__ __forceinline__ float over(float a, float d) { return a * __builtin_amdgcn_rcpf(d); }
__device__ __forceinline__ double over(double a, double d) { return a / d; }

template <class T> __device__ __forceinline__ sp<kG, const T> coefs(const DevLevel& L);
template <> __device__ __forceinline__ sp<kG, const double> coefs<double>(const DevLevel& L) {
  return gp(L.co64);
}
template <> __device__ __forceinline__ sp<kG, const float> coefs<float>(const DevLevel& L) {
  return gp(L.co32);
}

// NaN-propagating maximum (fmax ignores NaN).
__device__ __forceinline__ double nmax(double a, double b) {
  return (a != a) ? a : ((b != b) ? b : fmax(a, b));
}

// max(acc, |x|) for accumulators of magnitudes (acc >= 0 or a NaN with the
// sign bit clear): the bit patterns of non-negative doubles order as their
// values, and a NaN orders above +inf, so one 64-bit integer maximum
// propagates NaN without a branch.
__device__ __forceinline__ double amax(double acc, double x) {
  const unsigned long long a = (unsigned long long)__double_as_longlong(acc);
  const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(x));
  return __longlong_as_double((long long)(a > b ? a : b));
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide reduction of NV values; thread 0..NV-1 writes out[q].
template <int NV, bool SUM>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double* __restrict__ out) {
  __shared__ double sm[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double x = SUM ? wave_sum(v[q]) : wave_max(v[q]);
    if (lane == 0) sm[q][w] = x;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double a = sm[threadIdx.x][0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
      a = SUM ? a + sm[threadIdx.x][i] : nmax(a, sm[threadIdx.x][i]);
    out[threadIdx.x] = a;
  }
}

// Chequer-layout index of (m, j, k) in the half-row of colour c.
__device__ __forceinline__ int cidx(int m, int j, int k, int c, int hx, int ny) {
  return ((k * ny + j) * 2 + c) * hx + m;
}

// Seven chequer indices of cell (i,j,k): centre, x-, x+, y-, y+, z-, z+.
struct Nb {
  int c, xm, xp, ym, yp, zm, zp;
};

__device__ __forceinline__ Nb neighbours(const DevLevel& L, int i, int j, int k) {
  const int hx = L.hx, ny = L.ny, nz = L.nz;
  const int col = (i + j) & 1, oth = col ^ 1, m = i >> 1, r = i & 1;
  int ml = m - 1 + r; if (ml < 0) ml = hx - 1;
  int mr = m + r;     if (mr >= hx) mr = 0;
  const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
  const int km = (k == 0) ? nz - 1 : k - 1, kp = (k == nz - 1) ? 0 : k + 1;
  Nb n;
  n.c = cidx(m, j, k, col, hx, ny);
  const int ro = cidx(0, j, k, oth, hx, ny);
  n.xm = ro + ml;
  n.xp = ro + mr;
  n.ym = cidx(m, jm, k, oth, hx, ny);
  n.yp = cidx(m, jp, k, oth, hx, ny);
  n.zm = cidx(m, j, km, col, hx, ny);
  n.zp = cidx(m, j, kp, col, hx, ny);
  return n;
}

// Decode a chequer element index into (i, j, k).
// floor(n / d) by the invariant divisor of a level (DevLevel::gmhx/gshx,
// gmny/gsny: Granlund-Montgomery multiplier and shifts, exact for every
// 32-bit n).
__device__ __forceinline__ int gm_div(int n, unsigned m, int sh) {
  const unsigned un = (unsigned)n, t = __umulhi(m, un);
  return (int)((t + ((un - t) >> (sh & 0xff))) >> (sh >> 8));
}

__device__ __forceinline__ void decode(const DevLevel& L, int e, int& i, int& j, int& k) {
  const int t = gm_div(e, L.gmhx, L.gshx);
  const int m = e - t * L.hx;
  const int c = t & 1;
  const int row = t >> 1;
  k = gm_div(row, L.gmny, L.gsny);
  j = row - k * L.ny;
  i = 2 * m + ((c + j) & 1);
}

// Off-diagonal stencil sum, in a fixed order: x-, x+, y-, y+, z-, z+.
template <class T, class CP>
__device__ __forceinline__ T offsum(CP co, const DevLevel& L, int i, int j, int k, T uxm, T uxp,
                                    T uym, T uyp, T uzm, T uzp) {
  const int nx = L.nx, ny = L.ny;
  const auto cy = co + 3 * nx;
  const auto cz = cy + 3 * ny;
  T s = co[i] * uxm + co[nx + i] * uxp;
  s += cy[j] * uym;
  s += cy[ny + j] * uyp;
  s += cz[k] * uzm;
  s += cz[L.nz + k] * uzp;
  return s;
}

template <class T, class CP>
__device__ __forceinline__ T diag_at(CP co, const DevLevel& L, int i, int j, int k) {
  const int nx = L.nx, ny = L.ny;
  const auto cy = co + 3 * nx;
  const auto cz = cy + 3 * ny;
  return (cz[2 * L.nz + k] + cy[2 * ny + j]) + co[2 * nx + i];
}

// ---------------------------------------------------------------------------
// Point red-black Gauss-Seidel half-sweep (mgMK/mg_relax.F90:146-168):
// u = (f - sum w_n u_n)/diag on the cells with (i+j+k)&1 == parity.  Element
// e enumerates the cells of the colour; in plane k the colour is column
// colour (parity+k)&1, so consecutive e are one contiguous half-row.
template <class T, class UP, class FP, class CP>
__device__ __forceinline__ void rb_point_mjk(const DevLevel& L, UP u, FP f, CP co, int m, int j,
                                             int k, int parity, int zero_guess) {
  const int c = (k + parity) & 1;
  const int i = 2 * m + ((c + j) & 1);
  const Nb n = neighbours(L, i, j, k);
  const T d = diag_at<T>(co, L, i, j, k);
  T val;
  if (zero_guess & 1) {
    val = over(f[n.c], d);
  } else {
    const T s = offsum<T>(co, L, i, j, k, u[n.xm], u[n.xp], u[n.ym], u[n.yp], u[n.zm], u[n.zp]);
    val = over(f[n.c] - s, d);
  }
  u[n.c] = val;
}

template <class T, class UP, class FP, class CP>
__device__ __forceinline__ void rb_point_elem(const DevLevel& L, UP u, FP f, CP co, int e,
                                              int parity, int zero_guess) {
  const int row = gm_div(e, L.gmhx, L.gshx), m = e - row * L.hx;
  const int k = gm_div(row, L.gmny, L.gsny), j = row - k * L.ny;
  rb_point_mjk<T>(L, u, f, co, m, j, k, parity, zero_guess);
}

// The same with the element index decoded by multiply-high reciprocals
// (DevLevel::mhx, mny), exact for the small levels of the coarse tail.
__device__ __forceinline__ int fast_div(int e, unsigned mag) {
  return mag ? (int)__umulhi((unsigned)e, mag) : e;
}
template <class T, class UP, class FP, class CP>
__device__ __forceinline__ void rb_point_small(const DevLevel& L, UP u, FP f, CP co, int e,
                                               int parity, int zero_guess) {
  const int row = fast_div(e, L.mhx), m = e - row * L.hx;
  const int k = fast_div(row, L.mny), j = row - k * L.ny;
  rb_point_mjk<T>(L, u, f, co, m, j, k, parity, zero_guess);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_rb_point(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                     int parity, int zero_guess) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e < L.hx * L.ny * L.nz)
    rb_point_elem<T>(L, gfield<T>(L.u), gfield<const T>((const void*)L.f), coefs<T>(L), e, parity,
                     zero_flags(L, zero_guess));
}

// The same half-sweep marching in z (large levels): a thread owns position
// (m, j) of a plane, both halves, and K consecutive planes.  In plane k the
// colour's cells are in half h = (parity + k) & 1, so the thread updates one
// cell per plane, i = 2m + ((h + j) & 1).  Its z-neighbours are the same
// half of the planes k-1 and k+1 (the other colour), so the one above plane
// k is the one below plane k+2 and is loaded once.  Every load of the slab
// (the other colour only, which the sweep does not write) is issued before
// the first store, the coefficients of the thread's two columns and row once
// per thread, those of z (wave-uniform) from the scalar cache.  The update is
// rb_point_mjk's, in its order.
template <class T, int K>
__global__ __launch_bounds__(kBlock) void k_rb_march(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                     int parity, int zero_guess) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  const int zg = zero_flags(L, zero_guess);
  const int hx = L.hx, ny = L.ny, nz = L.nz, nx = L.nx;
  const int npos = hx * ny, ncb = (npos + kBlock - 1) / kBlock;
  const int lb = b - ll.prefix[t];
  const int bk = lb / ncb, pos = (lb - bk * ncb) * kBlock + threadIdx.x;
  if (pos >= npos) return;
  const int j = gm_div(pos, L.gmhx, L.gshx), m = pos - j * hx;
  const int k0 = bk * K;
  const sp<kG, const T> co = coefs<T>(L);
  const sp<kG, const T> cy = co + 3 * nx;
  const sp<kG, const T> cz = cy + 3 * ny;
  const auto u = gfield<T>(L.u);
  const auto f = gfield<const T>((const void*)L.f);
  // the thread's two columns (half h: i = 2m + ((h + j) & 1))
  int ml[2], mr[2];
  T wxl[2], wxr[2], dxx[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = 2 * m + ((h + j) & 1), r = i & 1;
    ml[h] = m - 1 + r; if (ml[h] < 0) ml[h] = hx - 1;
    mr[h] = m + r;     if (mr[h] >= hx) mr[h] = 0;
    wxl[h] = co[i];
    wxr[h] = co[nx + i];
    dxx[h] = co[2 * nx + i];
  }
  const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
  const T wys = cy[j], wyn = cy[ny + j], dyy = cy[2 * ny + j];
  const bool nbr = !(zg & 1);
  // plane k of half h, row jj: index of position mm
  auto at = [&](int k, int h, int jj, int mm) { return ((k * ny + jj) * 2 + h) * hx + mm; };
  T fv[K], xl[K], xr[K], ys[K], yn[K], zp[K], zm01[2];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int k = min(k0 + s, nz - 1);              // a short last slab re-reads its last plane
    const int h = (parity + k0 + s) & 1, o = h ^ 1;
    fv[s] = f[at(k, h, j, m)];
    if (nbr) {
      xl[s] = u[at(k, o, j, ml[h])];
      xr[s] = u[at(k, o, j, mr[h])];
      ys[s] = u[at(k, o, jm, m)];
      yn[s] = u[at(k, o, jp, m)];
      zp[s] = u[at(k == nz - 1 ? 0 : k + 1, h, j, m)];
    }
  }
  if (nbr) {
    // below the slab's first two planes
    const int h0 = (parity + k0) & 1;
    zm01[0] = u[at(k0 == 0 ? nz - 1 : k0 - 1, h0, j, m)];
    zm01[1] = u[at(k0, h0 ^ 1, j, m)];
  }
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int k = k0 + s;
    if (k >= nz) break;
    const int h = (parity + k) & 1;
    const T d = (cz[2 * nz + k] + dyy) + dxx[h];
    T val;
    if (!nbr) {
      val = over(fv[s], d);
    } else {
      const T zmv = s < 2 ? zm01[s] : zp[s - 2];
      T a = wxl[h] * xl[s] + wxr[h] * xr[s];
      a += wys * ys[s];
      a += wyn * yn[s];
      a += cz[k] * zmv;
      a += cz[nz + k] * zp[s];
      val = over(fv[s] - a, d);
    }
    u[at(k, h, j, m)] = val;
  }
}

// ---------------------------------------------------------------------------
// z-line red-black Gauss-Seidel half-sweep (mgMK/mg_relax.F90:423-600, LJ):
// the columns (i,j) of one colour ((i+j)&1 == parity) solve their
// tridiagonal z-system with the other colour's current values on the
// right-hand side.  A workgroup takes 64 columns (one per lane) and splits
// each column into P segments of S cells (wave p owns segment p).  Every
// segment eliminates its rows once, forward, for three right-hand sides:
// its data, and unit couplings to the last value of the segment below and
// to the first value of the segment above.  Its end values are then linear
// in those two unknowns; the 2P interface unknowns are solved by parallel
// cyclic reduction in LDS, and one back substitution per segment gives its
// values.  The result equals a Thomas solve of the whole column up to
// rounding.  Periodic z: the wrap couplings are lagged (the other end's old
// value on the right-hand side), in the order of the oracle's restatement.
// A row's coupling past either end of the column (a wrap or a wall ghost)
// only ever multiplies an interface value that does not exist, which the
// reduction takes as zero, so the rows are used as the tables hold them.

// Row access at a uniform row base plus a per-lane 32-bit byte offset: the
// global load/store form with a 64-bit scalar base and a 32-bit vector
// offset, so a row's address costs no vector arithmetic.
template <class T> __device__ __forceinline__ T ld_row(sp<kG, const T> row, unsigned boff) {
  return *(sp<kG, const T>)((sp<kG, const char>)row + boff);
}
template <class T> __device__ __forceinline__ void st_row(sp<kG, T> row, unsigned boff, T v) {
  *(sp<kG, T>)((sp<kG, char>)row + boff) = v;
}
// The same, non-temporal: for data that is not read again before it would
// have left the L2 (the streaming stages of the mixed outer step).
template <class T> __device__ __forceinline__ void st_row_nt(sp<kG, T> row, unsigned boff, T v) {
  __builtin_nontemporal_store(v, (T*)(sp<kG, T>)((sp<kG, char>)row + boff));
}
template <class T> __device__ __forceinline__ T ld_row_nt(sp<kG, const T> row, unsigned boff) {
  return __builtin_nontemporal_load((const T*)(sp<kG, const T>)((sp<kG, const char>)row + boff));
}

// NBR false: the other colour is zero (zero_guess bit 0, the first
// half-sweep from u = 0) and is not read.  LEAN: the register-lean form of
// the segment elimination (below).  H: a wave holds two segments of 32
// columns (lanes 0-31 segment 2w, lanes 32-63 segment 2w+1), so a launch of
// few column groups spreads over twice the CUs; p and lane (the column
// within the workgroup's 32) are then per lane, and the z rows are staged
// in LDS.  Every lane does the same operations in either shape.
template <class T, int S, int C, bool NBR, bool LEAN, bool H>
__device__ __forceinline__ void zline_group(const DevLevel& L, int parity, int zero_guess, int col0,
                                            int p, int lane, int P, T* __restrict__ lds, bool act) {
  const int hx = L.hx, ny = L.ny, nz = L.nz, nx = L.nx;
  const int ncol = hx * ny;
  const bool live = act && col0 < ncol;
  const int col = live ? col0 : ncol - 1;
  const int j = gm_div(col, L.gmhx, L.gshx), m = col - j * hx;
  const int c = parity, oth = parity ^ 1;
  const int i = 2 * m + ((c + j) & 1), r = i & 1;
  int ml = m - 1 + r; if (ml < 0) ml = hx - 1;
  int mr = m + r;     if (mr >= hx) mr = 0;
  const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
  const sp<kG, const T> co = coefs<T>(L);
  const sp<kG, const T> cy = co + 3 * nx;
  const sp<kG, const T> cz = cy + 3 * ny;
  const T wxl = co[i], wxr = co[nx + i], wys = cy[j], wyn = cy[ny + j];
  const T dxy = co[2 * nx + i] + cy[2 * ny + j];
  const sp<kG, T> u = gfield<T>(L.u);
  const sp<kG, const T> f = gfield<const T>((const void*)L.f);
  const int plane = ny * 2 * hx;
  // byte offsets within a plane (the same in every row)
  constexpr unsigned W = sizeof(T);
  const unsigned bown = (unsigned)cidx(m, j, 0, c, hx, ny) * W;
  const unsigned xo = (unsigned)cidx(0, j, 0, oth, hx, ny);
  const unsigned bxl = (xo + ml) * W, bxr = (xo + mr) * W;
  const unsigned bso = (unsigned)cidx(m, jm, 0, oth, hx, ny) * W;
  const unsigned bno = (unsigned)cidx(m, jp, 0, oth, hx, ny) * W;
  const bool pz = L.pz != 0;
  const int Pi = (nz + S - 1) / S;            // segments of this item (<= blockDim.y)
  const int ks = p * S;
  const int len = max(0, min(S, nz - ks));    // 0 for spare waves of a shorter item
  // A (sub-diagonal), E (super-diagonal) and the z part of the diagonal
  // depend on k only and p is wave-uniform, so they are scalar values
  // (uniform loads issued with the rest, read without memory latency in the
  // serial elimination chain); the lane's x/y diagonal share dxy is added
  // at use.  All loads -- these rows, the lagged wrap values and the
  // right-hand-side data -- are unconditional (every address is in bounds,
  // uses are selected) and issued before the first wait: a sweep pays one
  // memory round trip.
  // FP64 rows do not fit in scalar registers, and with H a wave's two
  // segments have different rows: they are staged in LDS by lanes 0..S-1
  // (of each half) and read back as broadcasts (LDS_COEF).
  constexpr bool LDS_COEF = sizeof(T) == 8 || H;
  constexpr int CW = H ? 32 : 64;             // columns of a workgroup
  constexpr int SR = LDS_COEF ? 1 : S;
  T A[SR], E[SR], Dz[SR];
  T* const cA = lds + (sizeof(T) == 8 ? 6 : 12) * P * CW + p * 3 * S;
  T* const cE = cA + S;
  T* const cD = cE + S;
  T za = T(0), ze = T(0), zd = T(0);
  if constexpr (LDS_COEF) {
    const int kq = min(ks + (lane & (S - 1)), nz - 1);
    za = cz[kq];
    ze = cz[nz + kq];
    zd = cz[2 * nz + kq];
  } else {
    // rows kc..kc+S-1 from one base (three 16-word scalar loads); the tables
    // end with kZSeg spare words, so a ragged or spare segment stays in bounds
    const sp<kG, const T> cr = cz + min(ks, nz - 1);
#pragma unroll
    for (int q = 0; q < S; ++q) {
      A[q] = cr[q];
      E[q] = cr[nz + q];
      Dz[q] = cr[2 * nz + q];
    }
  }
  // zero_guess bit 1: this colour's old values are zero (no lagged wrap)
  const bool lag = pz && !(zero_guess & 2);
  const T cz_first = cz[0], cz_last = cz[2 * nz - 1];
  const T ubot = ld_row<T>(u + (nz - 1) * plane, bown), utop = ld_row<T>(u, bown);
  const T lagb = (lag && p == 0) ? cz_first * ubot : T(0);
  const T lagt = (lag && ks + len == nz) ? cz_last * utop : T(0);
  T g[S];
#pragma unroll
  for (int q0 = 0; q0 < S; q0 += C) {
    // keep this chunk's loads behind the previous chunk's arithmetic
    if (q0 > 0) __builtin_amdgcn_sched_barrier(0);
    T fv[C], xl[C], xr[C], ys[C], yn[C];
#pragma unroll
    for (int qq = 0; qq < C; ++qq) {
      const int kb = min(ks + q0 + qq, nz - 1) * plane;   // rows clamped: in bounds
      fv[qq] = ld_row<T>(f + kb, bown);
      if constexpr (NBR) {
        const sp<kG, const T> ur = u + kb;
        xl[qq] = ld_row<T>(ur, bxl);
        xr[qq] = ld_row<T>(ur, bxr);
        ys[qq] = ld_row<T>(ur, bso);
        yn[qq] = ld_row<T>(ur, bno);
      }
    }
    if (LDS_COEF && q0 == 0) {
      if (lane < S) {
        cA[lane] = za;
        cE[lane] = ze;
        cD[lane] = zd;
      }
      __syncthreads();
    }
#pragma unroll
    for (int qq = 0; qq < C; ++qq) {
      if constexpr (NBR) {
        T sn = wxl * xl[qq] + wxr * xr[qq];
        sn += wys * ys[qq];
        sn += wyn * yn[qq];
        g[q0 + qq] = fv[qq] - sn;
      } else {
        g[q0 + qq] = fv[qq];
      }
    }
  }
  g[0] -= lagb;                                            // row 0 of the column (p == 0)
#define A(q) (LDS_COEF ? cA[q] : A[LDS_COEF ? 0 : (q)])
#define E(q) (LDS_COEF ? cE[q] : E[LDS_COEF ? 0 : (q)])
#define D(q) ((LDS_COEF ? cD[q] : Dz[LDS_COEF ? 0 : (q)]) + dxy)
  const T a_first = A(0);                                  // coupling to u_{ks-1}
  // phase 1: forward elimination of the segment's rows for the data (g, in
  // place), the coupling a_first to u_{ks-1} entering row 0 (v) and the
  // coupling E(len-1) to u_{ks+len} entering the last row (w).  The first
  // values of the three solutions accumulate during the pass,
  // x_0 = sum_q pi_q x'_q with pi_q = prod_{i<q} (-E(i) rd_i); the last
  // values are the last row's.
  // Short form: the eliminated v (vv) and super-diagonal (cp) are stored.
  // LEAN form (two values per row instead of three, so a 1024-lane block
  // fits in 64 registers and two blocks share a CU): only the pivot
  // reciprocals rd are stored, and v and cp are recomputed from them with
  // the same operations.  Both forms round alike, so every launch gives the
  // same values whichever form it takes.
  // A full segment (n == S, every segment but the last of a ragged column)
  // instantiates the passes with a constant n, so the unrolled chains carry
  // no branches.
  constexpr int SS = LEAN ? 1 : S, SL = LEAN ? S : 1;
  T vv[SS], cp[SS], rdv[SL];
  T y0 = T(0), yl = T(0), v0 = T(0), vl = T(0), w0 = T(0), wl = T(0);
  auto factor = [&](const int len) {
    T cq = T(0), y = T(0), v = T(0), pi = T(1);
#pragma unroll
    for (int q = 0; q < S; ++q) {
      if (q < len) {
        if (q == len - 1) g[q] -= lagt;                    // row nz-1 (last segment)
        const T a = (q == 0) ? T(0) : A(q);
        const T rd = rcp(D(q) - a * cq);
        y = (g[q] - a * y) * rd;
        v = (q == 0) ? a_first * rd : -(a * rd) * v;
        cq = E(q) * rd;
        g[q] = y;
        if (q == 0) { y0 = y; v0 = v; }
        else { y0 += pi * y; v0 += pi * v; }
        if (q == len - 1) { yl = y; vl = v; wl = cq; w0 = pi * cq; }
        else pi = (q == 0) ? -cq : pi * -cq;
        if constexpr (LEAN) {
          rdv[q] = rd;
        } else {
          vv[q] = v;
          cp[q] = cq;
        }
      }
    }
  };
  if (sizeof(T) == 4 && len == S) factor(S);
  else if (len > 0) factor(len);
  // phase 2: reduced system X_p = (t_p, b_p) (first/last values of segment p)
  //   t_p + v0_p b_{p-1} + w0_p t_{p+1} = y0_p
  //   b_p + vl_p b_{p-1} + wl_p t_{p+1} = yl_p
  // with (y, v, w) the segment's solutions for the three right-hand sides,
  // taken at its first and last rows.  Parallel cyclic reduction over the
  // segments: at distance d every segment substitutes b_{p-d} (row b of
  // p-d) and t_{p+d} (row t of p+d); the couplings keep their shape (to
  // b_{p-2d} and t_{p+2d} only), so after ceil(log2 P) steps every segment
  // holds its own values.  Segments past either end of the column
  // contribute zeros.  Every wave takes part.
  // FP32 double-buffers the state in LDS (one barrier per step); FP64 uses
  // one buffer and two barriers, keeping the block within 64 KiB of LDS.
  const int sl = p * CW + lane;
  T bprev = T(0), tnext = T(0);
  constexpr int kZBuf = sizeof(T) == 8 ? 1 : 2;
  if (Pi > 1) {
    const int PS = P * CW;
    T at = v0, ab = vl, ct = w0, cb = wl, rt = y0, rb = yl;
    int buf = 0;
    for (int d = 1; d < Pi; d <<= 1, buf ^= (kZBuf - 1)) {
      T* W = lds + buf * 6 * PS;
      W[sl] = at; W[PS + sl] = ab; W[2 * PS + sl] = ct;
      W[3 * PS + sl] = cb; W[4 * PS + sl] = rt; W[5 * PS + sl] = rb;
      __syncthreads();
      const bool hm = p - d >= 0, hp = p + d < Pi;
      const int sm = hm ? (p - d) * CW + lane : sl, spp = hp ? (p + d) * CW + lane : sl;
      // row b of segment p-d, row t of segment p+d
      const T abm = W[PS + sm], cbm = W[3 * PS + sm], rbm = W[5 * PS + sm];
      const T atp = W[spp], ctp = W[2 * PS + spp], rtp = W[4 * PS + spp];
      const T ab_m = hm ? abm : T(0), cb_m = hm ? cbm : T(0), rb_m = hm ? rbm : T(0);
      const T at_p = hp ? atp : T(0), ct_p = hp ? ctp : T(0), rt_p = hp ? rtp : T(0);
      if (kZBuf == 1) __syncthreads();   // every read done before the next writes
      const T m11 = T(1) - at * cb_m, m12 = -ct * at_p;
      const T m21 = -ab * cb_m, m22 = T(1) - cb * at_p;
      const T id = rcp(m11 * m22 - m12 * m21);
      const T qt = rt - at * rb_m - ct * rt_p, qb = rb - ab * rb_m - cb * rt_p;
      const T lt = -at * ab_m, lb = -ab * ab_m;     // couplings to b_{p-2d}
      const T ut = -ct * ct_p, ub = -cb * ct_p;     // couplings to t_{p+2d}
      at = id * (m22 * lt - m12 * lb);
      ab = id * (m11 * lb - m21 * lt);
      ct = id * (m22 * ut - m12 * ub);
      cb = id * (m11 * ub - m21 * ut);
      rt = id * (m22 * qt - m12 * qb);
      rb = id * (m11 * qb - m21 * qt);
    }
    T* W = lds + buf * 6 * PS;
    W[sl] = rt;
    W[PS + sl] = rb;
    __syncthreads();
    bprev = (p > 0) ? W[PS + (p - 1) * CW + lane] : T(0);
    tnext = (p < Pi - 1) ? W[(p + 1) * CW + lane] : T(0);
    __syncthreads();   // the group's LDS may be rewritten by its next call
  }
  // phase 3: back substitution of g - v bprev - w tnext (LEAN: v is
  // recomputed from rd by the forward pass's own operations first)
  auto backsub = [&](const int len) {
    if constexpr (LEAN) {
      if (p > 0) {
        T v = T(0);
#pragma unroll
        for (int q = 0; q < S; ++q) {
          if (q < len) {
            v = (q == 0) ? a_first * rdv[q] : -(A(q) * rdv[q]) * v;
            g[q] -= v * bprev;
          }
        }
      }
    }
    T x = T(0);
#pragma unroll
    for (int q = S - 1; q >= 0; --q) {
      if (q < len) {
        if constexpr (LEAN) {
          const T cq = E(q) * rdv[q];
          x = (q == len - 1) ? g[q] - cq * tnext : g[q] - cq * x;
        } else {
          const T gq = g[q] - vv[q] * bprev;
          x = (q == len - 1) ? gq - cp[q] * tnext : gq - cp[q] * x;
        }
        if (live) st_row<T>(u + (ks + q) * plane, bown, x);
      }
    }
  };
  if (sizeof(T) == 4 && len == S) backsub(S);
  else if (len > 0) backsub(len);
#undef A
#undef E
#undef D
}

// The LEAN FP32 form is built for 8 waves per SIMD (64 registers), so two
// 1024-lane blocks share a CU and one's loads overlap the other's solve.
// With H a workgroup has at most 8 waves and a CU holds one: up to 256
// registers.
template <class T, int S, int C, bool NBR, bool LEAN, bool H>
__global__ __launch_bounds__(H ? 512 : 1024, H ? 2 : ((LEAN && sizeof(T) == 4) ? 8 : 4)) void k_zline(
    const DevLevel* __restrict__ Lt, LaunchList ll, int parity, int zero_guess) {
  extern __shared__ unsigned char lds_raw[];
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  // the first sweep of a mixed cycle's finest level runs as two launches,
  // from zero (NBR false) and continuing the iterate (kZgWarm): one returns
  if (!NBR && (zero_guess & kZgDyn) && L.warm) return;
  if ((zero_guess & kZgWarm) && !L.warm) return;
  // threadIdx.y is constant across a wave (blockDim.x == 64): say so to the compiler
  const int seg = __builtin_amdgcn_readfirstlane(threadIdx.y);
  if constexpr (H) {
    const int cl = threadIdx.x & 31;
    zline_group<T, S, C, NBR, LEAN, true>(L, parity, zero_flags(L, zero_guess),
                                          (b - ll.prefix[t]) * 32 + cl, 2 * seg + (threadIdx.x >> 5),
                                          cl, 2 * blockDim.y, reinterpret_cast<T*>(lds_raw), true);
  } else {
    zline_group<T, S, C, NBR, LEAN, false>(L, parity, zero_flags(L, zero_guess),
                                           (b - ll.prefix[t]) * 64 + threadIdx.x, seg, threadIdx.x,
                                           blockDim.y, reinterpret_cast<T*>(lds_raw), true);
  }
}

template <class T> struct TailNz;
template <> struct TailNz<float> { static constexpr int value = kTailNz; };
template <> struct TailNz<double> { static constexpr int value = kTailNz64; };

// Short columns (nz <= NZ) inside the coarse tail: one thread per column,
// the whole column in registers, plain Thomas (the reference's line solve,
// mgMK/mg_relax.F90:541-595); column index col = j*hx + m.  Every load (the
// right-hand-side data and the z coefficients) is issued before the
// elimination; with EXACT (nz == NZ) the chains carry no predicates.
template <class T, int NZ, bool EXACT, class UP, class FP, class CP>
__device__ __forceinline__ void zline_column(const DevLevel& L, UP u, FP f, CP co, int parity,
                                             int zero_guess, int col) {
  const int hx = L.hx, ny = L.ny, nx = L.nx;
  const int nz = EXACT ? NZ : L.nz;
  // A compiler barrier: it keeps the solve in a scheduling region of its own,
  // apart from the caller's step dispatch.  Without it the coarse tail's
  // line-solve steps run about 1.3x slower on gfx950 (measured: 60 -> 46 us
  // per FP32 tail at 16^3..4^3).
  asm volatile("" ::: "memory");
  const int m = col % hx, j = col / hx;
  const int c = parity, oth = parity ^ 1;
  const int i = 2 * m + ((c + j) & 1), r = i & 1;
  int ml = m - 1 + r; if (ml < 0) ml = hx - 1;
  int mr = m + r;     if (mr >= hx) mr = 0;
  const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
  const auto cy = co + 3 * nx;
  const auto cz = cy + 3 * ny;
  const T wxl = co[i], wxr = co[nx + i], wys = cy[j], wyn = cy[ny + j];
  const T dxy = co[2 * nx + i] + cy[2 * ny + j];
  const int plane = ny * 2 * hx;
  const int own = cidx(m, j, 0, c, hx, ny);
  const int xo = cidx(0, j, 0, oth, hx, ny);
  const int so = cidx(m, jm, 0, oth, hx, ny);
  const int no = cidx(m, jp, 0, oth, hx, ny);
  const bool pz = L.pz != 0;
  const bool lag = pz && !(zero_guess & 2);
  const bool nbr = !(zero_guess & 1);
  const T ubot = u[(nz - 1) * plane + own], utop = u[own];
  const T lagb = lag ? cz[0] * ubot : T(0);
  const T lagt = lag ? cz[2 * nz - 1] * utop : T(0);
  // z coefficients held in registers for columns of up to 16 cells
  constexpr bool HOIST = NZ <= 16;
  constexpr int NH = HOIST ? NZ : 1;
  T g[NZ], za[NH], ze[NH], zd[NH];
#pragma unroll
  for (int q = 0; q < NZ; ++q) {
    const int k = EXACT ? q : min(q, nz - 1);     // rows clamped: every load in bounds
    const int kb = k * plane;
    const T fv = f[kb + own];
    const T xl = u[kb + xo + ml], xr = u[kb + xo + mr], ys = u[kb + so], yn = u[kb + no];
    if constexpr (HOIST) {
      za[q] = cz[k];
      ze[q] = cz[nz + k];
      zd[q] = cz[2 * nz + k];
    }
    T sn = wxl * xl + wxr * xr;
    sn += wys * ys;
    sn += wyn * yn;
    T v = nbr ? fv - sn : fv;
    v -= (q == 0) ? lagb : T(0);
    v -= (q == nz - 1) ? lagt : T(0);
    g[q] = v;
  }
  T cpa[NZ];
  T cp = T(0), yf = T(0);
#pragma unroll
  for (int q = 0; q < NZ; ++q) {
    if (EXACT || q < nz) {
      const T a = (q == 0) ? T(0) : (HOIST ? za[HOIST ? q : 0] : cz[q]);
      const T e = (q == nz - 1) ? T(0) : (HOIST ? ze[HOIST ? q : 0] : cz[nz + q]);
      const T rd = rcp(((HOIST ? zd[HOIST ? q : 0] : cz[2 * nz + q]) + dxy) - a * cp);
      yf = (g[q] - a * yf) * rd;
      cp = e * rd;
      g[q] = yf;
      cpa[q] = cp;
    }
  }
  T x = T(0);
#pragma unroll
  for (int q = NZ - 1; q >= 0; --q) {
    if (EXACT || q < nz) {
      x = g[q] - cpa[q] * x;
      u[q * plane + own] = x;
    }
  }
}

// Dispatch on the column length: an exact instantiation for full columns
// of NZ, NZ/2 and NZ/4 cells, the predicated one otherwise.
template <class T, int NZ, class UP, class FP, class CP>
__device__ __forceinline__ void zline_column_n(const DevLevel& L, UP u, FP f, CP co, int parity,
                                               int zero_guess, int col) {
  const int nz = L.nz;
  if (nz == NZ) zline_column<T, NZ, true>(L, u, f, co, parity, zero_guess, col);
  else if (NZ >= 8 && nz == NZ / 2) zline_column<T, (NZ >= 8 ? NZ / 2 : NZ), true>(L, u, f, co, parity, zero_guess, col);
  else if (NZ >= 16 && nz == NZ / 4) zline_column<T, (NZ >= 16 ? NZ / 4 : NZ), true>(L, u, f, co, parity, zero_guess, col);
  else zline_column<T, NZ, false>(L, u, f, co, parity, zero_guess, col);
}

// The standalone short-column sweep (tests of the tail's line solve).
template <class T, class UP, class FP, class CP>
__device__ __forceinline__ void zline_column_any(const DevLevel& L, UP u, FP f, CP co, int parity,
                                                 int zero_guess, int col) {
  if (L.nz <= 16 || TailNz<T>::value <= 16) zline_column_n<T, 16>(L, u, f, co, parity, zero_guess, col);
  else zline_column_n<T, TailNz<T>::value>(L, u, f, co, parity, zero_guess, col);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_zline_col(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                      int parity, int zero_guess) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e < L.hx * L.ny)
    zline_column_any<T>(L, gfield<T>(L.u), gfield<const T>((const void*)L.f), coefs<T>(L), parity,
                        zero_flags(L, zero_guess), e);
}

// ---------------------------------------------------------------------------
// Fused residual + volume-weighted restriction (mgMK/mg_restr.F90:124-164):
// element e is a coarse cell in natural (I fastest) order, so the fine reads
// of a wave are contiguous half-rows; the fine residual is never stored.
// FX, FY, FZ: the coarsening factors (1 or 2), compile-time so that the
// children loop unrolls and its loads are issued together.
template <class T, int FX, int FY, int FZ, class UP, class FP, class CP, class VP, class CFP>
__device__ __forceinline__ void restrict_elem(const DevLevel& F, UP u, FP f, CP co, VP vol,
                                              const DevLevel& C, CFP fc, int e) {
  // nx = 2 hx: e / nx = (e >> 1) / hx
  const int JK = gm_div(e >> 1, C.gmhx, C.gshx), I = e - JK * C.nx;
  const int K = gm_div(JK, C.gmny, C.gsny), J = JK - K * C.ny;
  const auto vx = vol;
  const auto vy = vx + F.nx;
  const auto vz = vy + F.ny;
  T acc = T(0), vs = T(0);
#pragma unroll
  for (int cz = 0; cz < FZ; ++cz) {
    const int k = K * FZ + cz;
#pragma unroll
    for (int cy = 0; cy < FY; ++cy) {
      const int j = J * FY + cy;
#pragma unroll
      for (int cx = 0; cx < FX; ++cx) {
        const int i = I * FX + cx;
        const Nb n = neighbours(F, i, j, k);
        const T s = offsum<T>(co, F, i, j, k, u[n.xm], u[n.xp], u[n.ym], u[n.yp], u[n.zm], u[n.zp]);
        const T rr = f[n.c] - (s + diag_at<T>(co, F, i, j, k) * u[n.c]);
        const T v = T(vz[k] * vy[j] * vx[i]);
        acc += v * rr;
        vs += v;
      }
    }
  }
  fc[cidx(I >> 1, J, K, (I + J) & 1, C.hx, C.ny)] = acc / vs;
}

template <class T, class UP, class FP, class CP, class VP, class CFP>
__device__ __forceinline__ void restrict_any(const DevLevel& F, UP u, FP f, CP co, VP vol,
                                             const DevLevel& C, CFP fc, const DevXfer& x, int e) {
  const int code = (x.fx - 1) * 4 + (x.fy - 1) * 2 + (x.fz - 1);
  switch (code) {
    case 7: restrict_elem<T, 2, 2, 2>(F, u, f, co, vol, C, fc, e); break;
    case 6: restrict_elem<T, 2, 2, 1>(F, u, f, co, vol, C, fc, e); break;
    case 5: restrict_elem<T, 2, 1, 2>(F, u, f, co, vol, C, fc, e); break;
    case 4: restrict_elem<T, 2, 1, 1>(F, u, f, co, vol, C, fc, e); break;
    case 3: restrict_elem<T, 1, 2, 2>(F, u, f, co, vol, C, fc, e); break;
    case 2: restrict_elem<T, 1, 2, 1>(F, u, f, co, vol, C, fc, e); break;
    case 1: restrict_elem<T, 1, 1, 2>(F, u, f, co, vol, C, fc, e); break;
    default: restrict_elem<T, 1, 1, 1>(F, u, f, co, vol, C, fc, e); break;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_restrict(const DevLevel* __restrict__ Lt,
                                                     const DevXfer* __restrict__ X, LaunchList ll) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevXfer x = X[ll.items[t]];
  const DevLevel& C = Lt[x.coarse];
  const DevLevel& F = Lt[x.fine];
  if (level_off(F)) return;
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e < C.ncell)
    restrict_any<T>(F, gfield<const T>((const void*)F.u), gfield<const T>((const void*)F.f),
                    coefs<T>(F), gp(F.vol), C, gfield<T>(C.f), x, e);
}

// Residual + restriction of a transfer with factor 2 in every direction and
// fine sizes exactly twice the coarse ones, marching in z.  A thread owns
// coarse column (I, J) over a slab of restrict222_k coarse planes: fine
// columns i in {2I, 2I+1}, j in {2J, 2J+1}.  It walks their fine planes
// upward with the centre values of the planes below, at and above in
// registers; per coarse cell it loads the centres of the two planes above,
// b of its two planes and the eight in-plane neighbours outside its 2x2
// block (16 loads per fine plane, against 32 for restrict_elem).  The
// neighbours inside the block are its own centres, at the indices
// neighbours() gives, and the outside ones take its wraps.  Each fine
// residual is formed with offsum/diag_at as in restrict_elem, and the coarse
// value sums its eight in the same order (z, then y, then x), so the result
// is restrict_elem's.
//
// SK (DevXfer::rskip): after the fine level's pre-smoothing, whose last
// half-sweep solved the equations of its colour-1 cells from their
// neighbours, those cells' residuals are zero but for rounding; they are
// taken as zero, and the loads only they need (their b and their outside
// neighbours) are not in the code.  Colour 1 of cell (2I+cx, 2J+cy, 2K+h):
// (cx+cy) odd for the z-line sweep, (cx+cy+h) odd for the point sweep.
template <int SK>
__device__ __forceinline__ constexpr bool restrict_keeps(int cx, int cy, int h) {
  return SK == kXferSkipLine ? ((cx + cy) & 1) == 0
       : SK == kXferSkipPoint ? ((cx + cy + h) & 1) == 0 : true;
}

template <class T, int SK>
__device__ __forceinline__ void restrict222_cols(const DevLevel& F, const DevLevel& C, int I, int J,
                                                 int K0, int K1) {
  const auto u = gfield<const T>((const void*)F.u);
  const auto f = gfield<const T>((const void*)F.f);
  const auto co = coefs<T>(F);
  const auto vx = gp(F.vol);
  const auto vy = vx + F.nx;
  const auto vz = vy + F.ny;
  const int hx = F.hx, ny = F.ny, nz = F.nz;
  const int plane = 2 * hx * ny;
  // in-plane offsets: own cells o[cy][cx], outside neighbours
  int o[2][2], oxm[2], oxp[2], oym[2], oyp[2];
#pragma unroll
  for (int cy = 0; cy < 2; ++cy)
#pragma unroll
    for (int cx = 0; cx < 2; ++cx) {
      const int i = 2 * I + cx, j = 2 * J + cy;
      o[cy][cx] = cidx(i >> 1, j, 0, (i + j) & 1, hx, ny);
    }
#pragma unroll
  for (int cy = 0; cy < 2; ++cy) {
    const int j = 2 * J + cy;
    // x- of i = 2I (r = 0: m - 1, wrapped) and x+ of i = 2I+1 (r = 1: m + 1, wrapped)
    const int ml = I == 0 ? hx - 1 : I - 1, mr = I + 1 >= hx ? 0 : I + 1;
    oxm[cy] = cidx(0, j, 0, ((2 * I + j) & 1) ^ 1, hx, ny) + ml;
    oxp[cy] = cidx(0, j, 0, ((2 * I + 1 + j) & 1) ^ 1, hx, ny) + mr;
  }
#pragma unroll
  for (int cx = 0; cx < 2; ++cx) {
    const int i = 2 * I + cx;
    const int j0 = 2 * J, j1 = 2 * J + 1;
    const int jm = j0 == 0 ? ny - 1 : j0 - 1, jp = j1 == ny - 1 ? 0 : j1 + 1;
    oym[cx] = cidx(i >> 1, jm, 0, ((i + j0) & 1) ^ 1, hx, ny);
    oyp[cx] = cidx(i >> 1, jp, 0, ((i + j1) & 1) ^ 1, hx, ny);
  }
  auto ld4 = [&](int k, T (&v)[2][2]) {
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
      for (int cx = 0; cx < 2; ++cx) v[cy][cx] = u[k * plane + o[cy][cx]];
  };
  T um[2][2], uc[2][2];
  const int k00 = 2 * K0;
  ld4(k00 == 0 ? nz - 1 : k00 - 1, um);
  ld4(k00, uc);
  for (int K = K0; K < K1; ++K) {
    const int ka = 2 * K, kb = ka + 1, kc = kb == nz - 1 ? 0 : kb + 1;
    // every load of the coarse cell first
    T ua[2][2], ub[2][2], fv[2][2][2], xm[2][2], xp[2][2], ym[2][2], yp[2][2];
    ld4(kb, ua);
    ld4(kc, ub);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int kk = ka + h;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        // fv[h][cy][cx]; xm/xp[h][cy] serve cells (0, cy)/(1, cy), ym/yp[h][cx]
        // cells (cx, 0)/(cx, 1)
        fv[h][q][0] = restrict_keeps<SK>(0, q, h) ? f[kk * plane + o[q][0]] : T(0);
        fv[h][q][1] = restrict_keeps<SK>(1, q, h) ? f[kk * plane + o[q][1]] : T(0);
        xm[h][q] = restrict_keeps<SK>(0, q, h) ? u[kk * plane + oxm[q]] : T(0);
        xp[h][q] = restrict_keeps<SK>(1, q, h) ? u[kk * plane + oxp[q]] : T(0);
        ym[h][q] = restrict_keeps<SK>(q, 0, h) ? u[kk * plane + oym[q]] : T(0);
        yp[h][q] = restrict_keeps<SK>(q, 1, h) ? u[kk * plane + oyp[q]] : T(0);
      }
    }
    T acc = T(0), vs = T(0);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = ka + h;
      const T (&zm)[2][2] = h ? uc : um;
      const T (&cc)[2][2] = h ? ua : uc;
      const T (&zp)[2][2] = h ? ub : ua;
#pragma unroll
      for (int cy = 0; cy < 2; ++cy) {
        const int j = 2 * J + cy;
#pragma unroll
        for (int cx = 0; cx < 2; ++cx) {
          const int i = 2 * I + cx;
          const T v = T(vz[k] * vy[j] * vx[i]);
          if (restrict_keeps<SK>(cx, cy, h)) {
            const T uxm = cx ? cc[cy][0] : xm[h][cy];
            const T uxp = cx ? xp[h][cy] : cc[cy][1];
            const T uym = cy ? cc[0][cx] : ym[h][cx];
            const T uyp = cy ? yp[h][cx] : cc[1][cx];
            const T s = offsum<T>(co, F, i, j, k, uxm, uxp, uym, uyp, zm[cy][cx], zp[cy][cx]);
            const T rr = fv[h][cy][cx] - (s + diag_at<T>(co, F, i, j, k) * cc[cy][cx]);
            acc += v * rr;
          }
          vs += v;
        }
      }
    }
    gfield<T>(C.f)[cidx(I >> 1, J, K, (I + J) & 1, C.hx, C.ny)] = acc / vs;
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
      for (int cx = 0; cx < 2; ++cx) {
        um[cy][cx] = ua[cy][cx];
        uc[cy][cx] = ub[cy][cx];
      }
  }
}

// use_skip: take the transfer's rskip rule (the V-cycle's restriction after
// pre-smoothing); 0 forms every residual.
template <class T>
__global__ __launch_bounds__(kBlock) void k_restrict222(const DevLevel* __restrict__ Lt,
                                                        const DevXfer* __restrict__ X,
                                                        LaunchList ll, int use_skip) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevXfer x = X[ll.items[t]];
  const DevLevel& C = Lt[x.coarse];
  const DevLevel& F = Lt[x.fine];
  if (level_off(F)) return;
  const int cols = C.nx * C.ny;
  const int ncb = (cols + kBlock - 1) / kBlock;
  const int lb = b - ll.prefix[t];
  const int bk = lb / ncb, col = (lb - bk * ncb) * kBlock + threadIdx.x;
  if (col >= cols) return;
  const int J = gm_div(col >> 1, C.gmhx, C.gshx), I = col - J * C.nx;   // col / nx, nx = 2 hx
  const int kr = restrict222_k(C);
  const int K0 = bk * kr, K1 = min(K0 + kr, C.nz);
  const int sk = use_skip ? x.rskip : kXferSkipNone;
  if (sk == kXferSkipLine) restrict222_cols<T, kXferSkipLine>(F, C, I, J, K0, K1);
  else if (sk == kXferSkipPoint) restrict222_cols<T, kXferSkipPoint>(F, C, I, J, K0, K1);
  else restrict222_cols<T, kXferSkipNone>(F, C, I, J, K0, K1);
}

// ---------------------------------------------------------------------------
// Trilinear prolongation-add (mgMK/mg_cor.F90:102-159): fine u += sum of
// 27,9,9,3,9,3,3,1 /64 of the eight nearest coarse cells; coarse ghosts are
// periodic wraps, Neumann copies or Dirichlet zeros.
//
// One thread per dual cell (prolong_duals): the fine cells between coarse
// centres I and I+1 in each direction share the same eight coarse values,
// which are loaded once.  Per factor-2 direction the dual d covers fine
// cells 2I+1 (weights 3/4 on I, 1/4 on I+1) and 2I+2 (3/4 on I+1, 1/4 on I)
// with I = d (periodic, I+1 and 2I+2 wrapped) or I = d-1 (walls: the half
// duals d = 0 and d = nc hold one fine cell; the missing coarse value is
// the ghost -- the neighbour's value for Neumann, zero for Dirichlet).  A
// factor-1 direction maps cell to cell with weight 1.  Each fine value is
// summed nearest-first: x, then y, then z.
//
// skip (DevXfer::skip, kXferSkip*): the cells that the fine level's first
// post-smoothing half-sweep rewrites from the other colour and b alone are
// neither read nor written -- their correction would be overwritten unread,
// so the result of that half-sweep, and of everything after it, is the same.
struct PDir {
  int fi0, fi1;     // fine indices (fine cell 0 is nearest to coarse centre 0)
  bool fv0, fv1;    // fine cell exists
  int ci0, ci1;     // coarse indices
  bool cz0, cz1;    // coarse value is a Dirichlet zero ghost
  __device__ int fi(int h) const { return h ? fi1 : fi0; }
  __device__ bool fv(int h) const { return h ? fv1 : fv0; }
  __device__ int ci(int h) const { return h ? ci1 : ci0; }
  __device__ bool cz(int h) const { return h ? cz1 : cz0; }
};

template <int F>
__device__ __forceinline__ PDir pdir(int d, int nc, int per, int bc0, int bc1) {
  PDir r;
  if (F == 1) {
    r.fi0 = r.fi1 = d;
    r.fv0 = true; r.fv1 = false;
    r.ci0 = r.ci1 = d;
    r.cz0 = r.cz1 = false;
    return r;
  }
  const int nf = 2 * nc;
  if (per) {
    const int I = d;
    r.fi0 = 2 * I + 1;
    r.fi1 = (2 * I + 2 == nf) ? 0 : 2 * I + 2;
    r.fv0 = r.fv1 = true;
    r.ci0 = I;
    r.ci1 = (d + 1 == nc) ? 0 : d + 1;
    r.cz0 = r.cz1 = false;
  } else {
    const int I = d - 1;
    r.fi0 = 2 * I + 1;
    r.fi1 = 2 * I + 2;
    r.fv0 = I >= 0;
    r.fv1 = I + 1 < nc;
    // a missing coarse centre takes the ghost rule of its wall
    r.ci0 = (I >= 0) ? I : 0;
    r.ci1 = (I + 1 < nc) ? I + 1 : nc - 1;
    r.cz0 = (I < 0) && bc0 == 2;
    r.cz1 = (I + 1 >= nc) && bc1 == 2;
  }
  return r;
}

// Whether fine cell (h, q, a) of a 2:2:2 dual is corrected under skip rule
// SK.  Fine cell 0 of a factor-2 direction has an odd index (2I+1) and cell
// 1 an even one (2I+2, or 0 wrapped), so the colour of each cell is fixed by
// its position in the dual: the skipped cells are known at compile time,
// and their loads and stores are not in the code.
template <int SK>
__device__ __forceinline__ constexpr bool prolong_keeps(int h, int q, int a) {
  return SK == kXferSkipNone ? true
       : (SK == kXferSkipLine || SK == kXferSkipLineWrap) ? (((1 - h) + (1 - q)) & 1) != 0
       : (((1 - h) + (1 - q) + (1 - a)) & 1) != 0;
}

template <class T, int FX, int FY, int FZ, int SK = kXferSkipNone, class UP, class CUP>
__device__ __forceinline__ void prolong_dual(const DevLevel& F, UP u, const DevLevel& C, CUP uc,
                                             const DevXfer& x, int e) {
  static_assert(SK == kXferSkipNone || (FX == 2 && FY == 2 && FZ == 2), "skip: 2:2:2 duals only");
  constexpr int NX = FX == 2 ? 2 : 1, NY = FY == 2 ? 2 : 1, NZ = FZ == 2 ? 2 : 1;
  const int ndx = prolong_duals_dir(C.nx, FX, F.px), ndy = prolong_duals_dir(C.ny, FY, F.py);
  // e = (dz * ndy + dy) * ndx + dx, divided by the transfer's invariant divisors
  const int dyz = gm_div(e, x.pmx, x.psx), dx = e - dyz * ndx;
  const int dz = gm_div(dyz, x.pmy, x.psy), dy = dyz - dz * ndy;
  const PDir X = pdir<FX>(dx, C.nx, F.px, F.bcx0, F.bcx1);
  const PDir Y = pdir<FY>(dy, C.ny, F.py, F.bcy0, F.bcy1);
  const PDir Z = pdir<FZ>(dz, C.nz, F.pz, F.bcz0, F.bcz1);
  // every load first: the coarse values v[z][y][x] (index 0: the near
  // centre of fine cell 0) and the fine cells' old values (a missing cell's
  // index is clamped to its partner; it is not stored; a skipped cell is
  // neither loaded nor stored)
  T v[NZ][NY][NX], fold[NZ][NY][NX];
  int fid[NZ][NY][NX];
  bool keep[NZ][NY][NX];
  // with periodic z the z-line sweep's lagged wrap reads its own colour at
  // rows 0 and nz-1: both are cells of the last z dual
  const bool zwrap = SK == kXferSkipLineWrap && Z.fi(NZ - 1) == 0;
#pragma unroll
  for (int a = 0; a < NZ; ++a)
#pragma unroll
    for (int q = 0; q < NY; ++q)
#pragma unroll
      for (int h = 0; h < NX; ++h) {
        const int I = X.ci(h), J = Y.ci(q), K = Z.ci(a);
        // a Dirichlet ghost contributes zero: a 0/1 factor, not a branch,
        // so that all loads stay in one batch (the same 0*value as a zero
        // interpolation weight)
        const T c = uc[cidx(I >> 1, J, K, (I + J) & 1, C.hx, C.ny)];
        v[a][q][h] = c * ((X.cz(h) || Y.cz(q) || Z.cz(a)) ? T(0) : T(1));
        const int i = X.fv(h) ? X.fi(h) : X.fi(h ^ (NX - 1));
        const int j = Y.fv(q) ? Y.fi(q) : Y.fi(q ^ (NY - 1));
        const int k = Z.fv(a) ? Z.fi(a) : Z.fi(a ^ (NZ - 1));
        fid[a][q][h] = cidx(i >> 1, j, k, (i + j) & 1, F.hx, F.ny);
        keep[a][q][h] = X.fv(h) && Y.fv(q) && Z.fv(a) && (prolong_keeps<SK>(h, q, a) || zwrap);
        fold[a][q][h] = (prolong_keeps<SK>(h, q, a) || zwrap) ? u[fid[a][q][h]] : T(0);
      }
  // keep the old values' loads in the batch above (an empty register
  // constraint: otherwise a load is sunk into its predicated store)
#pragma unroll
  for (int a = 0; a < NZ; ++a)
#pragma unroll
    for (int q = 0; q < NY; ++q)
#pragma unroll
      for (int h = 0; h < NX; ++h) asm volatile("" : "+v"(fold[a][q][h]));
  // separable interpolation, nearest first: along x, then y, then z
  T ix[NZ][NY][NX], iy[NZ][NY][NX];
#pragma unroll
  for (int a = 0; a < NZ; ++a)
#pragma unroll
    for (int q = 0; q < NY; ++q)
#pragma unroll
      for (int h = 0; h < NX; ++h)
        ix[a][q][h] = (NX == 2) ? T(0.75) * v[a][q][h] + T(0.25) * v[a][q][h ^ (NX - 1)] : v[a][q][h];
#pragma unroll
  for (int a = 0; a < NZ; ++a)
#pragma unroll
    for (int q = 0; q < NY; ++q)
#pragma unroll
      for (int h = 0; h < NX; ++h)
        iy[a][q][h] = (NY == 2) ? T(0.75) * ix[a][q][h] + T(0.25) * ix[a][q ^ (NY - 1)][h] : ix[a][q][h];
#pragma unroll
  for (int a = 0; a < NZ; ++a)
#pragma unroll
    for (int q = 0; q < NY; ++q)
#pragma unroll
      for (int h = 0; h < NX; ++h) {
        const T add = (NZ == 2) ? T(0.75) * iy[a][q][h] + T(0.25) * iy[a ^ (NZ - 1)][q][h] : iy[a][q][h];
        if (keep[a][q][h]) u[fid[a][q][h]] = fold[a][q][h] + add;
      }
}

template <class T, class UP, class CUP>
__device__ __forceinline__ void prolong_any(const DevLevel& F, UP u, const DevLevel& C, CUP uc,
                                            const DevXfer& x, int e, int skip = kXferSkipNone) {
  const int code = (x.fx - 1) * 4 + (x.fy - 1) * 2 + (x.fz - 1);
  switch (code) {
    case 7:
      // the skip rules apply to 2:2:2 transfers (DevXfer::skip)
      if (skip == kXferSkipLine) prolong_dual<T, 2, 2, 2, kXferSkipLine>(F, u, C, uc, x, e);
      else if (skip == kXferSkipLineWrap) prolong_dual<T, 2, 2, 2, kXferSkipLineWrap>(F, u, C, uc, x, e);
      else if (skip == kXferSkipPoint) prolong_dual<T, 2, 2, 2, kXferSkipPoint>(F, u, C, uc, x, e);
      else prolong_dual<T, 2, 2, 2>(F, u, C, uc, x, e);
      break;
    case 6: prolong_dual<T, 2, 2, 1>(F, u, C, uc, x, e); break;
    case 5: prolong_dual<T, 2, 1, 2>(F, u, C, uc, x, e); break;
    case 4: prolong_dual<T, 2, 1, 1>(F, u, C, uc, x, e); break;
    case 3: prolong_dual<T, 1, 2, 2>(F, u, C, uc, x, e); break;
    case 2: prolong_dual<T, 1, 2, 1>(F, u, C, uc, x, e); break;
    case 1: prolong_dual<T, 1, 1, 2>(F, u, C, uc, x, e); break;
    default: prolong_dual<T, 1, 1, 1>(F, u, C, uc, x, e); break;
  }
}

// use_skip: leave the cells of the transfer's skip rule (the V-cycle's
// prolongation, followed by the fine level's post-smoothing); 0 corrects
// every cell.
template <class T>
__global__ __launch_bounds__(kBlock) void k_prolong(const DevLevel* __restrict__ Lt,
                                                    const DevXfer* __restrict__ X, LaunchList ll,
                                                    int use_skip) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevXfer x = X[ll.items[t]];
  const DevLevel& F = Lt[x.fine];
  const DevLevel& C = Lt[x.coarse];
  if (level_off(F)) return;
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e < prolong_duals(F, C, x))
    prolong_any<T>(F, gfield<T>(F.u), C, gfield<const T>((const void*)C.u), x, e,
                   use_skip ? x.skip : kXferSkipNone);
}

// ---------------------------------------------------------------------------
// Coarsest level: u = Minv f with a precomputed FP64 (pseudo-)inverse
// stored transposed (row r of the product reads column r: coalesced), FP64
// accumulation in eight independent partial sums.
template <class T, class UP, class FP, class MP>
__device__ __forceinline__ void coarse_row(const DevLevel& L, UP u, FP f, MP Mt, int r) {
  const int n = L.ncell;
  const auto M = Mt + r;
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int s = 0;
  // 16 matrix loads in flight per row: a large coarsest level (an 8 MB
  // inverse at 1024 cells) is read by one workgroup, and with fewer loads
  // per thread its rate is set by the memory latency.  Element s+q adds to
  // acc[(s+q) & 7] in increasing s, as in the 8-wide loop below.
  for (; s + 16 <= n; s += 16) {
    double m[16];
    T fv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      m[q] = M[(s + q) * n];
      fv[q] = f[s + q];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q & 7] += m[q] * (double)fv[q];
  }
  for (; s + 8 <= n; s += 8) {
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] += M[(s + q) * n] * (double)f[s + q];
  }
  for (; s < n; ++s) acc[0] += M[s * n] * (double)f[s];
  const double sum = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  u[r] = (T)sum;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_coarse(const DevLevel* __restrict__ Lt, LaunchList ll) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  const int r = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (r < L.ncell) coarse_row<T>(L, gfield<T>(L.u), gfield<const T>((const void*)L.f), gp(L.minvT), r);
}

// ---------------------------------------------------------------------------
// Coarse tail: the whole sub-V-cycle below a level (pre-smoothing and
// residual restriction down to the coarsest level, its exact solve, then
// prolongation and post-smoothing back up -- the order of
// mgMK/mg_cycle.F90) for one instance in one workgroup of 64*kTailWaves
// threads, with workgroup barriers between the steps instead of dependent
// launches.  The tail levels' u, f, coefficient tables and cell widths live
// in LDS (the first level's f, and u when it is an input, are copied in,
// its u copied out); the element routines are the standalone kernels' ones
// instantiated with LDS pointers.  One workgroup per item (DevTail).
// Global -> LDS copy of n values (converted to D) by `nthr` threads: each
// thread issues its (up to) 8 loads before any store, so a copy costs one
// memory round trip per 8*nthr values instead of one per nthr.
template <class D, class SRC, class DST>
__device__ __forceinline__ void stage_copy(DST dst, SRC src, int n, int t, int nthr) {
  for (int e0 = t; e0 < n; e0 += 8 * nthr) {
    D v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = (D)src[min(e0 + q * nthr, n - 1)];
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (e0 + q * nthr < n) dst[e0 + q * nthr] = v[q];
  }
}

template <class T, int NZ, bool ALL222>
// NZ = 0 is the point-smoother variant: no line-solve code (whose column
// registers would cap the workgroup at 512 lanes), and 1024 lanes.
// phase 0: the whole step list; 1: the steps before the coarse solve, then
// every level's u and f saved to its global arrays; 2: those restored, the
// steps after the coarse solve (run by k_coarse_wide in between).
__global__ __launch_bounds__(NZ == 0 ? 64 * kTailWavesPoint : 64 * kTailWaves) void k_tail(const DevLevel* __restrict__ Lt,
                                                          const DevXfer* __restrict__ X,
                                                          const DevTail* __restrict__ Tt,
                                                          LaunchList ll, int phase) {
  extern __shared__ unsigned char lds_raw[];
  __shared__ DevLevel Ls[kTailMaxLevels];
  const DevTail& tl = Tt[ll.items[blockIdx.x]];
  if (level_off(Lt[tl.lev[0]])) return;
  const int tid = threadIdx.y * 64 + threadIdx.x, nthr = 64 * blockDim.y;
  const int nl = tl.nlev;
  const sp<kS, unsigned char> base = (sp<kS, unsigned char>)lds_raw;
  auto U = [&](int d) { return (sp<kS, T>)(base + tl.lds_u[d]); };
  auto Fv = [&](int d) { return (sp<kS, T>)(base + tl.lds_f[d]); };
  auto CO = [&](int d) { return (sp<kS, const T>)(base + tl.lds_co[d]); };
  auto VOL = [&](int d) { return (sp<kS, const double>)(base + tl.lds_vol[d]); };
  if (tid < nl) Ls[tid] = Lt[tl.lev[tid]];
  // stage the tables of every tail level (one wave per level, in parallel),
  // the coarsest inverse when it fits, and the first level's fields
  const int wv = threadIdx.y, ln = threadIdx.x;
  for (int d = wv; d < nl; d += blockDim.y) {
    const DevLevel& G = Lt[tl.lev[d]];
    const int nco = 3 * (G.nx + G.ny + G.nz), nv = G.nx + G.ny + G.nz;
    stage_copy<T>((sp<kS, T>)(base + tl.lds_co[d]), coefs<T>(G), nco, ln, 64);
    if (d + 1 < nl) stage_copy<double>((sp<kS, double>)(base + tl.lds_vol[d]), gp(G.vol), nv, ln, 64);
  }
  if (tl.lds_minv >= 0) {
    const DevLevel& G = Lt[tl.lev[nl - 1]];
    stage_copy<T>((sp<kS, T>)(base + tl.lds_minv), gp(G.minvT), G.ncell * G.ncell, tid, nthr);
  }
  const DevLevel& G0 = Lt[tl.lev[0]];
  if (phase == 2) {
    for (int d = 0; d < nl; ++d) {
      const DevLevel& G = Lt[tl.lev[d]];
      stage_copy<T>(U(d), gfield<const T>((const void*)G.u), G.ncell, tid, nthr);
      if (d + 1 < nl) stage_copy<T>(Fv(d), gfield<const T>((const void*)G.f), G.ncell, tid, nthr);
    }
  } else {
    stage_copy<T>(Fv(0), gfield<const T>((const void*)G0.f), G0.ncell, tid, nthr);
    if (tl.copy_u_in) stage_copy<T>(U(0), gfield<const T>((const void*)G0.u), G0.ncell, tid, nthr);
  }
  __syncthreads();
  // the coarse solve's place in the step list (phases 1 and 2 stop or start there)
  int ci = 0;
  while (ci < tl.nsteps && (tl.steps[ci] & 7) != kStepCoarse) ++ci;
  const int st0 = phase == 2 ? ci + 1 : 0, st1 = phase == 1 ? ci : tl.nsteps;
  // one routine per step type, so that each element routine is instantiated
  // once (the kernel stays small enough for the instruction cache)
  for (int st = st0; st < st1; ++st) {
    const int code = tl.steps[st];
    const int type = code & 7, d = (code >> 3) & 15, par = (code >> 7) & 1, zg = (code >> 8) & 3;
    const DevLevel& L = Ls[d];
    if (type == kStepHalf) {
      if (NZ > 0 && tl.zline) {
        if constexpr (NZ > 0) {
          const int n = L.hx * L.ny;
          for (int e = tid; e < n; e += nthr) zline_column_n<T, NZ>(L, U(d), Fv(d), CO(d), par, zg, e);
        }
      } else {
        const int n = L.hx * L.ny * L.nz;
        for (int e = tid; e < n; e += nthr) rb_point_small<T>(L, U(d), Fv(d), CO(d), e, par, zg);
      }
    } else if (type == kStepRestrict) {
      const DevXfer x = X[tl.xf[d]];
      const DevLevel& C = Ls[d + 1];
      for (int e = tid; e < C.ncell; e += nthr) {
        if (ALL222) restrict_elem<T, 2, 2, 2>(L, U(d), Fv(d), CO(d), VOL(d), C, Fv(d + 1), e);
        else restrict_any<T>(L, U(d), Fv(d), CO(d), VOL(d), C, Fv(d + 1), x, e);
      }
    } else if (type == kStepCoarse) {
      if (tl.lds_minv >= 0) {
        const sp<kS, const T> M = (sp<kS, const T>)(base + tl.lds_minv);
        for (int r = tid; r < L.ncell; r += nthr) coarse_row<T>(L, U(d), Fv(d), M, r);
      } else {
        for (int r = tid; r < L.ncell; r += nthr) coarse_row<T>(L, U(d), Fv(d), gp(L.minvT), r);
      }
    } else if (type == kStepProlong) {
      const DevXfer x = X[tl.xf[d]];
      const DevLevel& C = Ls[d + 1];
      const int nd = prolong_duals(L, C, x);
      for (int e = tid; e < nd; e += nthr) {
        if (ALL222) prolong_dual<T, 2, 2, 2>(L, U(d), C, U(d + 1), x, e);
        else prolong_any<T>(L, U(d), C, U(d + 1), x, e);
      }
    } else {
      const sp<kS, T> u = U(d);
      for (int e = tid; e < L.ncell; e += nthr) u[e] = T(0);
    }
    __syncthreads();
  }
  if (phase == 1) {
    for (int d = 0; d < nl; ++d) {
      const DevLevel& G = Lt[tl.lev[d]];
      const sp<kG, T> gu = gfield<T>(G.u), gf = gfield<T>(G.f);
      const sp<kS, T> lu = U(d), lf = Fv(d);
      for (int e = tid; e < G.ncell; e += nthr) {
        gu[e] = lu[e];
        gf[e] = lf[e];
      }
    }
    return;
  }
  const sp<kG, T> gu = gfield<T>(G0.u);
  const sp<kS, T> lu = U(0);
  for (int e = tid; e < G0.ncell; e += nthr) gu[e] = lu[e];
}

// Coarsest level of a split tail: u = M f with the row-major (pseudo-)inverse,
// one wave per row (FP64 partial sums over the lane's columns in a fixed
// order, then a fixed butterfly), so a large inverse is read by the whole
// device rather than by one workgroup.
template <class T>
__global__ __launch_bounds__(kBlock) void k_coarse_wide(const DevLevel* __restrict__ Lt, LaunchList ll) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  const int n = L.ncell;
  const int r = (b - ll.prefix[t]) * (kBlock / 64) + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= n) return;
  const sp<kG, const double> M = gp(L.minv) + (size_t)r * n;
  const sp<kG, const T> f = gfield<const T>((const void*)L.f);
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int s = lane;
  for (; s + 7 * 64 < n; s += 8 * 64) {
    double m[8], fv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      m[q] = M[s + q * 64];
      fv[q] = (double)f[s + q * 64];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] += m[q] * fv[q];
  }
  for (int q = 0; s < n; s += 64, ++q) acc[q & 7] += M[s] * (double)f[s];
  double sum = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  sum = wave_sum(sum);
  if (lane == 0) gfield<T>(L.u)[r] = (T)sum;
}

// ---------------------------------------------------------------------------
// FP64 finest-level kernels.
// Each thread handles kFinePer cells e = e0 + s*kBlock (coalesced), keeps
// running maxima, and the block writes one partial per value; the finalize
// kernel folds the partials per item.
template <int NV, class F>
__device__ __forceinline__ void fine_loop(const DevLevel& L, int lb, double* __restrict__ partial,
                                          int b, F&& body) {
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  const int e0 = lb * kBlock * kFinePer + threadIdx.x;
#pragma unroll
  for (int s = 0; s < kFinePer; ++s) {
    const int e = e0 + s * kBlock;
    if (e < L.ncell) body(e, v);
  }
  block_reduce_store<NV, false>(v, partial + (size_t)b * NV);
}

// {max|u - u2|, max|u|}
__global__ __launch_bounds__(kBlock) void k_change_max(const DevLevel* __restrict__ Lt,
                                                       LaunchList ll, double* __restrict__ partial) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  const auto u = gfield<const double>((const void*)L.u);
  const auto u2 = gfield<const double>((const void*)L.u2);
  fine_loop<2>(L, b - ll.prefix[t], partial, b, [&](int e, double* v) {
    v[0] = amax(v[0], u[e] - u2[e]);
    v[1] = amax(v[1], u[e]);
  });
}

// FP64 finest-level stencil kernels, marching in z.  A thread owns one
// position col = (j*2 + colour)*hx + m of a plane and walks kMarchK planes,
// keeping the centre values at k-1, k, k+1 in registers: every value of the
// iterate is read from memory once per slab (plus one halo plane at each
// slab end), the four in-plane neighbours come from the cache, and the
// index arithmetic is done once per thread instead of once per cell.
//   kResMax:        {max|b - A u|, max|b|}
//   kDefect32:      r = b - A u -> FP32 right-hand side of the correction
//                   level; {max|r|, max|b|}
//   kCorrDefect32:  p_out = p_in + e, r = b - A p_out -> FP32; p_in is never
//                   written, so neighbouring threads read consistent values;
//                   {max|r|, max|e|, max|p_out|}
// The stencil sums are ordered as in offsum/diag_at.
enum { kResMax = 0, kDefect32 = 1, kCorrDefect32 = 2 };
#ifndef AMG_FINE_DEPTH
#define AMG_FINE_DEPTH 2
#endif
constexpr int kFineDepth = AMG_FINE_DEPTH;

// The marching loop of one thread.  Every field pointer is a __restrict__
// parameter (the outputs never alias the inputs), and the loop is
// software-pipelined by hand: the loads of plane k+1 are issued before plane
// k is computed and stored, so a thread keeps two planes of loads in flight.
// In mixed mode the centre column's p and e are carried separately (e for
// the correction norm).
template <int MODE, int NV, bool SHIFT>
__device__ __forceinline__ void fine_march(sp<kG, const double> __restrict__ u,
                                           sp<kG, const float> __restrict__ ev,
                                           sp<kG, double> __restrict__ f, double cs,
                                           sp<kG, double> __restrict__ pout,
                                           sp<kG, float> __restrict__ r32,
                                           sp<kG, const double> __restrict__ cz, int nz, int plane,
                                           int col, int oxm, int oxp, int oym, int oyp, double wxm,
                                           double wxp, double wym, double wyp, double dxy_y,
                                           double dxy_x, int k0, int k1, double (&v)[NV]) {
  constexpr bool MIXED = MODE == kCorrDefect32;
  // loads of plane k: the four in-plane neighbours, b, and the centre at k+1
  struct Plane {
    double xm, xp, ym, yp, fb, pu;
    float xme, xpe, yme, ype, pe;
  };
  // rows are addressed as a plane base plus per-lane byte offsets that are
  // the same in every plane (no vector address arithmetic per load)
  const unsigned dxm = (unsigned)oxm * 8u, dxp = (unsigned)oxp * 8u, dym = (unsigned)oym * 8u,
                 dyp = (unsigned)oyp * 8u, dc = (unsigned)col * 8u;
  const unsigned sxm = (unsigned)oxm * 4u, sxp = (unsigned)oxp * 4u, sym = (unsigned)oym * 4u,
                 syp = (unsigned)oyp * 4u, sc = (unsigned)col * 4u;
  auto load = [&](int k) {
    Plane q;
    const int pb = k * plane;
    const int kp = (k == nz - 1) ? 0 : k + 1;
    const sp<kG, const double> ur = u + pb;
    q.xm = ld_row<double>(ur, dxm); q.xp = ld_row<double>(ur, dxp);
    q.ym = ld_row<double>(ur, dym); q.yp = ld_row<double>(ur, dyp);
    // cs: the compatibility shift of b, subtracted as b is loaded (the
    // first defect; the mixed step, which keeps b as ingested).  The mixed
    // step reads b once per cell and nothing else reads it during the step:
    // a non-temporal load (it was measured slower in the FP64 residual)
    const sp<kG, const double> fr = (sp<kG, const double>)(f + pb);
    q.fb = (MIXED ? ld_row_nt<double>(fr, dc) : ld_row<double>(fr, dc)) - cs;
    q.pu = ld_row<double>(u + kp * plane, dc);
    if constexpr (MIXED) {
      const sp<kG, const float> er = ev + pb;
      q.xme = ld_row<float>(er, sxm); q.xpe = ld_row<float>(er, sxp);
      q.yme = ld_row<float>(er, sym); q.ype = ld_row<float>(er, syp);
      q.pe = ld_row<float>(ev + kp * plane, sc);
    } else {
      q.xme = q.xpe = q.yme = q.ype = q.pe = 0.0f;
    }
    return q;
  };
  auto sum = [&](double a, float b) -> double {
    if constexpr (MIXED) return a + (double)b;
    else return a;
  };
  const int km0 = (k0 == 0) ? nz - 1 : k0 - 1;
  double vm = sum(u[km0 * plane + col], MIXED ? ev[km0 * plane + col] : 0.0f);
  const float ec0 = MIXED ? ev[k0 * plane + col] : 0.0f;
  double vc = sum(u[k0 * plane + col], ec0);
  float ec = ec0;
  auto step = [&](int k, const Plane& cur) {
    const int pbk = k * plane;
    const double vp = sum(cur.pu, cur.pe);
    double s = wxm * sum(cur.xm, cur.xme) + wxp * sum(cur.xp, cur.xpe);
    s += wym * sum(cur.ym, cur.yme);
    s += wyp * sum(cur.yp, cur.ype);
    s += cz[k] * vm;
    s += cz[nz + k] * vp;
    const double diag = (cz[2 * nz + k] + dxy_y) + dxy_x;
    const double res = cur.fb - (s + diag * vc);
    // the shifted b (FP64: the V-cycle's finest right-hand side) and the
    // FP32 defect are read by later kernels only
    if constexpr (SHIFT && MODE == kResMax) st_row_nt<double>(f + pbk, dc, cur.fb);
    if constexpr (MODE == kCorrDefect32) {
      // the new p is read in the next cycle and the FP32 defect by the next
      // V-cycle's first sweep, both after many more bytes than the L2 holds
      st_row_nt<double>(pout + pbk, dc, vc);
      st_row_nt<float>(r32 + pbk, sc, (float)res);
      v[0] = amax(v[0], res);
      v[1] = amax(v[1], (double)ec);
      v[2] = amax(v[2], vc);
    } else {
      if constexpr (MODE == kDefect32) st_row_nt<float>(r32 + pbk, sc, (float)res);
      v[0] = amax(v[0], res);
      v[1] = amax(v[1], cur.fb);
    }
    vm = vc;
    vc = vp;
    ec = cur.pe;
  };
  // a ring of kFineDepth plane buffers, indexed statically (the loop is
  // unrolled by the ring size, so no register copies): the loads of plane
  // k + kFineDepth - 1 are issued before plane k is computed, and the
  // prefetch is unconditional (past the slab it reloads the last plane).
  // Loads complete in issue order, so the wait before a plane's arithmetic
  // leaves the kFineDepth - 1 younger planes in flight.
  // A full slab is unrolled completely and has no branch, so the ring needs
  // no register copies at a loop back edge (they would wait for the loads in
  // flight) and no load is sunk past a branch toward its use.
  constexpr int D = kFineDepth;
  if (k1 - k0 == kMarchK) {
    Plane R[D];
#pragma unroll
    for (int s = 0; s < D - 1; ++s) R[s] = load(k0 + s);
#pragma unroll
    for (int s = 0; s < kMarchK; ++s) {
      if (s + D - 1 < kMarchK) R[(s + D - 1) % D] = load(k0 + s + D - 1);
      // keep the prefetch ahead of the arithmetic: the scheduler would
      // otherwise move the loads toward their use
      __builtin_amdgcn_sched_barrier(0);
      step(k0 + s, R[s % D]);
    }
  } else {
    // the short last slab when nz is not a multiple of kMarchK
    Plane a = load(k0);
    for (int k = k0; k < k1; ++k) {
      const Plane n = load(min(k + 1, k1 - 1));
      step(k, a);
      a = n;
    }
  }
}

// SHIFT: the first defect of a solve, which takes the compatibility shift
// of a singular instance from csum (sum V b per item); FP64 writes the
// shifted b back, mixed keeps b and records the shift in the instance's
// bshift word, which its outer steps subtract as they load b.
template <int MODE, bool SHIFT>
__global__ __launch_bounds__(kBlock) void k_fine(DevLevel* __restrict__ Lt, LaunchList ll,
                                                 const double* __restrict__ csum,
                                                 double* __restrict__ partial) {
  constexpr int NV = MODE == kCorrDefect32 ? 3 : 2;
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  // the first cycle of a pair: no correction (its gate reads no norms)
  if (MODE == kCorrDefect32 && L.skipc) return;
  // the FP32 finest level (mixed); an FP64 solve has none and reads no correction
  const DevLevel& S = Lt[L.link >= 0 ? L.link : ll.items[t]];
  const int hx = L.hx, nx = L.nx, ny = L.ny, nz = L.nz;
  // the current FP64 pressure and its ping-pong partner
  const void* pin = L.pcur ? L.u2 : L.u;
  void* pout = L.pcur ? L.u : L.u2;
  const int plane = 2 * hx * ny;
  const int ncb = (plane + kBlock - 1) / kBlock;
  const int lb = b - ll.prefix[t];
  const int bk = lb / ncb, bc = lb - bk * ncb;
  const int col = bc * kBlock + threadIdx.x;
  const int k0 = bk * kMarchK, k1 = min(k0 + kMarchK, nz);
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  if (col < plane) {
    const int row = gm_div(col, L.gmhx, L.gshx), m = col - row * hx, c = row & 1, j = row >> 1;
    const int i = 2 * m + ((c + j) & 1), r = i & 1, oth = c ^ 1;
    int ml = m - 1 + r; if (ml < 0) ml = hx - 1;
    int mr = m + r;     if (mr >= hx) mr = 0;
    const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
    const auto co = gp(L.co64);
    const auto cy = co + 3 * nx;
    const auto cz = cy + 3 * ny;
    const double cs = MODE == kCorrDefect32 ? L.bshift : (SHIFT && L.nullsp) ? csum[t] / L.vsum : 0.0;
    if (SHIFT && MODE == kDefect32 && lb == 0 && threadIdx.x == 0) Lt[ll.items[t]].bshift = cs;
    fine_march<MODE, NV, SHIFT>(gfield<const double>(pin),
                         gfield<const float>((const void*)S.u), gfield<double>(L.f), cs,
                         gfield<double>(pout),
                         gfield<float>(S.f), cz, nz, plane, col, (j * 2 + oth) * hx + ml,
                         (j * 2 + oth) * hx + mr, (jm * 2 + oth) * hx + m, (jp * 2 + oth) * hx + m,
                         co[i], co[nx + i], cy[j], cy[ny + j], cy[2 * ny + j], co[2 * nx + i], k0, k1, v);
  }
  block_reduce_store<NV, false>(v, partial + (size_t)b * NV);
}

// State words an instance's levels take at the end of a cycle (a negative
// value leaves that word as it is; pflip toggles pcur).
struct GateWords {
  int stop, warm, skipc, pflip;
};

// Writes w on the levels of instance inst among entries l0, l0+step, ...
// of the table (the threads of a block split the table between them).  The
// entries' instance numbers are read eight at a time before any write, so
// a walk pays a memory round trip per eight entries, not per entry.
__device__ void set_words(DevLevel* Lt, int nL, int inst, GateWords w, int l0, int step) {
  constexpr int U = 8;
  for (int lb = l0; lb < nL; lb += U * step) {
    int iv[U];
#pragma unroll
    for (int j = 0; j < U; ++j) iv[j] = lb + j * step < nL ? Lt[lb + j * step].inst : -1;
#pragma unroll
    for (int j = 0; j < U; ++j) {
      if (iv[j] != inst) continue;
      DevLevel& d = Lt[lb + j * step];
      if (w.stop >= 0) d.stop = w.stop;
      if (w.warm >= 0) d.warm = w.warm;
      if (w.skipc >= 0) d.skipc = w.skipc;
      if (w.pflip) d.pcur ^= 1;
    }
  }
}

// What a gate check reads besides the norms of its own fold: the gate, the
// instance's state words and the norms in memory.  k_finalize loads it
// before it folds, so these round trips overlap the fold's.
struct GatePre {
  double o[4], oc[2];
  double bnorm, tol, rprev;
  int gstop, cycles, n, maxcy, rule, fp64, report, gpair, skipc, warm;
  int gslot, inst;
};
__device__ __forceinline__ GatePre gate_pre(const ItemGate* g, const double* out, const double* outc,
                                            const DevLevel& L) {
  GatePre P;
  for (int q = 0; q < 4; ++q) P.o[q] = out[q];
  P.oc[0] = outc[0];
  P.oc[1] = outc[1];
  P.gstop = g->stop;
  P.cycles = g->cycles;
  P.n = g->count;
  P.maxcy = g->maxcy;
  P.rule = g->rule;
  P.fp64 = g->fp64;
  P.report = g->report;
  P.gpair = g->pair;
  P.bnorm = g->bnorm;
  P.tol = g->tol;
  P.rprev = g->rprev;
  P.skipc = L.skipc;
  P.warm = L.warm;
  P.gslot = L.gslot;
  P.inst = L.inst;
  return P;
}

// End of a gated cycle of one instance (one thread; L its item's level,
// the FP64 finest): the first cycle of a pair only hands over to the
// second; otherwise the correction's norms are recorded and the stop rule
// of the host's norm step (Context::post_norms) applied to them with the
// same double operations, so the host's replay reaches the same decision;
// then the next cycle is planned (paired or not).  Returns the state words
// the instance's levels take.
// P holds what it reads (gate_pre, with the norms of the calling fold put
// in); only stores remain here.
__device__ GateWords item_gate_check(ItemGate* g, const GatePre& P) {
  const double* o = P.o;
  const double* oc = P.oc;
  const int gstop = P.gstop, cycles = P.cycles + 1, n = P.n, maxcy = P.maxcy;
  const int rule = P.rule, fp64 = P.fp64, report = P.report, gpair = P.gpair;
  const double bnorm = P.bnorm, tol = P.tol, rprev = P.rprev;
  const int skipc = P.skipc, warm = P.warm;
  if (gstop) return {-1, -1, -1, 0};
  g->cycles = cycles;
  if (skipc) return {-1, 1, 0, 0};
  const int m = warm ? 2 : 1;
  double* h = g->hist[n % kItemRing];
  for (int q = 0; q < 4; ++q) h[q] = o[q];
  h[4] = report ? oc[0] : 0.0;
  h[5] = report ? oc[1] : 0.0;
  g->hcyc[n % kItemRing] = m;
  const double rmax = o[0];
  double relmax;
  if (fp64) relmax = report ? (oc[1] > 0.0 ? oc[0] / oc[1] : 0.0) : __longlong_as_double(0x7ff8000000000000ll);
  else relmax = o[2] > 0.0 ? o[1] / o[2] : 0.0;
  const double relres = bnorm > 0.0 ? rmax / bnorm : rmax;
  g->count = n + 1;
  int stop = 0;
  if (!isfinite(rmax)) {
    stop = 1;
  } else {
    const double crit = (rule == AMG_STOP_RELMAX) ? relmax : relres;
    if (crit <= tol || cycles >= maxcy) stop = 1;
  }
  const bool pair = !stop && gpair && cycles + 2 <= maxcy && pair_next(relres, rprev, m, tol);
  g->rprev = relres;
  g->stop = stop;
  return {stop, 0, pair ? 1 : 0, fp64 ? 0 : 1};
}

// One thread per item: the gate of a starting solve, from its first defect
// (the decisions the host's replay, Context::absorb_gate, takes with the
// same operations), and the plan of its first cycle.
__global__ void k_gate_init(LaunchList ll, const double* __restrict__ out, ItemGate* ig,
                            DevLevel* Lt, int nL, int maxcy, int rule, double tol, int report) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ll.nitems) return;
  const DevLevel& L = Lt[ll.items[t]];
  ItemGate* g = ig + L.gslot;
  const double r = out[(size_t)t * 4], b = out[(size_t)t * 4 + 1];
  const int fp64 = L.link < 0;
  g->count = 0;
  g->cycles = 0;
  g->maxcy = maxcy;
  g->rule = rule;
  g->fp64 = fp64;
  g->report = fp64 && report;
  g->pair = L.pairok;
  g->tol = tol;
  g->bnorm = b;
  g->res0 = b > 0.0 ? r / b : r;
  g->rprev = g->res0;
  int stop = 0;
  if (!isfinite(r) || !isfinite(b)) stop = 1;
  else if (rule == AMG_STOP_RESIDUAL && (b == 0.0 || g->res0 <= tol)) stop = 1;
  g->stop = stop;
  const bool pair = !stop && g->pair && maxcy >= 2 && g->res0 * kPairRho0 > tol;
  set_words(Lt, nL, L.inst, {stop, 0, pair ? 1 : 0, 0}, 0, 1);
}

// One block per item: thread t folds partials b0+t, b0+t+kFinBlock, ... in
// order, then a fixed-shape tree combines the threads (deterministic).  With
// per-instance gates, thread 0 of every item's block then ends its cycle and
// the block's threads write the resulting state words, one table entry each.
constexpr int kFinBlock = 1024;
__global__ __launch_bounds__(kFinBlock) void k_finalize(LaunchList ll, const double* __restrict__ partial,
                                                        int nvals, int op_sum, double* __restrict__ out,
                                                        int ostride, GateArgs ga) {
  __shared__ double sm[kFinBlock / 64];
  const int t = blockIdx.x;
  const int b0 = ll.prefix[t], b1 = ll.prefix[t + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // the gate's inputs, loaded before the fold (thread 0), and the instance
  // number of every thread's first table entry (never written on the device)
  GatePre gpre;
  int inst0 = -1;
  if (ga.ig) {
    if (threadIdx.x < ga.nL) inst0 = ga.Lt[threadIdx.x].inst;
    if (threadIdx.x == 0) {
      const DevLevel& L = ga.Lt[ll.items[t]];
      gpre = gate_pre(ga.ig + L.gslot, ga.out + (size_t)t * 4, ga.outc + (size_t)t * 2, L);
    }
  }
  // one pass over the partials for every value (nvals <= 4), kFinU
  // partials per thread and round: their loads (clamped into the item, so
  // unconditional) are issued before they are folded, in order
  constexpr int kFinU = 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = b0 + threadIdx.x; b < b1; b += kFinBlock * kFinU) {
    double x[kFinU][4];
#pragma unroll
    for (int r = 0; r < kFinU; ++r) {
      const int bb = min(b + r * kFinBlock, b1 - 1);
#pragma unroll
      for (int q = 0; q < 4; ++q) x[r][q] = partial[(size_t)bb * nvals + min(q, nvals - 1)];
    }
#pragma unroll
    for (int r = 0; r < kFinU; ++r) {
      if (b + r * kFinBlock < b1) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (q < nvals) acc[q] = op_sum ? acc[q] + x[r][q] : nmax(acc[q], x[r][q]);
      }
    }
  }
  double fr[4] = {0.0, 0.0, 0.0, 0.0};   // thread 0: the folded values
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q >= nvals) break;
    double a = acc[q];
    a = op_sum ? wave_sum(a) : wave_max(a);
    if (lane == 0) sm[w] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = sm[0];
      for (int i = 1; i < kFinBlock / 64; ++i) r = op_sum ? r + sm[i] : nmax(r, sm[i]);
      out[(size_t)t * ostride + q] = r;
      fr[q] = r;
    }
    __syncthreads();
  }
  if (ga.ig) {
    __shared__ GateWords gw;
    __shared__ int ginst;
    if (threadIdx.x == 0) {
      // the gate's norms that this fold wrote: its values, not the loads
      // made before it
      const double* mine = out + (size_t)t * ostride;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q >= nvals) break;
        for (int r = 0; r < 4; ++r)
          if (ga.out + (size_t)t * 4 + r == mine + q) gpre.o[r] = fr[q];
        for (int r = 0; r < 2; ++r)
          if (ga.outc + (size_t)t * 2 + r == mine + q) gpre.oc[r] = fr[q];
      }
      // (the descriptor's words as loaded before the fold: the stores to
      // out may alias the table as far as the compiler knows)
      gw = item_gate_check(ga.ig + gpre.gslot, gpre);
      ginst = gpre.inst;
    }
    __syncthreads();
    // the first entry of each thread from its prefetched instance number
    const GateWords wv = gw;
    const int gi = ginst;
    if (threadIdx.x < ga.nL && inst0 == gi) {
      DevLevel& d = ga.Lt[threadIdx.x];
      if (wv.stop >= 0) d.stop = wv.stop;
      if (wv.warm >= 0) d.warm = wv.warm;
      if (wv.skipc >= 0) d.skipc = wv.skipc;
      if (wv.pflip) d.pcur ^= 1;
    }
    set_words(ga.Lt, ga.nL, gi, wv, threadIdx.x + kFinBlock, kFinBlock);
  }
}

// ---------------------------------------------------------------------------
// Layout conversions (FP64).
__device__ __forceinline__ size_t fidx(int i, int j, int k, int nx, int ny) {
  return ((size_t)(k + 1) * (ny + 2) + (j + 1)) * (nx + 2) + (i + 1);
}

__global__ __launch_bounds__(kBlock) void k_from_fortran(const DevLevel* __restrict__ Lt,
                                                         LaunchList ll, PtrList ptr,
                                                         int which, int add_fold) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e >= L.ncell) return;
  int i, j, k;
  decode(L, e, i, j, k);
  double v = gp(ptr.p[t])[fidx(i, j, k, L.nx, L.ny)];
  if (add_fold) {
    const auto bf = gp(L.bcfold);
    v += (bf[i] + bf[L.nx + j]) + bf[L.nx + L.ny + k];
  }
  gfield<double>(which == 0 ? L.u : L.f)[e] = v;
}

// Ingest of a solve: b = rhs + boundary fold and u = phi (reference layout
// -> chequer), with the partial sum of V*b of the block's cells.  A thread
// takes cell pairs (2m, j, k), (2m+1, j, k): one decode per pair, the pair's
// two reference-layout values are adjacent (a wave reads contiguous rows),
// and the two cells go to the two colours' half-rows at position m.  Each
// thread folds kWsumPer cells (kWsumPer/2 pairs) in a fixed order.
__global__ __launch_bounds__(kBlock) void k_ingest(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                   PtrList rhs, PtrList phi,
                                                   double* __restrict__ partial) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  constexpr int kPairs = kWsumPer / 2;
  const int np = L.ncell >> 1;                        // pairs: hx * ny * nz
  const int p0 = (b - ll.prefix[t]) * kBlock * kPairs + threadIdx.x;
  const auto vol = gp(L.vol);
  const auto bf = gp(L.bcfold);
  const auto src = gp(rhs.p[t]);
  const auto ph = gp(phi.p[t]);
  const auto f = gfield<double>(L.f);
  const auto u = gfield<double>(L.u);
  const int nx = L.nx, ny = L.ny, hx = L.hx;
  double v[1] = {0.0};
  // every load is unconditional (pair indices clamped into the item), so
  // each chunk of kIngestU pairs issues its loads before their uses; pairs
  // past the item's end are masked at the stores and the sum
  constexpr int kIngestU = 4;
#pragma unroll
  for (int s0 = 0; s0 < kPairs; s0 += kIngestU) {
    if (s0 > 0) __builtin_amdgcn_sched_barrier(0);
    double s0v[kIngestU], s1v[kIngestU], q0v[kIngestU], q1v[kIngestU], fb0[kIngestU], fb1[kIngestU];
    double w0[kIngestU], w1[kIngestU];
    int e0[kIngestU], e1[kIngestU];
#pragma unroll
    for (int q = 0; q < kIngestU; ++q) {
      const int pp = p0 + (s0 + q) * kBlock;
      const int pc = pp < np ? pp : np - 1;
      const int row = gm_div(pc, L.gmhx, L.gshx), m = pc - row * hx;
      const int k = gm_div(row, L.gmny, L.gsny), j = row - k * ny;
      const int i = 2 * m;
      const size_t fi = fidx(i, j, k, nx, ny);
      s0v[q] = src[fi];
      s1v[q] = src[fi + 1];
      q0v[q] = ph[fi];
      q1v[q] = ph[fi + 1];
      fb0[q] = (bf[i] + bf[nx + j]) + bf[nx + ny + k];
      fb1[q] = (bf[i + 1] + bf[nx + j]) + bf[nx + ny + k];
      const double wyz = vol[nx + ny + k] * vol[nx + j];
      w0[q] = wyz * vol[i];
      w1[q] = wyz * vol[i + 1];
      const int rb = (k * ny + j) * 2;
      e0[q] = (rb + (j & 1)) * hx + m;                 // colour of (2m, j): j & 1
      e1[q] = (rb + ((j + 1) & 1)) * hx + m;
    }
#pragma unroll
    for (int q = 0; q < kIngestU; ++q) {
      const bool in = p0 + (s0 + q) * kBlock < np;
      const double bv0 = s0v[q] + fb0[q], bv1 = s1v[q] + fb1[q];
      // the first defect reads b and p back only after the whole ingest,
      // more than the caches hold: non-temporal stores (the ingest takes 6
      // us longer at 256^3, the first defect 9-13 us less)
      if (in) {
        __builtin_nontemporal_store(bv0, (double*)(f + e0[q]));
        __builtin_nontemporal_store(bv1, (double*)(f + e1[q]));
        __builtin_nontemporal_store(q0v[q], (double*)(u + e0[q]));
        __builtin_nontemporal_store(q1v[q], (double*)(u + e1[q]));
      }
      v[0] += in ? w0[q] * bv0 : 0.0;
      v[0] += in ? w1[q] * bv1 : 0.0;
    }
  }
  block_reduce_store<1, true>(v, partial + b);
}

// Weighted sums sum V x (V = dx_i dy_j dz_k), marching in z like the
// finest-level stencil kernels (march_blocks per item): a thread owns
// position col of a plane and kMarchK planes, sums dz_k x_k over them in
// order and weights that sum by dx_i dy_j; one partial per block.  The dz_k
// are wave-uniform (scalar loads), so a cell costs one vector load.
__global__ __launch_bounds__(kBlock) void k_wsum(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                 int which, double* __restrict__ partial) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const int hx = L.hx, ny = L.ny, nz = L.nz, nx = L.nx;
  const int plane = 2 * hx * ny;
  const int ncb = (plane + kBlock - 1) / kBlock;
  const int lb = b - ll.prefix[t];
  const int bk = lb / ncb, col = (lb - bk * ncb) * kBlock + threadIdx.x;
  const int k0 = bk * kMarchK, k1 = min(k0 + kMarchK, nz);
  const auto vol = gp(L.vol);
  const auto rho = gp(L.rho);
  // which 0: the current pressure of an FP64 finest entry (its ping-pong
  // state word pcur says which buffer holds it); 1: f; 2: rho(z)
  const auto x = gfield<const double>((const void*)(which == 0 ? (L.pcur ? L.u2 : L.u) : L.f));
  double v[1] = {0.0};
  if (col < plane) {
    const int row = gm_div(col, L.gmhx, L.gshx), m = col - row * hx, j = row >> 1;
    const int i = 2 * m + (((row & 1) + j) & 1);
    const auto vz = vol + nx + ny;
    double s = 0.0;
    if (k1 - k0 == kMarchK) {
      double xv[kMarchK];
#pragma unroll
      for (int q = 0; q < kMarchK; ++q)
        xv[q] = which == 2 ? rho[k0 + q] : x[(size_t)(k0 + q) * plane + col];
#pragma unroll
      for (int q = 0; q < kMarchK; ++q) s += vz[k0 + q] * xv[q];
    } else {
      for (int k = k0; k < k1; ++k) s += vz[k] * (which == 2 ? rho[k] : x[(size_t)k * plane + col]);
    }
    v[0] = (vol[i] * vol[nx + j]) * s;
  }
  block_reduce_store<1, true>(v, partial + b);
}

__global__ void k_zero(double* __restrict__ a, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = 0.0;
}

// dst interior = u, minus the null-space component (c[t]/vrsum) rho(z) of
// singular items when c is given
__global__ __launch_bounds__(kBlock) void k_to_fortran(const DevLevel* __restrict__ Lt,
                                                       LaunchList ll, PtrList ptr,
                                                       const double* __restrict__ c) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e >= L.ncell) return;
  int i, j, k;
  decode(L, e, i, j, k);
  // the current pressure (pcur: which buffer of the ping-pong pair holds it)
  double x = gfield<const double>((const void*)(L.pcur ? L.u2 : L.u))[e];
  if (c && L.nullsp) x -= (c[t] / L.vrsum) * gp(L.rho)[k];
  gp(ptr.p[t])[fidx(i, j, k, L.nx, L.ny)] = x;
}

__global__ __launch_bounds__(kBlock) void k_copy_u(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                   const int* __restrict__ active) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  if (active && !active[L.inst]) return;
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e >= L.ncell) return;
  gfield<double>(L.u2)[e] = gfield<const double>((const void*)L.u)[e];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_fill_zero(const DevLevel* __restrict__ Lt, LaunchList ll) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e < L.ncell) gfield<T>(L.u)[e] = T(0);
}

// ---------------------------------------------------------------------------
// Ghost layers (serial path of mgMK/mg_bdry.F90:71-190).
__global__ void k_ghost_wrap(double* __restrict__ a, int nx, int ny, int nz, int dir) {
  const int sx = nx + 2, sy = ny + 2, sz = nz + 2;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (dir == 0) {
    if (e >= sy * sz) return;
    const size_t base = (size_t)e * sx;
    a[base + 0] = a[base + nx];
    a[base + nx + 1] = a[base + 1];
  } else if (dir == 1) {
    if (e >= sx * sz) return;
    const int i = e % sx, k = e / sx;
    const size_t pl = (size_t)k * sy * sx;
    a[pl + i] = a[pl + (size_t)ny * sx + i];
    a[pl + (size_t)(ny + 1) * sx + i] = a[pl + (size_t)sx + i];
  } else {
    if (e >= sx * sy) return;
    const size_t pl = (size_t)sx * sy;
    a[e] = a[(size_t)nz * pl + e];
    a[(size_t)(nz + 1) * pl + e] = a[pl + e];
  }
}

// Every ghost cell of a ghosted field in one pass, with the values of the
// serial sequence (mgMK/mg_bdry.F90:71-190): the three periodic wraps over
// full planes, then the non-periodic faces in the order 1 X2max, 2 X3min,
// 3 X2min, 4 X3max, 5 X1min, 6 X1max, each from the layer inside it as the
// earlier passes left it.  A wrap only copies, so after the wraps a ghost
// cell holds the interior value at its coordinates wrapped one by one.  A
// face rule reads the cell one step inside its face, which the last earlier
// face whose layer holds that cell wrote, and so on: walking the faces from
// the last down and stepping inside each face whose layer holds the current
// cell gives that chain, whose rules are then applied forward to the
// wrapped interior value.  Interior cells are only read, and a cell on an
// edge or corner is computed alike by every thread that owns it.
// GhostBC: per face 1..6, type (0 periodic: no rule, 1 Neumann, 2
// Dirichlet), the value v and v*dxc(edge cell) (the Neumann offset, a
// product made on the host so the rule is one rounding, x -/+ vd).
__global__ __launch_bounds__(kBlock) void k_ghost_fill(double* __restrict__ a, int nx, int ny, int nz,
                                                       GhostBC bc) {
  // (32-bit index arithmetic: the ghost cells of any grid this library
  // accepts number far below 2^31)
  const int sx = nx + 2, sy = ny + 2, sz = nz + 2;
  const int nxf = 2 * sy * sz, nyf = 2 * sx * sz, nzf = 2 * sx * sy;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (e < nxf) {
    const int r = e >> 1;
    i = (e & 1) ? nx + 1 : 0; k = r / sy; j = r - k * sy;
  } else if (e < nxf + nyf) {
    const int r = (e - nxf) >> 1;
    j = ((e - nxf) & 1) ? ny + 1 : 0; k = r / sx; i = r - k * sx;
  } else if (e < nxf + nyf + nzf) {
    const int r = (e - nxf - nyf) >> 1;
    k = ((e - nxf - nyf) & 1) ? nz + 1 : 0; j = r / sx; i = r - j * sx;
  } else {
    return;
  }
  int chain[6], nc = 0;
  int ci = i, cj = j, ck = k;
  for (int f = 6; f >= 1; --f) {
    if (!bc.type[f]) continue;
    bool in = false;
    switch (f) {
      case 1: in = cj == ny + 1; if (in) cj = ny; break;
      case 2: in = ck == 0; if (in) ck = 1; break;
      case 3: in = cj == 0; if (in) cj = 1; break;
      case 4: in = ck == nz + 1; if (in) ck = nz; break;
      case 5: in = ci == 0; if (in) ci = 1; break;
      default: in = ci == nx + 1; if (in) ci = nx; break;
    }
    if (in) chain[nc++] = f;
  }
  auto wrap = [](int c, int n) { return c == 0 ? n : (c == n + 1 ? 1 : c); };
  double x = a[((size_t)wrap(ck, nz) * sy + wrap(cj, ny)) * sx + wrap(ci, nx)];
  for (int q = nc - 1; q >= 0; --q) {
    const int f = chain[q];
    if (bc.type[f] == 2) x = bc.v[f];
    else if (f == 1 || f == 4 || f == 6) x = x + bc.vd[f];
    else x = x - bc.vd[f];
  }
  a[((size_t)k * sy + j) * sx + i] = x;
}

// ---------------------------------------------------------------------------
// Projection step (dns/NSsolver.F90:2204-2219, :2057-2082; dns/Routines.F90:123-144).
__global__ __launch_bounds__(kBlock) void k_psource(const double* __restrict__ U,
                                                    const double* __restrict__ V,
                                                    const double* __restrict__ W,
                                                    double* __restrict__ S,
                                                    const double* __restrict__ rho,
                                                    const double* __restrict__ rdxc,
                                                    const double* __restrict__ rdyc,
                                                    const double* __restrict__ rdzc, int nx, int ny,
                                                    int nz, double rdt) {
  const int sx = nx + 2, sy = ny + 2, sz = nz + 2;
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= sx * sy * sz) return;
  const int i = e % sx, jk = e / sx, j = jk % sy, k = jk / sy;
  if (i == 0 || j == 0 || k == 0 || i == nx + 1 || j == ny + 1 || k == nz + 1) {
    S[e] = 0.0;
    return;
  }
  const size_t sxy = (size_t)sx * sy;
  const double rk = rho[k], rp = rho[k + 1], rm = rho[k - 1];
  S[e] = ((rk * U[e] - rk * U[e - 1]) * rdxc[i] + (rk * V[e] - rk * V[e - sx]) * rdyc[j] +
          (0.5 * (rp + rk) * W[e] - 0.5 * (rk + rm) * W[e - sxy]) * rdzc[k]) * rdt;
}

__global__ __launch_bounds__(kBlock) void k_pgrad(const double* __restrict__ p,
                                                  double* __restrict__ U, double* __restrict__ V,
                                                  double* __restrict__ W,
                                                  const double* __restrict__ rho,
                                                  const double* __restrict__ rdxe,
                                                  const double* __restrict__ rdye,
                                                  const double* __restrict__ rdze, int nx, int ny,
                                                  int nz, double dt, int wallmask) {
  const int e0 = blockIdx.x * kBlock + threadIdx.x;
  if (e0 >= nx * ny * nz) return;
  const int i = e0 % nx + 1, jk = e0 / nx, j = jk % ny + 1, k = jk / ny + 1;
  const int sx = nx + 2, sy = ny + 2;
  const size_t sxy = (size_t)sx * sy;
  const size_t e = (size_t)k * sxy + (size_t)j * sx + i;
  const double rk = rho[k], rp = rho[k + 1];
  const double pc = p[e];
  // the last face of a direction with a pressure-Neumann upper wall is the
  // wall itself: its normal velocity is the velocity boundary condition
  if (!((wallmask & 1) && i == nx)) U[e] = U[e] - (p[e + 1] / rk - pc / rk) * rdxe[i] * dt;
  if (!((wallmask & 2) && j == ny)) V[e] = V[e] - (p[e + sx] / rk - pc / rk) * rdye[j] * dt;
  if (!((wallmask & 4) && k == nz)) W[e] = W[e] - (p[e + sxy] / rp - pc / rk) * rdze[k] * dt;
}

// Partials per block: {sum |div| vol, sum vol, max |div|, index of the max}.
// Divergence report (dns/Routines.F90:123-173): a block covers
// kDivPer*kBlock consecutive interior cells (x fastest), each thread
// kDivPer of them at a stride of kBlock (coalesced rows), folded in a fixed
// order.  Per block: {sum |div| V, sum V, max |div|, its cell index}; the
// maximum keeps the smallest index on ties (the reference's first-found
// order).
__device__ __forceinline__ void div_better(double& b, long long& bi, double ob, long long oi) {
  if (ob > b || (ob == b && oi < bi)) { b = ob; bi = oi; }
}

__global__ __launch_bounds__(kBlock) void k_divergence(const double* __restrict__ U,
                                                       const double* __restrict__ V,
                                                       const double* __restrict__ W,
                                                       double* __restrict__ div,
                                                       const double* __restrict__ rdxc,
                                                       const double* __restrict__ rdyc,
                                                       const double* __restrict__ rdzc, int nx,
                                                       int ny, int nz, double* __restrict__ partial) {
  const long long n = (long long)nx * ny * nz;
  const long long base = (long long)blockIdx.x * kBlock * kDivPer + threadIdx.x;
  const int sx = nx + 2, sy = ny + 2;
  const size_t sxy = (size_t)sx * sy;
  double s[2] = {0.0, 0.0};
  double best = -1.0;
  long long bidx = (long long)1 << 62;
#pragma unroll 4
  for (int q = 0; q < kDivPer; ++q) {
    const long long e0 = base + (long long)q * kBlock;
    if (e0 < n) {
      const int i = (int)(e0 % nx) + 1;
      const long long jk = e0 / nx;
      const int j = (int)(jk % ny) + 1, k = (int)(jk / ny) + 1;
      const size_t e = (size_t)k * sxy + (size_t)j * sx + i;
      const double d = (U[e] - U[e - 1]) * rdxc[i] + (V[e] - V[e - sx]) * rdyc[j] +
                       (W[e] - W[e - sxy]) * rdzc[k];
      if (div) div[e] = d;
      const double vinv = rdxc[i] * rdyc[j] * rdzc[k];
      s[0] += fabs(d) / vinv;
      s[1] += 1.0 / vinv;
      div_better(best, bidx, fabs(d), e0);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_xor(best, off, 64);
    const long long oi = __shfl_xor(bidx, off, 64);
    div_better(best, bidx, ob, oi);
  }
  __shared__ double sb[kBlock / 64];
  __shared__ long long si[kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sb[w] = best; si[w] = bidx; }
  double* out = partial + (size_t)blockIdx.x * 4;
  block_reduce_store<2, true>(s, out);   // contains __syncthreads
  if (threadIdx.x == 0) {
    double b = sb[0];
    long long bi = si[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) div_better(b, bi, sb[q], si[q]);
    out[2] = b;
    out[3] = (double)bi;
  }
}

// One block of kDivFin threads: thread t folds partials t, t+kDivFin, ...
// in order, then a fixed-shape tree combines the threads (deterministic).
constexpr int kDivFin = 1024;
__global__ __launch_bounds__(kDivFin) void k_divergence_finalize(const double* __restrict__ partial,
                                                                 int nb, double* __restrict__ out) {
  __shared__ double s0s[kDivFin / 64], s1s[kDivFin / 64], bs[kDivFin / 64];
  __shared__ long long is[kDivFin / 64];
  double s0 = 0.0, s1 = 0.0, b = -1.0;
  long long bi = (long long)1 << 62;
  for (int q = threadIdx.x; q < nb; q += kDivFin) {
    const double* p = partial + (size_t)q * 4;
    s0 += p[0];
    s1 += p[1];
    div_better(b, bi, p[2], (long long)p[3]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    const double ob = __shfl_xor(b, off, 64);
    const long long oi = __shfl_xor(bi, off, 64);
    div_better(b, bi, ob, oi);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { s0s[w] = s0; s1s[w] = s1; bs[w] = b; is[w] = bi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a0 = s0s[0], a1 = s1s[0], ab = bs[0];
    long long ai = is[0];
    for (int q = 1; q < kDivFin / 64; ++q) {
      a0 += s0s[q];
      a1 += s1s[q];
      div_better(ab, ai, bs[q], is[q]);
    }
    out[0] = a0; out[1] = a1; out[2] = ab; out[3] = (double)ai;
  }
}

// ---------------------------------------------------------------------------
// Separable direct solve of a coarse level (host set-up in amg_direct.cpp):
// u = A^-1 f with A = Ax + Ay + Az and every 1-D operator diagonalised,
// A_d = V_d diag(l_d) V_d^-1.  Three passes, all dense products on the
// matrix cores, in the compute type CT of the level's tables: FP64
// (v_mfma_f64_16x16x4f64) for an FP64 V-cycle, FP32 (v_mfma_f32_16x16x4f32)
// for the FP32 V-cycle of a mixed solve, whose coarse correction needs no
// more (at 64^3 the FP32 passes are the faster):
//   k_direct_xy<FWD>  per plane k:      f^ = Vy^-1 f_k Vx^-T          (-> dscr)
//   k_direct_z        per 16 modes:     x^ = Vz (inv .* (Vz^-1 f^))    (-> dscr + n)
//   k_direct_xy<inv>  per plane k:      u_k = Vy x^_k Vx^T
// with inv = 1/(lx + ly + lz) per eigenmode (0 for a singular operator's
// null mode: the solution then has sum V u = 0).  Operands are staged in
// LDS; rows are padded by one word against bank conflicts.
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// The 16x16x4 product of each compute type: lane l holds A[i][k] and B[k][j]
// with i, j = l % 16 and k = l / 16; result register r of lane l holds
// D[row(l / 16, r)][l % 16] (the two types lay D out differently).
template <class CT> struct Mfma16;
template <> struct Mfma16<double> {
  typedef dbl4 V;
  static __device__ __forceinline__ V op(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct Mfma16<float> {
  typedef flt4 V;
  static __device__ __forceinline__ V op(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// LDS of the plane passes and of the z pass (es: bytes per element)
__host__ __device__ inline size_t direct_lds_xy(int nx, int ny, size_t es) {
  return ((size_t)ny * (nx + 1) + (size_t)ny * (ny + 1) + (size_t)(nx + ny) * (kDirectCols + 1)) * es;
}
__host__ __device__ inline size_t direct_lds_z(int nzp, size_t es) {
  return ((size_t)2 * nzp * (kDirectCols + 1) + (size_t)2 * nzp * (nzp + 1)) * es;
}
__host__ __device__ inline int direct_nzp(int nz) { return (nz + 15) & ~15; }

// C = A B for an nr x nc output, A (nr x nq, row stride as) and B (nq x nc,
// row stride bs) in LDS; nr, nc, nq multiples of 16.  A wave takes one
// 16x16 tile at a time (Mfma16 lane layout).  The operands of four products
// are read before them.
template <class CT, class AP, class BP, class F>
__device__ __forceinline__ void mfma_products(AP A, int as, BP B, int bs, int nr, int nc, int nq,
                                              F&& store) {
  typedef Mfma16<CT> M;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int tr = nr >> 4, tc = nc >> 4;
  for (int tile = wave; tile < tr * tc; tile += nw) {
    const int r0 = (tile / tc) << 4, c0 = (tile % tc) << 4;
    typename M::V acc = {CT(0), CT(0), CT(0), CT(0)};
    for (int q0 = 0; q0 < nq; q0 += 16) {
      CT av[4], bv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        av[q] = A[(r0 + li) * as + q0 + 4 * q + lk];
        bv[q] = B[(q0 + 4 * q + lk) * bs + c0 + li];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) acc = M::op(av[q], bv[q], acc);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) store(r0 + M::row(lk, r), c0 + li, acc[r]);
  }
}

// Global -> LDS staging of three arrays at once, N loads of each in flight
// per thread before any store (a loop of load-store pairs would wait out a
// memory round trip per element): store_a(e, load_a(e)) for e < na, and
// the same for b and c.
template <int N, class CT, class LA, class SA, class LB, class SB, class LC, class SC>
__device__ __forceinline__ void stage3(int na, LA&& la, SA&& sa, int nb, LB&& lb, SB&& sb, int nc,
                                       LC&& lc, SC&& sc) {
  const int bd = blockDim.x, step = N * bd;
  const int n = max(na, max(nb, nc));
  for (int e0 = threadIdx.x; e0 < n; e0 += step) {
    CT va[N], vb[N], vc[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
      const int e = e0 + q * bd;
      va[q] = la(min(e, na - 1));
      vb[q] = lb(min(e, nb - 1));
      vc[q] = lc(min(e, nc - 1));
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
      const int e = e0 + q * bd;
      if (e < na) sa(e, va[q]);
      if (e < nb) sb(e, vb[q]);
      if (e < nc) sc(e, vc[q]);
    }
  }
}

// The plane passes, one workgroup per (plane k, slab of kDirectCols columns
// c0..): FWD: the slab of f^_k = Vy^-1 f_k Vx^-T into dscr, from plane k of
// f (level precision T); otherwise the slab of u_k = Vy x^_k Vx^T from
// dscr + n.  Either is My (X Mx[:, slab]) with the whole plane X staged.
// The tables (dmat, dinv, dscr) hold CT values.
template <class T, class CT, bool FWD>
__global__ __launch_bounds__(kBlock) void k_direct_xy(const DevLevel* __restrict__ Lt, LaunchList ll) {
  extern __shared__ double dl[];
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  const int nx = L.nx, ny = L.ny, hx = L.hx, sx = nx + 1, sy = ny + 1;
  constexpr int sc = kDirectCols + 1;
  const int ns = nx / kDirectCols, q = b - ll.prefix[t], k = q / ns, c0 = (q - k * ns) * kDirectCols;
  const sp<kS, CT> X = (sp<kS, CT>)(sp<kS, double>)dl;   // the plane, ny x nx (stride sx)
  const sp<kS, CT> My = X + ny * sx;         // FWD: Vy^-1, else Vy (ny x ny, stride sy)
  const sp<kS, CT> Mx = My + ny * sy;        // columns c0.. of Vx^-T / Vx^T (nx x 16)
  const sp<kS, CT> Y = Mx + nx * sc;         // X Mx (ny x 16)
  const auto m = gp(reinterpret_cast<const CT*>(L.dmat));
  const size_t ox = FWD ? 0 : (size_t)nx * nx;
  const size_t oy = (size_t)2 * nx * nx + (FWD ? 0 : (size_t)ny * ny);
  auto lmx = [&](int e) { return m[ox + (size_t)(e / kDirectCols) * nx + c0 + e % kDirectCols]; };
  auto smx = [&](int e, CT v) { Mx[(e / kDirectCols) * sc + e % kDirectCols] = v; };
  auto lmy = [&](int e) { return m[oy + e]; };
  // row and column of a staged element by the level's invariant divisors
  // (e / nx = (e / 2) / hx)
  auto smy = [&](int e, CT v) {
    const int r = gm_div(e, L.gmny, L.gsny);
    My[r * sy + e - r * ny] = v;
  };
  auto sxp = [&](int e, CT v) {
    const int r = gm_div(e >> 1, L.gmhx, L.gshx);
    X[r * sx + e - r * nx] = v;
  };
  const auto scr = gp(reinterpret_cast<CT*>(L.dscr));
  if constexpr (FWD) {
    const auto f = gfield<const T>((const void*)L.f);
    auto lf = [&](int e) {
      const int j = gm_div(e >> 1, L.gmhx, L.gshx), i = e - j * nx;
      return (CT)f[cidx(i >> 1, j, k, (i + j) & 1, hx, ny)];
    };
    stage3<16, CT>(nx * ny, lf, sxp, ny * ny, lmy, smy, nx * kDirectCols, lmx, smx);
  } else {
    const auto in = scr + (size_t)L.ncell + (size_t)k * ny * nx;
    stage3<16, CT>(nx * ny, [&](int e) { return in[e]; }, sxp, ny * ny, lmy, smy, nx * kDirectCols, lmx,
                   smx);
  }
  __syncthreads();
  mfma_products<CT>(X, sx, Mx, sc, ny, kDirectCols, nx, [&](int r, int c, CT v) { Y[r * sc + c] = v; });
  __syncthreads();
  if constexpr (FWD) {
    const auto out = scr + (size_t)k * ny * nx + c0;
    mfma_products<CT>(My, sy, Y, sc, ny, kDirectCols, ny, [&](int r, int c, CT v) { out[r * nx + c] = v; });
  } else {
    const auto u = gfield<T>(L.u);
    mfma_products<CT>(My, sy, Y, sc, ny, kDirectCols, ny, [&](int r, int c, CT v) {
      const int i = c0 + c;
      u[cidx(i >> 1, r, k, (i + r) & 1, hx, ny)] = (T)v;
    });
  }
}

// The z pass, one workgroup per slab of kDirectCols modes m0.. (all k):
// x^ = Vz (inv .* (Vz^-1 f^)).
template <class CT>
__global__ __launch_bounds__(kBlock) void k_direct_z(const DevLevel* __restrict__ Lt, LaunchList ll) {
  extern __shared__ double dl[];
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (level_off(L)) return;
  const int nx = L.nx, ny = L.ny, nz = L.nz, nzp = direct_nzp(nz);
  const int plane = nx * ny, m0 = (b - ll.prefix[t]) * kDirectCols, sz = nzp + 1;
  constexpr int sc = kDirectCols + 1;
  const sp<kS, CT> S = (sp<kS, CT>)(sp<kS, double>)dl;   // nzp x 16 (stride 17)
  const sp<kS, CT> Tm = S + nzp * sc;        // nzp x 16
  const sp<kS, CT> Wi = Tm + nzp * sc;       // Vz^-1 (nzp x nzp, stride sz)
  const sp<kS, CT> W = Wi + nzp * sz;        // Vz
  const auto m = gp(reinterpret_cast<const CT*>(L.dmat));
  const size_t oz = (size_t)2 * nx * nx + (size_t)2 * ny * ny;
  const auto fh = gp(reinterpret_cast<const CT*>(L.dscr));
  const auto inv = gp(reinterpret_cast<const CT*>(L.dinv));
  auto zrow = [&](int e) { return gm_div(e, L.gmzp, L.gszp); };   // e / nzp
  stage3<16, CT>(
      nzp * nzp, [&](int e) { return m[oz + e]; }, [&](int e, CT v) { Wi[zrow(e) * sz + e - zrow(e) * nzp] = v; },
      nzp * nzp, [&](int e) { return m[oz + (size_t)nzp * nzp + e]; },
      [&](int e, CT v) { W[zrow(e) * sz + e - zrow(e) * nzp] = v; }, nzp * kDirectCols,
      [&](int e) {
        const int k = min(e / kDirectCols, nz - 1);
        return fh[(size_t)k * plane + m0 + e % kDirectCols];
      },
      [&](int e, CT v) { S[(e / kDirectCols) * sc + e % kDirectCols] = e / kDirectCols < nz ? v : CT(0); });
  __syncthreads();
  mfma_products<CT>(Wi, sz, S, sc, nzp, kDirectCols, nzp, [&](int r, int c, CT v) {
    Tm[r * sc + c] = r < nz ? v * inv[(size_t)r * plane + m0 + c] : CT(0);
  });
  __syncthreads();
  const auto xo = gp(reinterpret_cast<CT*>(L.dscr)) + (size_t)L.ncell;
  mfma_products<CT>(W, sz, Tm, sc, nzp, kDirectCols, nzp, [&](int r, int c, CT v) {
    if (r < nz) xo[(size_t)r * plane + m0 + c] = v;
  });
}

inline unsigned nb(size_t n, int bs = kBlock) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// Clears the state words of the instances of the items of ll.
__global__ void k_gate_reset(DevLevel* __restrict__ Lt, int nL, LaunchList ll, int only_stopped) {
  for (int l = threadIdx.x; l < nL; l += blockDim.x) {
    bool mine = false;
    for (int t = 0; t < ll.nitems; ++t) mine = mine || Lt[ll.items[t]].inst == Lt[l].inst;
    if (only_stopped && !Lt[l].stop) mine = false;
    if (mine) {
      Lt[l].stop = 0;
      Lt[l].warm = 0;
      Lt[l].skipc = 0;
      Lt[l].pcur = 0;
    }
  }
}

// ---- launchers --------------------------------------------------------------
#define LAUNCH(kern, grid, blk, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(blk), 0, s, __VA_ARGS__)

// Compute units of the current device (cached per device).
static int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cus[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

void launch_rb_march(bool fp64, const DevLevel* L, const LaunchList& ll, int parity, int zero_guess,
                     int planes, hipStream_t s) {
#define RM(T, K) LAUNCH((k_rb_march<T, K>), ll.nblocks, kBlock, L, ll, parity, zero_guess)
  if (fp64) {
    if (planes == 8) RM(double, 8); else if (planes == 4) RM(double, 4); else RM(double, 2);
  } else {
    if (planes == 8) RM(float, 8); else if (planes == 4) RM(float, 4); else RM(float, 2);
  }
#undef RM
}
void launch_rb_point(bool fp64, const DevLevel* L, const LaunchList& ll, int parity, int zero_guess,
                     hipStream_t s) {
  if (fp64) LAUNCH(k_rb_point<double>, ll.nblocks, kBlock, L, ll, parity, zero_guess);
  else LAUNCH(k_rb_point<float>, ll.nblocks, kBlock, L, ll, parity, zero_guess);
}
static void zline_variant(bool fp64, const DevLevel* L, const LaunchList& ll, int parity,
                          int zero_guess, int segs, hipStream_t s, int form, bool nbr, bool half) {
  // half: two segments of 32 columns per wave, segs rounded up to even
  const int wy = half ? (segs + 1) / 2 : segs, ps = half ? 2 * wy : segs, cw = half ? 32 : 64;
  const dim3 blk(64, wy);
  // reduced-system state (two buffers in FP32, one in FP64) + staged rows
  const size_t lds = (size_t)((fp64 ? 6 : 12) * cw + 3 * kZSeg) * ps * (fp64 ? 8 : 4);
#define ZL(T, CH, NBR, LEAN, H)                                                                   \
  hipLaunchKernelGGL((k_zline<T, kZSeg, CH, NBR, LEAN, H>), dim3(ll.nblocks), blk, lds, s, L, ll, \
                     parity, zero_guess)
  // FP64 always takes the lean form (the short one does not fit its
  // registers).  FP32 takes it when there are more blocks than CUs, so that
  // blocks can pair up; a smaller launch is latency-bound and takes the
  // shorter dependent chains of the short form (and, with half, twice the
  // workgroups).
  if (fp64) {
    if (half) {
      if (nbr) ZL(double, kZChunk64, true, false, true);
      else ZL(double, kZChunk64, false, false, true);
    } else {
      if (nbr) ZL(double, kZChunk64, true, true, false);
      else ZL(double, kZChunk64, false, true, false);
    }
  } else if (half) {
    if (nbr) ZL(float, kZChunk32, true, false, true);
    else ZL(float, kZChunk32, false, false, true);
  } else if (form == 2 || (form == 0 && ll.nblocks > device_cus())) {
    if (nbr) ZL(float, kZChunk32, true, true, false);
    else ZL(float, kZChunk32, false, true, false);
  } else {
    if (nbr) ZL(float, kZChunk32, true, false, false);
    else ZL(float, kZChunk32, false, false, false);
  }
#undef ZL
}
void launch_zline(bool fp64, const DevLevel* L, const LaunchList& ll, int parity, int zero_guess,
                  int segs, hipStream_t s, int form) {
  // A launch of at most half as many column groups as the device has CUs
  // (and columns of two segments or more) runs 32 columns per workgroup:
  // every item's block range doubles.  Forced forms keep 64.
  const bool half = form == 0 && segs >= 2 && 2 * ll.nblocks <= device_cus();
  LaunchList hl;
  if (half) {
    hl = ll;
    for (int t = 0; t <= ll.nitems; ++t) hl.prefix[t] = 2 * ll.prefix[t];
    hl.nblocks = 2 * ll.nblocks;
  }
  const LaunchList& l2 = half ? hl : ll;
  // kZgDyn with a zero other colour: the sweep from zero, then the ordinary
  // sweep for the instances whose cycle continues the iterate; each kernel
  // returns at once for the instances of the other
  if ((zero_guess & kZgDyn) && (zero_guess & 1)) {
    zline_variant(fp64, L, l2, parity, zero_guess, segs, s, form, false, half);
    zline_variant(fp64, L, l2, parity, kZgWarm, segs, s, form, true, half);
    return;
  }
  zline_variant(fp64, L, l2, parity, zero_guess, segs, s, form, !(zero_guess & 1), half);
}
void launch_zline_col(bool fp64, const DevLevel* L, const LaunchList& ll, int parity,
                      int zero_guess, hipStream_t s) {
  if (fp64) LAUNCH(k_zline_col<double>, ll.nblocks, kBlock, L, ll, parity, zero_guess);
  else LAUNCH(k_zline_col<float>, ll.nblocks, kBlock, L, ll, parity, zero_guess);
}
void launch_restrict(bool fp64, const DevLevel* L, const DevXfer* X, const LaunchList& ll,
                     hipStream_t s) {
  if (fp64) LAUNCH(k_restrict<double>, ll.nblocks, kBlock, L, X, ll);
  else LAUNCH(k_restrict<float>, ll.nblocks, kBlock, L, X, ll);
}
void launch_restrict222(bool fp64, const DevLevel* L, const DevXfer* X, const LaunchList& ll,
                        hipStream_t s, bool use_skip) {
  if (fp64) LAUNCH(k_restrict222<double>, ll.nblocks, kBlock, L, X, ll, use_skip ? 1 : 0);
  else LAUNCH(k_restrict222<float>, ll.nblocks, kBlock, L, X, ll, use_skip ? 1 : 0);
}
void launch_prolong(bool fp64, const DevLevel* L, const DevXfer* X, const LaunchList& ll,
                    hipStream_t s, bool use_skip) {
  if (fp64) LAUNCH(k_prolong<double>, ll.nblocks, kBlock, L, X, ll, use_skip ? 1 : 0);
  else LAUNCH(k_prolong<float>, ll.nblocks, kBlock, L, X, ll, use_skip ? 1 : 0);
}
size_t direct_lds_limit() {
  static size_t limit = 0;
  if (limit) return limit;
  int dev = 0, v = 0;
  size_t lim = 64 * 1024;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
      v > 0)
    lim = (size_t)v;
  const void* kernels[] = {reinterpret_cast<const void*>(&k_direct_xy<double, double, true>),
                           reinterpret_cast<const void*>(&k_direct_xy<float, float, true>),
                           reinterpret_cast<const void*>(&k_direct_xy<double, double, false>),
                           reinterpret_cast<const void*>(&k_direct_xy<float, float, false>),
                           reinterpret_cast<const void*>(&k_direct_z<double>),
                           reinterpret_cast<const void*>(&k_direct_z<float>)};
  bool ok = true;
  for (const void* k : kernels)
    ok = ok && hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    lim = 64 * 1024;
  }
  limit = lim;
  return limit;
}
size_t direct_lds(int nx, int ny, int nz, bool fp64) {
  const size_t es = fp64 ? sizeof(double) : sizeof(float);
  return std::max(direct_lds_xy(nx, ny, es), direct_lds_z(direct_nzp(nz), es));
}
void launch_direct(bool fp64, const DevLevel* L, const LaunchList& planes, const LaunchList& slabs,
                   size_t lds, hipStream_t s) {
  const dim3 blk(kBlock);
  if (fp64) {
    hipLaunchKernelGGL((k_direct_xy<double, double, true>), dim3(planes.nblocks), blk, lds, s, L, planes);
    hipLaunchKernelGGL(k_direct_z<double>, dim3(slabs.nblocks), blk, lds, s, L, slabs);
    hipLaunchKernelGGL((k_direct_xy<double, double, false>), dim3(planes.nblocks), blk, lds, s, L, planes);
  } else {
    hipLaunchKernelGGL((k_direct_xy<float, float, true>), dim3(planes.nblocks), blk, lds, s, L, planes);
    hipLaunchKernelGGL(k_direct_z<float>, dim3(slabs.nblocks), blk, lds, s, L, slabs);
    hipLaunchKernelGGL((k_direct_xy<float, float, false>), dim3(planes.nblocks), blk, lds, s, L, planes);
  }
}
void launch_coarse(bool fp64, const DevLevel* L, const LaunchList& ll, hipStream_t s) {
  if (fp64) LAUNCH(k_coarse<double>, ll.nblocks, kBlock, L, ll);
  else LAUNCH(k_coarse<float>, ll.nblocks, kBlock, L, ll);
}
int tail_nz_class(bool fp64, int nz_max, bool zline) {
  if (!zline) return 0;
  if (nz_max <= 8) return 8;
  if (nz_max <= 16 || fp64) return 16;
  return 32;
}

#define TAIL_KERNEL(T, NZ, F) (k_tail<T, NZ, F>)
void launch_tail(bool fp64, int nz_class, bool all222, const DevLevel* L, const DevXfer* X,
                 const DevTail* T, const LaunchList& ll, size_t lds_bytes, hipStream_t s, int phase) {
  const dim3 grd(ll.nitems);
#define GO(TT, NZ)                                                                                   \
  do {                                                                                               \
    const dim3 blk(64, NZ == 0 ? kTailWavesPoint : kTailWaves);                                      \
    if (all222) hipLaunchKernelGGL(TAIL_KERNEL(TT, NZ, true), grd, blk, lds_bytes, s, L, X, T, ll, phase); \
    else hipLaunchKernelGGL(TAIL_KERNEL(TT, NZ, false), grd, blk, lds_bytes, s, L, X, T, ll, phase);       \
  } while (0)
  if (fp64) {
    if (nz_class == 0) GO(double, 0); else if (nz_class <= 8) GO(double, 8); else GO(double, 16);
  } else {
    if (nz_class == 0) GO(float, 0);
    else if (nz_class <= 8) GO(float, 8);
    else if (nz_class <= 16) GO(float, 16);
    else GO(float, 32);
  }
#undef GO
}

void launch_coarse_wide(bool fp64, const DevLevel* L, const LaunchList& ll, hipStream_t s) {
  if (fp64) LAUNCH(k_coarse_wide<double>, ll.nblocks, kBlock, L, ll);
  else LAUNCH(k_coarse_wide<float>, ll.nblocks, kBlock, L, ll);
}

size_t tail_lds_limit() {
  static size_t limit = 0;
  if (limit) return limit;
  int dev = 0, v = 0;
  size_t lim = 64 * 1024;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
      v > 0)
    lim = (size_t)v;
  // room for the static descriptor array of the kernel
  lim -= sizeof(DevLevel) * kTailMaxLevels + 256;
  const void* kernels[] = {
      reinterpret_cast<const void*>(&k_tail<double, 0, true>),
      reinterpret_cast<const void*>(&k_tail<double, 0, false>),
      reinterpret_cast<const void*>(&k_tail<float, 0, true>),
      reinterpret_cast<const void*>(&k_tail<float, 0, false>),
      reinterpret_cast<const void*>(&k_tail<double, 8, true>),
      reinterpret_cast<const void*>(&k_tail<double, 8, false>),
      reinterpret_cast<const void*>(&k_tail<double, 16, true>),
      reinterpret_cast<const void*>(&k_tail<double, 16, false>),
      reinterpret_cast<const void*>(&k_tail<float, 8, true>),
      reinterpret_cast<const void*>(&k_tail<float, 8, false>),
      reinterpret_cast<const void*>(&k_tail<float, 16, true>),
      reinterpret_cast<const void*>(&k_tail<float, 16, false>),
      reinterpret_cast<const void*>(&k_tail<float, 32, true>),
      reinterpret_cast<const void*>(&k_tail<float, 32, false>)};
  bool ok = true;
  for (const void* k : kernels)
    ok = ok && hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    lim = 64 * 1024 - sizeof(DevLevel) * kTailMaxLevels - 256;
  }
  limit = lim;
  return limit;
}
// (k_fine takes the table writable: the mixed first defect records bshift)
void launch_residual_max(const DevLevel* L, const LaunchList& ll, double* partial, hipStream_t s) {
  LAUNCH((k_fine<kResMax, false>), ll.nblocks, kBlock, const_cast<DevLevel*>(L), ll, nullptr, partial);
}
void launch_first_defect(bool fp64, const DevLevel* L, const LaunchList& ll, const double* c,
                         double* partial, hipStream_t s) {
  DevLevel* W = const_cast<DevLevel*>(L);
  if (fp64) LAUNCH((k_fine<kResMax, true>), ll.nblocks, kBlock, W, ll, c, partial);
  else LAUNCH((k_fine<kDefect32, true>), ll.nblocks, kBlock, W, ll, c, partial);
}
void launch_ingest(const DevLevel* L, const LaunchList& ll, const PtrList& rhs, const PtrList& phi,
                   double* partial, hipStream_t s) {
  LAUNCH(k_ingest, ll.nblocks, kBlock, L, ll, rhs, phi, partial);
}
void launch_change_max(const DevLevel* L, const LaunchList& ll, double* partial, hipStream_t s) {
  LAUNCH(k_change_max, ll.nblocks, kBlock, L, ll, partial);
}
void launch_correct_defect32(const DevLevel* L, const LaunchList& ll, double* partial,
                             hipStream_t s) {
  LAUNCH((k_fine<kCorrDefect32, false>), ll.nblocks, kBlock, const_cast<DevLevel*>(L), ll, nullptr,
         partial);
}
void launch_zero(double* a, int n, hipStream_t s) {
  LAUNCH(k_zero, nb((size_t)n), kBlock, a, n);
}
void launch_gate_reset(DevLevel* L, int nL, const LaunchList& ll, hipStream_t s, bool only_stopped) {
  hipLaunchKernelGGL(k_gate_reset, dim3(1), dim3(256), 0, s, L, nL, ll, only_stopped ? 1 : 0);
}
void launch_gate_init(const LaunchList& ll, const double* out, ItemGate* ig, DevLevel* L, int nL,
                      int maxcy, int rule, double tol, int report, hipStream_t s) {
  hipLaunchKernelGGL(k_gate_init, dim3(1), dim3(kMaxItems), 0, s, ll, out, ig, L, nL, maxcy, rule,
                     tol, report);
}
void launch_finalize(const LaunchList& ll, const double* partial, int nvals, int op_sum, double* out,
                     int ostride, hipStream_t s, const GateArgs* gate) {
  GateArgs ga{};
  if (gate) ga = *gate;
  LAUNCH(k_finalize, ll.nitems, kFinBlock, ll, partial, nvals, op_sum, out, ostride, ga);
}
void launch_from_fortran(const DevLevel* L, const LaunchList& ll, const PtrList& ptr, int which,
                         int add_fold, hipStream_t s) {
  LAUNCH(k_from_fortran, ll.nblocks, kBlock, L, ll, ptr, which, add_fold);
}
void launch_wsum(const DevLevel* L, const LaunchList& ll, int which, double* partial, hipStream_t s) {
  LAUNCH(k_wsum, ll.nblocks, kBlock, L, ll, which, partial);
}
void launch_to_fortran(const DevLevel* L, const LaunchList& ll, const PtrList& ptr, const double* c,
                       hipStream_t s) {
  LAUNCH(k_to_fortran, ll.nblocks, kBlock, L, ll, ptr, c);
}
void launch_copy_u(const DevLevel* L, const LaunchList& ll, const int* active, hipStream_t s) {
  LAUNCH(k_copy_u, ll.nblocks, kBlock, L, ll, active);
}
void launch_fill_zero(bool fp64, const DevLevel* L, const LaunchList& ll, hipStream_t s) {
  if (fp64) LAUNCH(k_fill_zero<double>, ll.nblocks, kBlock, L, ll);
  else LAUNCH(k_fill_zero<float>, ll.nblocks, kBlock, L, ll);
}
void launch_ghost_wrap(double* a, int nx, int ny, int nz, int dir, hipStream_t s) {
  const size_t n = dir == 0 ? (size_t)(ny + 2) * (nz + 2)
                 : dir == 1 ? (size_t)(nx + 2) * (nz + 2) : (size_t)(nx + 2) * (ny + 2);
  LAUNCH(k_ghost_wrap, nb(n), kBlock, a, nx, ny, nz, dir);
}
void launch_ghost_fill(double* a, int nx, int ny, int nz, const GhostBC& bc, hipStream_t s) {
  const size_t n = 2 * ((size_t)(ny + 2) * (nz + 2) + (size_t)(nx + 2) * (nz + 2) +
                        (size_t)(nx + 2) * (ny + 2));
  LAUNCH(k_ghost_fill, nb(n), kBlock, a, nx, ny, nz, bc);
}
void launch_psource(const double* U, const double* V, const double* W, double* SRC, const double* rho,
                    const double* rdxc, const double* rdyc, const double* rdzc, int nx, int ny, int nz,
                    double rdt, hipStream_t s) {
  LAUNCH(k_psource, nb((size_t)(nx + 2) * (ny + 2) * (nz + 2)), kBlock, U, V, W, SRC, rho, rdxc, rdyc,
         rdzc, nx, ny, nz, rdt);
}
void launch_pgrad(const double* p, double* U, double* V, double* W, const double* rho,
                  const double* rdxe, const double* rdye, const double* rdze, int nx, int ny, int nz,
                  double dt, int wallmask, hipStream_t s) {
  LAUNCH(k_pgrad, nb((size_t)nx * ny * nz), kBlock, p, U, V, W, rho, rdxe, rdye, rdze, nx, ny, nz, dt,
         wallmask);
}
int divergence_blocks(int nx, int ny, int nz) {
  return (int)nb((size_t)nx * ny * nz, kBlock * kDivPer);
}
void launch_divergence(const double* U, const double* V, const double* W, double* div,
                       const double* rdxc, const double* rdyc, const double* rdzc, int nx, int ny,
                       int nz, double* partial, hipStream_t s) {
  LAUNCH(k_divergence, divergence_blocks(nx, ny, nz), kBlock, U, V, W, div, rdxc, rdyc, rdzc, nx, ny,
         nz, partial);
}
void launch_divergence_finalize(const double* partial, int nblocks, double* out, hipStream_t s) {
  LAUNCH(k_divergence_finalize, 1, kDivFin, partial, nblocks, out);
}

}  // namespace amg
