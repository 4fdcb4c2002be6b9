// amg_kernels.hip -- gfx950 kernels of the multigrid pressure solve.
//
// Every V-cycle kernel runs over a work list: one launch covers the same
// step of several instances/levels (LaunchList in amg_internal.h), so a
// single solve and a batch of uneven solves share the same code.  All
// kernels are HBM-streaming stencils (about 0.5 flop/byte): no MFMA.
#include "amg_internal.h"

#include <cmath>

namespace amg {
namespace {

__device__ __forceinline__ int find_item(const int* prefix, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
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

template <class T> __device__ __forceinline__ const T* coefs(const DevLevel& L);
template <> __device__ __forceinline__ const double* coefs<double>(const DevLevel& L) { return L.co64; }
template <> __device__ __forceinline__ const float* coefs<float>(const DevLevel& L) { return L.co32; }

// NaN-propagating maximum (fmax ignores NaN).
__device__ __forceinline__ double nmax(double a, double b) {
  return (a != a) ? a : ((b != b) ? b : fmax(a, b));
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
__device__ __forceinline__ size_t cidx(int m, int j, int k, int c, int hx, int ny) {
  return ((size_t)(k * ny + j) * 2 + c) * hx + m;
}

// Seven chequer indices of cell (i,j,k): centre, x-, x+, y-, y+, z-, z+.
struct Nb {
  size_t c, xm, xp, ym, yp, zm, zp;
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
  const size_t ro = cidx(0, j, k, oth, hx, ny);
  n.xm = ro + ml;
  n.xp = ro + mr;
  n.ym = cidx(m, jm, k, oth, hx, ny);
  n.yp = cidx(m, jp, k, oth, hx, ny);
  n.zm = cidx(m, j, km, col, hx, ny);
  n.zp = cidx(m, j, kp, col, hx, ny);
  return n;
}

// Decode a chequer element index into (i, j, k).
__device__ __forceinline__ void decode(const DevLevel& L, int e, int& i, int& j, int& k) {
  const int m = e % L.hx;
  const int t = e / L.hx;
  const int c = t & 1;
  const int row = t >> 1;
  j = row % L.ny;
  k = row / L.ny;
  i = 2 * m + ((c + j) & 1);
}

// Off-diagonal stencil sum, in a fixed order: x-, x+, y-, y+, z-, z+.
template <class T>
__device__ __forceinline__ T offsum(const T* __restrict__ co, const DevLevel& L, int i, int j, int k,
                                    T uxm, T uxp, T uym, T uyp, T uzm, T uzp) {
  const int nx = L.nx, ny = L.ny;
  const T* cy = co + 3 * nx;
  const T* cz = cy + 3 * ny;
  T s = co[i] * uxm + co[nx + i] * uxp;
  s += cy[j] * uym;
  s += cy[ny + j] * uyp;
  s += cz[k] * uzm;
  s += cz[L.nz + k] * uzp;
  return s;
}

template <class T>
__device__ __forceinline__ T diag_at(const T* __restrict__ co, const DevLevel& L, int i, int j, int k) {
  const int nx = L.nx, ny = L.ny;
  const T* cy = co + 3 * nx;
  const T* cz = cy + 3 * ny;
  return (cz[2 * L.nz + k] + cy[2 * ny + j]) + co[2 * nx + i];
}

// ---------------------------------------------------------------------------
// Point red-black Gauss-Seidel half-sweep (mgMK/mg_relax.F90:146-168):
// u = (f - sum w_n u_n)/diag on the cells with (i+j+k)&1 == parity.  Element
// e enumerates the cells of the colour; in plane k the colour is column
// colour (parity+k)&1, so consecutive e are one contiguous half-row.
template <class T>
__device__ __forceinline__ void rb_point_elem(const DevLevel& L, int e, int parity, int zero_guess) {
  const int hx = L.hx, ny = L.ny;
  const int m = e % hx, row = e / hx, j = row % ny, k = row / ny;
  const int c = (k + parity) & 1;
  const int i = 2 * m + ((c + j) & 1);
  const T* __restrict__ co = coefs<T>(L);
  T* __restrict__ u = static_cast<T*>(L.u);
  const T* __restrict__ f = static_cast<const T*>(L.f);
  const Nb n = neighbours(L, i, j, k);
  const T d = diag_at<T>(co, L, i, j, k);
  T val;
  if (zero_guess & 1) {
    val = f[n.c] / d;
  } else {
    const T s = offsum<T>(co, L, i, j, k, u[n.xm], u[n.xp], u[n.ym], u[n.yp], u[n.zm], u[n.zp]);
    val = (f[n.c] - s) / d;
  }
  u[n.c] = val;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_rb_point(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                     int parity, int zero_guess) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e < L.hx * L.ny * L.nz) rb_point_elem<T>(L, e, parity, zero_guess);
}

// ---------------------------------------------------------------------------
// z-line red-black Gauss-Seidel half-sweep (mgMK/mg_relax.F90:423-538; the
// reference solves each line with Thomas, :541-595): on every z-column of
// colour (i+j)&1 == parity solve exactly
//     wzm_k u_{k-1} + (dz_k + dx_i + dy_j) u_k + wzp_k u_{k+1} = g_k,
//     g = f - (x/y neighbour terms, other colour).
// A periodic z wrap is lagged (old values), as the reference lags ghosts.
//
// Partitioned solve: the column is cut into segments of S cells; wave p of a
// column group owns segment p of 64 columns (one column per lane, the
// segment's RHS in registers).  Each segment T_p is solved locally for its
// interface data -- the first/last entries of y = T_p^-1 g and of the two
// spikes v = T_p^-1 (a e_1), w = T_p^-1 (c e_S), from one LU (top-down) and
// one UL (bottom-up) sweep without storage -- then wave 0 solves the
// 2x2-block reduced system of segment end values in LDS, and every segment
// re-solves with its end couplings moved to the right-hand side.  All loads
// are contiguous half-rows of the chequer layout; nothing but u is written.
// One column group: 64 columns (lane = column) x P segments (wave p).  Every
// wave of the workgroup must call it together: it synchronises the
// workgroup when a column has more than one segment.  'lds' holds 6*P*64
// values of T for this group; 'act' false makes the group idle.
template <class T, int S>
__device__ __forceinline__ void zline_group(const DevLevel& L, int parity, int zero_guess, int col0,
                                            int p, int lane, int P, T* __restrict__ lds, bool act) {
  const int hx = L.hx, ny = L.ny, nz = L.nz, nx = L.nx;
  const int ncol = hx * ny;
  const bool live = act && col0 < ncol;
  const int col = live ? col0 : ncol - 1;
  const int m = col % hx, j = col / hx;
  const int c = parity, oth = parity ^ 1;
  const int i = 2 * m + ((c + j) & 1), r = i & 1;
  int ml = m - 1 + r; if (ml < 0) ml = hx - 1;
  int mr = m + r;     if (mr >= hx) mr = 0;
  const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
  const T* __restrict__ co = coefs<T>(L);
  const T* cy = co + 3 * nx;
  const T* cz = cy + 3 * ny;
  const T wxl = co[i], wxr = co[nx + i], wys = cy[j], wyn = cy[ny + j];
  const T dxy = co[2 * nx + i] + cy[2 * ny + j];
  T* __restrict__ u = static_cast<T*>(L.u);
  const T* __restrict__ f = static_cast<const T*>(L.f);
  const size_t plane = (size_t)ny * 2 * hx;
  const size_t own = cidx(m, j, 0, c, hx, ny);
  const size_t xo = cidx(0, j, 0, oth, hx, ny);
  const size_t so = cidx(m, jm, 0, oth, hx, ny);
  const size_t no = cidx(m, jp, 0, oth, hx, ny);
  const bool pz = L.pz != 0;
  const int Pi = (nz + S - 1) / S;            // segments of this item (<= blockDim.y)
  const int ks = p * S;
  const int len = max(0, min(S, nz - ks));    // 0 for spare waves of a shorter item
  // Rows of the segment, clamped so that every load below is in bounds and
  // the unrolled loops stay branch-free (all loads in flight at once).
  // A: sub-diagonal, E: super-diagonal, D: diagonal; the lagged periodic
  // wrap couplings are zero in the matrix.
  T A[S], E[S], D[S], g[S];
#pragma unroll
  for (int q = 0; q < S; ++q) {
    const int k = min(ks + q, nz - 1);
    A[q] = (pz && k == 0) ? T(0) : cz[k];
    E[q] = (pz && k == nz - 1) ? T(0) : cz[nz + k];
    D[q] = cz[2 * nz + k] + dxy;
  }
  // zero_guess bit 0: the other colour is zero (first half-sweep from u = 0);
  // bit 1: this colour's old values are zero (no lagged wrap coupling)
  const bool nbr = !(zero_guess & 1);
  const bool lag = pz && !(zero_guess & 2);
  const T lagb = (lag && p == 0) ? cz[0] * u[(size_t)(nz - 1) * plane + own] : T(0);
  const T lagt = (lag && ks + len == nz && len > 0) ? cz[2 * nz - 1] * u[own] : T(0);
#pragma unroll
  for (int q = 0; q < S; ++q) {
    const int k = min(ks + q, nz - 1);
    const size_t kb = (size_t)k * plane;
    T v = f[kb + own];
    if (nbr) {
      T s = wxl * u[kb + xo + ml] + wxr * u[kb + xo + mr];
      s += wys * u[kb + so];
      s += wyn * u[kb + no];
      v = v - s;
    }
    v -= (ks + q == 0) ? lagb : T(0);
    v -= (ks + q == nz - 1) ? lagt : T(0);
    g[q] = (q < len) ? v : T(0);
  }
  const T a_first = A[0];                                  // coupling to u_{ks-1}
  T c_last = T(0);                                         // coupling to u_{ks+len}
#pragma unroll
  for (int q = 0; q < S; ++q) c_last = (q == len - 1) ? E[q] : c_last;
  // phase 1: interface data of segment p
  T y0 = T(0), yl = T(0), v0 = T(0), vl = T(0), w0 = T(0), wl = T(0);
  if (len > 0) {
    // LU sweep: yl, vl, wl, w0
    T cp = T(0), yf = T(0), vf = T(0), prod = T(1);
#pragma unroll
    for (int q = 0; q < S; ++q) {
      if (q < len) {
        const T a = (q == 0) ? T(0) : A[q];
        const T rd = rcp(D[q] - a * cp);
        yf = (g[q] - a * yf) * rd;
        vf = (q == 0) ? a_first * rd : -a * vf * rd;
        if (q == len - 1) {
          wl = c_last * rd;
        } else {
          cp = E[q] * rd;
          prod *= -cp;
        }
      }
    }
    yl = yf;
    vl = vf;
    w0 = wl * prod;
    // UL sweep: y0, v0
    T ri = T(0), h = T(0);
#pragma unroll
    for (int q = S - 1; q >= 0; --q) {
      if (q < len) {
        if (q < len - 1) {
          const T ri2 = rcp(D[q] - E[q] * A[q + 1] * ri);
          h = (g[q] - E[q] * h) * ri2;
          ri = ri2;
        } else {
          ri = rcp(D[q]);
          h = g[q] * ri;
        }
      }
    }
    y0 = h;
    v0 = a_first * ri;
  }
  // phase 2: reduced system X_p = (t_p, b_p) (first/last values of segment p)
  //   t_p + v0_p b_{p-1} + w0_p t_{p+1} = y0_p
  //   b_p + vl_p b_{p-1} + wl_p t_{p+1} = yl_p
  T* Y0 = lds;
  T* YL = Y0 + P * 64;
  T* V0 = YL + P * 64;
  T* VL = V0 + P * 64;
  T* W0 = VL + P * 64;
  T* WL = W0 + P * 64;
  const int sl = p * 64 + lane;
  T bprev = T(0), tnext = T(0);
  if (Pi > 1) {
    Y0[sl] = y0; YL[sl] = yl; V0[sl] = v0; VL[sl] = vl; W0[sl] = w0; WL[sl] = wl;
    __syncthreads();
    if (p == 0) {
      // downward: b_{q-1} = beta - gamma t_q ;  t_q = tau - sigma t_{q+1}
      T beta = T(0), gamma = T(0);
      for (int q = 0; q < Pi; ++q) {
        const int s2 = q * 64 + lane;
        const T dn = T(1) / (T(1) - V0[s2] * gamma);
        const T tau = (Y0[s2] - V0[s2] * beta) * dn;
        const T sigma = W0[s2] * dn;
        const T nb = YL[s2] - VL[s2] * beta + VL[s2] * gamma * tau;
        const T ng = WL[s2] + VL[s2] * gamma * sigma;
        V0[s2] = tau; W0[s2] = sigma; VL[s2] = nb; WL[s2] = ng;
        beta = nb; gamma = ng;
      }
      // upward: t_P = 0
      T tn = T(0);
      for (int q = Pi - 1; q >= 0; --q) {
        const int s2 = q * 64 + lane;
        const T tq = V0[s2] - W0[s2] * tn;
        const T bq = VL[s2] - WL[s2] * tn;
        Y0[s2] = tq;
        YL[s2] = bq;
        tn = tq;
      }
    }
    __syncthreads();
    bprev = (p > 0) ? YL[(p - 1) * 64 + lane] : T(0);
    tnext = (p < Pi - 1) ? Y0[(p + 1) * 64 + lane] : T(0);
    __syncthreads();   // the group's LDS may be rewritten by its next call
  }
  // phase 3: re-solve the segment with its end couplings on the right-hand side
  if (len > 0) {
    T cpa[S];
    T cp = T(0), yf = T(0);
#pragma unroll
    for (int q = 0; q < S; ++q) {
      if (q < len) {
        T rhs = g[q];
        if (q == 0) rhs -= a_first * bprev;
        if (q == len - 1) rhs -= c_last * tnext;
        const T a = (q == 0) ? T(0) : A[q];
        const T rd = rcp(D[q] - a * cp);
        yf = (rhs - a * yf) * rd;
        cp = (q == len - 1) ? T(0) : E[q] * rd;
        g[q] = yf;
        cpa[q] = cp;
      }
    }
    T x = T(0);
#pragma unroll
    for (int q = S - 1; q >= 0; --q) {
      if (q < len) {
        x = g[q] - cpa[q] * x;
        if (live) u[(size_t)(ks + q) * plane + own] = x;
      }
    }
  }
}

template <class T, int S>
__global__ __launch_bounds__(1024) void k_zline(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                int parity, int zero_guess) {
  extern __shared__ unsigned char lds_raw[];
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  zline_group<T, S>(L, parity, zero_guess, (b - ll.prefix[t]) * 64 + threadIdx.x, threadIdx.y,
                    threadIdx.x, blockDim.y, reinterpret_cast<T*>(lds_raw), true);
}

template <class T> struct TailNz;
template <> struct TailNz<float> { static constexpr int value = kTailNz; };
template <> struct TailNz<double> { static constexpr int value = kTailNz64; };

// Short columns (nz <= TailNz) inside the coarse tail: one thread per
// column, the whole column in registers, plain Thomas (the reference's line
// solve, mgMK/mg_relax.F90:541-595); column index col = j*hx + m.
template <class T, int NZ>
__device__ __forceinline__ void zline_column(const DevLevel& L, int parity, int zero_guess, int col) {
  const int hx = L.hx, ny = L.ny, nz = L.nz, nx = L.nx;
  const int m = col % hx, j = col / hx;
  const int c = parity, oth = parity ^ 1;
  const int i = 2 * m + ((c + j) & 1), r = i & 1;
  int ml = m - 1 + r; if (ml < 0) ml = hx - 1;
  int mr = m + r;     if (mr >= hx) mr = 0;
  const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
  const T* __restrict__ co = coefs<T>(L);
  const T* cy = co + 3 * nx;
  const T* cz = cy + 3 * ny;
  const T wxl = co[i], wxr = co[nx + i], wys = cy[j], wyn = cy[ny + j];
  const T dxy = co[2 * nx + i] + cy[2 * ny + j];
  T* __restrict__ u = static_cast<T*>(L.u);
  const T* __restrict__ f = static_cast<const T*>(L.f);
  const size_t plane = (size_t)ny * 2 * hx;
  const size_t own = cidx(m, j, 0, c, hx, ny);
  const size_t xo = cidx(0, j, 0, oth, hx, ny);
  const size_t so = cidx(m, jm, 0, oth, hx, ny);
  const size_t no = cidx(m, jp, 0, oth, hx, ny);
  const bool pz = L.pz != 0;
  const bool lag = pz && !(zero_guess & 2);
  const T lagb = lag ? cz[0] * u[(size_t)(nz - 1) * plane + own] : T(0);
  const T lagt = lag ? cz[2 * nz - 1] * u[own] : T(0);
  T g[NZ], cpa[NZ];
#pragma unroll
  for (int q = 0; q < NZ; ++q) {
    const int k = min(q, nz - 1);
    const size_t kb = (size_t)k * plane;
    T v = f[kb + own];
    if (!(zero_guess & 1)) {
      T s = wxl * u[kb + xo + ml] + wxr * u[kb + xo + mr];
      s += wys * u[kb + so];
      s += wyn * u[kb + no];
      v = v - s;
    }
    v -= (q == 0) ? lagb : T(0);
    v -= (q == nz - 1) ? lagt : T(0);
    g[q] = v;
  }
  T cp = T(0), yf = T(0);
#pragma unroll
  for (int q = 0; q < NZ; ++q) {
    if (q < nz) {
      const T a = (q == 0) ? T(0) : cz[q];
      const T e = (q == nz - 1) ? T(0) : cz[nz + q];
      const T rd = rcp((cz[2 * nz + q] + dxy) - a * cp);
      yf = (g[q] - a * yf) * rd;
      cp = e * rd;
      g[q] = yf;
      cpa[q] = cp;
    }
  }
  T x = T(0);
#pragma unroll
  for (int q = NZ - 1; q >= 0; --q) {
    if (q < nz) {
      x = g[q] - cpa[q] * x;
      u[(size_t)q * plane + own] = x;
    }
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_zline_col(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                      int parity, int zero_guess) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e < L.hx * L.ny) zline_column<T, TailNz<T>::value>(L, parity, zero_guess, e);
}

// ---------------------------------------------------------------------------
// Fused residual + volume-weighted restriction (mgMK/mg_restr.F90:124-164):
// element e is a coarse cell in natural (I fastest) order, so the fine reads
// of a wave are contiguous half-rows; the fine residual is never stored.
template <class T>
__device__ __forceinline__ void restrict_elem(const DevLevel& F, const DevLevel& C, const DevXfer& x,
                                              int e) {
  const int I = e % C.nx, JK = e / C.nx, J = JK % C.ny, K = JK / C.ny;
  const T* __restrict__ co = coefs<T>(F);
  const T* __restrict__ u = static_cast<const T*>(F.u);
  const T* __restrict__ f = static_cast<const T*>(F.f);
  const double* vx = F.vol;
  const double* vy = vx + F.nx;
  const double* vz = vy + F.ny;
  T acc = T(0), vs = T(0);
  for (int cz = 0; cz < x.fz; ++cz) {
    const int k = K * x.fz + cz;
    for (int cy = 0; cy < x.fy; ++cy) {
      const int j = J * x.fy + cy;
      for (int cx = 0; cx < x.fx; ++cx) {
        const int i = I * x.fx + cx;
        const Nb n = neighbours(F, i, j, k);
        const T s = offsum<T>(co, F, i, j, k, u[n.xm], u[n.xp], u[n.ym], u[n.yp], u[n.zm], u[n.zp]);
        const T rr = f[n.c] - (s + diag_at<T>(co, F, i, j, k) * u[n.c]);
        const T v = T(vz[k] * vy[j] * vx[i]);
        acc += v * rr;
        vs += v;
      }
    }
  }
  T* __restrict__ fc = static_cast<T*>(C.f);
  fc[cidx(I >> 1, J, K, (I + J) & 1, C.hx, C.ny)] = acc / vs;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_restrict(const DevLevel* __restrict__ Lt,
                                                     const DevXfer* __restrict__ X, LaunchList ll) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevXfer x = X[ll.items[t]];
  const DevLevel& C = Lt[x.coarse];
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e < C.ncell) restrict_elem<T>(Lt[x.fine], C, x, e);
}

// 1-D prolongation weights of fine index i: coarse (I0, w0), (I1, w1).
__device__ __forceinline__ void pw(int i, int nc, int fac, int per, int bc0, int bc1, int& I0,
                                   int& I1, double& w0, double& w1) {
  if (fac == 1) { I0 = I1 = i; w0 = 1.0; w1 = 0.0; return; }
  I0 = i >> 1;
  int far = (i & 1) ? I0 + 1 : I0 - 1;
  w0 = 0.75; w1 = 0.25;
  if (far < 0 || far >= nc) {
    const int wall = (far < 0) ? bc0 : bc1;
    if (per) far = (far + nc) % nc;
    else if (wall == 2) { far = I0; w1 = 0.0; }
    else far = I0;
  }
  I1 = far;
}

// ---------------------------------------------------------------------------
// Trilinear prolongation-add (mgMK/mg_cor.F90:102-159): fine u += sum of
// 27,9,9,3,9,3,3,1 /64 of the eight nearest coarse cells; coarse ghosts are
// periodic wraps, Neumann copies or Dirichlet zeros.  Element e is a fine
// cell in chequer order (contiguous read-modify-write of the fine field).
template <class T>
__device__ __forceinline__ void prolong_elem(const DevLevel& F, const DevLevel& C, const DevXfer& x,
                                             int e) {
  int i, j, k;
  decode(F, e, i, j, k);
  int I[2], J[2], K[2];
  double wx[2], wy[2], wz[2];
  pw(i, C.nx, x.fx, F.px, F.bcx0, F.bcx1, I[0], I[1], wx[0], wx[1]);
  pw(j, C.ny, x.fy, F.py, F.bcy0, F.bcy1, J[0], J[1], wy[0], wy[1]);
  pw(k, C.nz, x.fz, F.pz, F.bcz0, F.bcz1, K[0], K[1], wz[0], wz[1]);
  const T* __restrict__ uc = static_cast<const T*>(C.u);
  T acc = T(0);
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    T sy = T(0);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const T v0 = uc[cidx(I[0] >> 1, J[q], K[a], (I[0] + J[q]) & 1, C.hx, C.ny)];
      const T v1 = uc[cidx(I[1] >> 1, J[q], K[a], (I[1] + J[q]) & 1, C.hx, C.ny)];
      const T sx = T(wx[0]) * v0 + T(wx[1]) * v1;
      sy += T(wy[q]) * sx;
    }
    acc += T(wz[a]) * sy;
  }
  T* __restrict__ u = static_cast<T*>(F.u);
  u[e] += acc;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_prolong(const DevLevel* __restrict__ Lt,
                                                    const DevXfer* __restrict__ X, LaunchList ll) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevXfer x = X[ll.items[t]];
  const DevLevel& F = Lt[x.fine];
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e < F.ncell) prolong_elem<T>(F, Lt[x.coarse], x, e);
}

// ---------------------------------------------------------------------------
// Coarsest level: u = Minv f with a precomputed FP64 (pseudo-)inverse
// (replaces the ipre+ipost sweeps of mgMK/mg_cycle.F90:117-148).  Minv is
// stored transposed so the reads Minv^T[s][r] of consecutive rows r coalesce.
template <class T>
__device__ __forceinline__ void coarse_row(const DevLevel& L, int r) {
  const int n = L.ncell;
  const T* __restrict__ f = static_cast<const T*>(L.f);
  const double* __restrict__ M = L.minvT + r;
  // 8 independent partial sums so that the loads of a row are in flight together
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int s = 0;
  for (; s + 8 <= n; s += 8) {
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] += M[(size_t)(s + q) * n] * (double)f[s + q];
  }
  for (; s < n; ++s) acc[0] += M[(size_t)s * n] * (double)f[s];
  const double sum = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  static_cast<T*>(L.u)[r] = (T)sum;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_coarse(const DevLevel* __restrict__ Lt, LaunchList ll) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const int r = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (r < L.ncell) coarse_row<T>(L, r);
}

// ---------------------------------------------------------------------------
// Coarse tail: the whole sub-V-cycle below a level (pre-smoothing and
// residual restriction down to the coarsest level, its exact solve, then
// prolongation and post-smoothing back up -- the order of
// mgMK/mg_cycle.F90) for one instance in one workgroup of 64*kTailWaves
// threads, with workgroup barriers between the steps instead of dependent
// launches.  The tail levels' u and f live in LDS (the first level's f is
// copied in, its u copied out); the element routines are the same as in the
// standalone kernels, reached through level descriptors whose field
// pointers point into LDS.  One workgroup per item (DevTail).
template <class T>
__global__ __launch_bounds__(64 * kTailWaves) void k_tail(const DevLevel* __restrict__ Lt,
                                                          const DevXfer* __restrict__ X,
                                                          const DevTail* __restrict__ Tt,
                                                          LaunchList ll) {
  extern __shared__ unsigned char lds_raw[];
  __shared__ DevLevel Ls[kTailMaxLevels];
  const DevTail& tl = Tt[ll.items[blockIdx.x]];
  const int tid = threadIdx.y * 64 + threadIdx.x, nthr = 64 * blockDim.y;
  const int nl = tl.nlev;
  T* buf = reinterpret_cast<T*>(lds_raw);
  if (tid < nl) {
    DevLevel d = Lt[tl.lev[tid]];
    d.u = buf + tl.lds_u[tid];
    d.f = buf + tl.lds_f[tid];
    Ls[tid] = d;
  }
  const DevLevel& G0 = Lt[tl.lev[0]];
  {
    const T* gf = static_cast<const T*>(G0.f);
    const T* gu = static_cast<const T*>(G0.u);
    for (int e = tid; e < G0.ncell; e += nthr) {
      buf[tl.lds_f[0] + e] = gf[e];
      if (tl.copy_u_in) buf[tl.lds_u[0] + e] = gu[e];
    }
  }
  __syncthreads();
  auto half = [&](const DevLevel& L, int parity, int zg) {
    if (tl.zline) {
      const int n = L.hx * L.ny;
      for (int e = tid; e < n; e += nthr) zline_column<T, TailNz<T>::value>(L, parity, zg, e);
    } else {
      const int n = L.hx * L.ny * L.nz;
      for (int e = tid; e < n; e += nthr) rb_point_elem<T>(L, e, parity, zg);
    }
    __syncthreads();
  };
  for (int d = 0; d + 1 < nl; ++d) {
    const DevLevel& L = Ls[d];
    const bool zf = d > 0 || tl.zero_first;
    if (zf && tl.pre[d] == 0) {
      for (int e = tid; e < L.ncell; e += nthr) static_cast<T*>(L.u)[e] = T(0);
      __syncthreads();
    }
    for (int s = 0; s < tl.pre[d]; ++s) {
      half(L, 0, (zf && s == 0) ? 3 : 0);
      half(L, 1, (zf && s == 0) ? 2 : 0);
    }
    const DevXfer x = X[tl.xf[d]];
    const DevLevel& C = Ls[d + 1];
    for (int e = tid; e < C.ncell; e += nthr) restrict_elem<T>(L, C, x, e);
    __syncthreads();
  }
  {
    const DevLevel& L = Ls[nl - 1];
    for (int r = tid; r < L.ncell; r += nthr) coarse_row<T>(L, r);
    __syncthreads();
  }
  for (int d = nl - 2; d >= 0; --d) {
    const DevLevel& L = Ls[d];
    const DevXfer x = X[tl.xf[d]];
    for (int e = tid; e < L.ncell; e += nthr) prolong_elem<T>(L, Ls[d + 1], x, e);
    __syncthreads();
    for (int s = 0; s < tl.post[d]; ++s) {
      half(L, 0, 0);
      half(L, 1, 0);
    }
  }
  T* gu = static_cast<T*>(G0.u);
  for (int e = tid; e < G0.ncell; e += nthr) gu[e] = buf[tl.lds_u[0] + e];
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

// {max|b - A u|, max|b|}
__global__ __launch_bounds__(kBlock) void k_residual_max(const DevLevel* __restrict__ Lt,
                                                         LaunchList ll,
                                                         double* __restrict__ partial) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const double* __restrict__ co = L.co64;
  const double* __restrict__ u = static_cast<const double*>(L.u);
  const double* __restrict__ f = static_cast<const double*>(L.f);
  fine_loop<2>(L, b - ll.prefix[t], partial, b, [&](int e, double* v) {
    int i, j, k;
    decode(L, e, i, j, k);
    const Nb n = neighbours(L, i, j, k);
    const double s = offsum<double>(co, L, i, j, k, u[n.xm], u[n.xp], u[n.ym], u[n.yp], u[n.zm], u[n.zp]);
    v[0] = nmax(v[0], fabs(f[n.c] - (s + diag_at<double>(co, L, i, j, k) * u[n.c])));
    v[1] = nmax(v[1], fabs(f[n.c]));
  });
}

// {max|u - u2|, max|u|}
__global__ __launch_bounds__(kBlock) void k_change_max(const DevLevel* __restrict__ Lt,
                                                       LaunchList ll, double* __restrict__ partial) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const double* __restrict__ u = static_cast<const double*>(L.u);
  const double* __restrict__ u2 = static_cast<const double*>(L.u2);
  fine_loop<2>(L, b - ll.prefix[t], partial, b, [&](int e, double* v) {
    v[0] = nmax(v[0], fabs(u[e] - u2[e]));
    v[1] = nmax(v[1], fabs(u[e]));
  });
}

// r = b - A u (FP64) -> FP32 right-hand side of the correction level; {max|r|, max|b|}
__global__ __launch_bounds__(kBlock) void k_defect32(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                     double* __restrict__ partial) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const double* __restrict__ co = L.co64;
  const double* __restrict__ u = static_cast<const double*>(L.u);
  const double* __restrict__ f = static_cast<const double*>(L.f);
  float* __restrict__ r32 = static_cast<float*>(Lt[L.link].f);
  fine_loop<2>(L, b - ll.prefix[t], partial, b, [&](int e, double* v) {
    int i, j, k;
    decode(L, e, i, j, k);
    const Nb n = neighbours(L, i, j, k);
    const double s = offsum<double>(co, L, i, j, k, u[n.xm], u[n.xp], u[n.ym], u[n.yp], u[n.zm], u[n.zp]);
    const double r = f[n.c] - (s + diag_at<double>(co, L, i, j, k) * u[n.c]);
    r32[e] = (float)r;
    v[0] = nmax(v[0], fabs(r));
    v[1] = nmax(v[1], fabs(f[n.c]));
  });
}

// p_out = p_in + e, then r = b - A p_out from p_in and e at the seven points
// (p_in is never written, so neighbouring threads read consistent values);
// {max|r|, max|e|, max|p_out|}
__global__ __launch_bounds__(kBlock) void k_correct_defect32(const DevLevel* __restrict__ Lt,
                                                             LaunchList ll,
                                                             double* __restrict__ partial) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const DevLevel& S = Lt[L.link];
  const double* __restrict__ co = L.co64;
  const double* __restrict__ p = static_cast<const double*>(L.u);
  const float* __restrict__ ev = static_cast<const float*>(S.u);
  const double* __restrict__ bb = static_cast<const double*>(L.f);
  double* __restrict__ pout = static_cast<double*>(L.u2);
  float* __restrict__ r32 = static_cast<float*>(S.f);
  fine_loop<3>(L, b - ll.prefix[t], partial, b, [&](int e, double* v) {
    int i, j, k;
    decode(L, e, i, j, k);
    const Nb n = neighbours(L, i, j, k);
#define PN(q) (p[n.q] + (double)ev[n.q])
    const double pc = PN(c);
    const double s = offsum<double>(co, L, i, j, k, PN(xm), PN(xp), PN(ym), PN(yp), PN(zm), PN(zp));
#undef PN
    const double r = bb[n.c] - (s + diag_at<double>(co, L, i, j, k) * pc);
    pout[n.c] = pc;
    r32[n.c] = (float)r;
    v[0] = nmax(v[0], fabs(r));
    v[1] = nmax(v[1], fabs((double)ev[n.c]));
    v[2] = nmax(v[2], fabs(pc));
  });
}

// One block per item: thread t folds partials b0+t, b0+t+256, ... in order,
// then a fixed-shape tree combines the threads (deterministic).
__global__ __launch_bounds__(kBlock) void k_finalize(LaunchList ll, const double* __restrict__ partial,
                                                     int nvals, int op_sum, double* __restrict__ out,
                                                     int ostride) {
  __shared__ double sm[kBlock / 64];
  const int t = blockIdx.x;
  const int b0 = ll.prefix[t], b1 = ll.prefix[t + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int q = 0; q < nvals; ++q) {
    double a = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += kBlock) {
      const double x = partial[(size_t)b * nvals + q];
      a = op_sum ? a + x : nmax(a, x);
    }
    a = op_sum ? wave_sum(a) : wave_max(a);
    if (lane == 0) sm[w] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = sm[0];
      for (int i = 1; i < kBlock / 64; ++i) r = op_sum ? r + sm[i] : nmax(r, sm[i]);
      out[(size_t)t * ostride + q] = r;
    }
    __syncthreads();
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
  double v = ptr.p[t][fidx(i, j, k, L.nx, L.ny)];
  if (add_fold) {
    const double* bf = L.bcfold;
    v += (bf[i] + bf[L.nx + j]) + bf[L.nx + L.ny + k];
  }
  static_cast<double*>(which == 0 ? L.u : L.f)[e] = v;
}

// Weighted sums: each thread folds kWsumPer cells (in a fixed order), one
// partial per block.
__global__ __launch_bounds__(kBlock) void k_wsum(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                 int which, double* __restrict__ partial) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const int e0 = (b - ll.prefix[t]) * kBlock * kWsumPer + threadIdx.x;
  double v[1] = {0.0};
  for (int s = 0; s < kWsumPer; ++s) {
    const int e = e0 + s * kBlock;
    if (e < L.ncell) {
      int i, j, k;
      decode(L, e, i, j, k);
      const double w = L.vol[L.nx + L.ny + k] * L.vol[L.nx + j] * L.vol[i];
      const double x = (which == 2) ? L.rho[k]
                                    : static_cast<const double*>(which == 0 ? L.u : L.f)[e];
      v[0] += w * x;
    }
  }
  block_reduce_store<1, true>(v, partial + b);
}

__global__ void k_zero(double* __restrict__ a, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_shift(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                  int which, const double* __restrict__ c,
                                                  int with_rho) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e >= L.ncell) return;
  double* x = static_cast<double*>(which == 0 ? L.u : L.f);
  double s = c[t] / (with_rho ? L.vrsum : L.vsum);
  if (with_rho) {
    int i, j, k;
    decode(L, e, i, j, k);
    s *= L.rho[k];
  }
  x[e] -= s;
}

__global__ __launch_bounds__(kBlock) void k_to_fortran(const DevLevel* __restrict__ Lt,
                                                       LaunchList ll, PtrList ptr) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e >= L.ncell) return;
  int i, j, k;
  decode(L, e, i, j, k);
  ptr.p[t][fidx(i, j, k, L.nx, L.ny)] = static_cast<const double*>(L.u)[e];
}

__global__ __launch_bounds__(kBlock) void k_copy_u(const DevLevel* __restrict__ Lt, LaunchList ll,
                                                   const int* __restrict__ active) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  if (active && !active[L.inst]) return;
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e >= L.ncell) return;
  static_cast<double*>(L.u2)[e] = static_cast<const double*>(L.u)[e];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_fill_zero(const DevLevel* __restrict__ Lt, LaunchList ll) {
  const int b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = find_item(ll.prefix, ll.nitems, b);
  const DevLevel& L = Lt[ll.items[t]];
  const int e = (b - ll.prefix[t]) * kBlock + threadIdx.x;
  if (e < L.ncell) static_cast<T*>(L.u)[e] = T(0);
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

// face: 1 X2max, 2 X3min, 3 X2min, 4 X3max, 5 X1min, 6 X1max; type 1 Neumann, 2 Dirichlet.
__global__ void k_ghost_face(double* __restrict__ a, int nx, int ny, int nz, int face, int type,
                             double v, double dedge) {
  const int sx = nx + 2, sy = ny + 2, sz = nz + 2;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  auto at = [&](int i, int j, int k) -> double& { return a[((size_t)k * sy + j) * sx + i]; };
  if (face == 1 || face == 3) {
    if (e >= sx * sz) return;
    const int i = e % sx, k = e / sx;
    if (face == 1) at(i, ny + 1, k) = (type == 1) ? at(i, ny, k) + v * dedge : v;
    else at(i, 0, k) = (type == 1) ? at(i, 1, k) - v * dedge : v;
  } else if (face == 2 || face == 4) {
    if (e >= sx * sy) return;
    const int i = e % sx, j = e / sx;
    if (face == 4) at(i, j, nz + 1) = (type == 1) ? at(i, j, nz) + v * dedge : v;
    else at(i, j, 0) = (type == 1) ? at(i, j, 1) - v * dedge : v;
  } else {
    if (e >= sy * sz) return;
    const int j = e % sy, k = e / sy;
    if (face == 6) at(nx + 1, j, k) = (type == 1) ? at(nx, j, k) + v * dedge : v;
    else at(0, j, k) = (type == 1) ? at(1, j, k) - v * dedge : v;
  }
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
__global__ __launch_bounds__(kBlock) void k_divergence(const double* __restrict__ U,
                                                       const double* __restrict__ V,
                                                       const double* __restrict__ W,
                                                       double* __restrict__ div,
                                                       const double* __restrict__ rdxc,
                                                       const double* __restrict__ rdyc,
                                                       const double* __restrict__ rdzc, int nx,
                                                       int ny, int nz, double* __restrict__ partial) {
  const int e0 = blockIdx.x * kBlock + threadIdx.x;
  double s[2] = {0.0, 0.0};
  double best = -1.0;
  long long bidx = (long long)1 << 62;
  if (e0 < nx * ny * nz) {
    const int i = e0 % nx + 1, jk = e0 / nx, j = jk % ny + 1, k = jk / ny + 1;
    const int sx = nx + 2, sy = ny + 2;
    const size_t sxy = (size_t)sx * sy;
    const size_t e = (size_t)k * sxy + (size_t)j * sx + i;
    const double d = (U[e] - U[e - 1]) * rdxc[i] + (V[e] - V[e - sx]) * rdyc[j] +
                     (W[e] - W[e - sxy]) * rdzc[k];
    if (div) div[e] = d;
    const double vinv = rdxc[i] * rdyc[j] * rdzc[k];
    s[0] = fabs(d) / vinv;
    s[1] = 1.0 / vinv;
    best = fabs(d);
    bidx = e0;
  }
  // max with the smallest index on ties (the reference's first-found order)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_xor(best, off, 64);
    const long long oi = __shfl_xor(bidx, off, 64);
    if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
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
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q)
      if (sb[q] > b || (sb[q] == b && si[q] < bi)) { b = sb[q]; bi = si[q]; }
    out[2] = b;
    out[3] = (double)bi;
  }
}

__global__ void k_divergence_finalize(const double* __restrict__ partial, int nb,
                                      double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double s0 = 0.0, s1 = 0.0, b = -1.0, bi = 0.0;
  for (int q = 0; q < nb; ++q) {
    const double* p = partial + (size_t)q * 4;
    s0 += p[0];
    s1 += p[1];
    if (p[2] > b || (p[2] == b && p[3] < bi)) { b = p[2]; bi = p[3]; }
  }
  out[0] = s0; out[1] = s1; out[2] = b; out[3] = bi;
}

inline unsigned nb(size_t n, int bs = kBlock) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ---- launchers --------------------------------------------------------------
#define LAUNCH(kern, grid, blk, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(blk), 0, s, __VA_ARGS__)

void launch_rb_point(bool fp64, const DevLevel* L, const LaunchList& ll, int parity, int zero_guess,
                     hipStream_t s) {
  if (fp64) LAUNCH(k_rb_point<double>, ll.nblocks, kBlock, L, ll, parity, zero_guess);
  else LAUNCH(k_rb_point<float>, ll.nblocks, kBlock, L, ll, parity, zero_guess);
}
void launch_zline(bool fp64, const DevLevel* L, const LaunchList& ll, int parity, int zero_guess,
                  int segs, hipStream_t s) {
  const dim3 blk(64, segs);
  const size_t lds = (size_t)6 * segs * 64 * (fp64 ? 8 : 4);
  if (fp64)
    hipLaunchKernelGGL((k_zline<double, kZSeg>), dim3(ll.nblocks), blk, lds, s, L, ll, parity, zero_guess);
  else
    hipLaunchKernelGGL((k_zline<float, kZSeg>), dim3(ll.nblocks), blk, lds, s, L, ll, parity, zero_guess);
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
void launch_prolong(bool fp64, const DevLevel* L, const DevXfer* X, const LaunchList& ll,
                    hipStream_t s) {
  if (fp64) LAUNCH(k_prolong<double>, ll.nblocks, kBlock, L, X, ll);
  else LAUNCH(k_prolong<float>, ll.nblocks, kBlock, L, X, ll);
}
void launch_coarse(bool fp64, const DevLevel* L, const LaunchList& ll, hipStream_t s) {
  if (fp64) LAUNCH(k_coarse<double>, ll.nblocks, kBlock, L, ll);
  else LAUNCH(k_coarse<float>, ll.nblocks, kBlock, L, ll);
}
void launch_tail(bool fp64, const DevLevel* L, const DevXfer* X, const DevTail* T,
                 const LaunchList& ll, size_t lds_bytes, hipStream_t s) {
  const dim3 blk(64, kTailWaves);
  if (fp64) hipLaunchKernelGGL(k_tail<double>, dim3(ll.nitems), blk, lds_bytes, s, L, X, T, ll);
  else hipLaunchKernelGGL(k_tail<float>, dim3(ll.nitems), blk, lds_bytes, s, L, X, T, ll);
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
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tail<double>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tail<float>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim) != hipSuccess) {
    (void)hipGetLastError();
    lim = 64 * 1024 - sizeof(DevLevel) * kTailMaxLevels - 256;
  }
  limit = lim;
  return limit;
}
void launch_residual_max(const DevLevel* L, const LaunchList& ll, double* partial, hipStream_t s) {
  LAUNCH(k_residual_max, ll.nblocks, kBlock, L, ll, partial);
}
void launch_change_max(const DevLevel* L, const LaunchList& ll, double* partial, hipStream_t s) {
  LAUNCH(k_change_max, ll.nblocks, kBlock, L, ll, partial);
}
void launch_defect32(const DevLevel* L, const LaunchList& ll, double* partial, hipStream_t s) {
  LAUNCH(k_defect32, ll.nblocks, kBlock, L, ll, partial);
}
void launch_correct_defect32(const DevLevel* L, const LaunchList& ll, double* partial,
                             hipStream_t s) {
  LAUNCH(k_correct_defect32, ll.nblocks, kBlock, L, ll, partial);
}
void launch_zero(double* a, int n, hipStream_t s) {
  LAUNCH(k_zero, nb((size_t)n), kBlock, a, n);
}
void launch_finalize(const LaunchList& ll, const double* partial, int nvals, int op_sum, double* out,
                     int ostride, hipStream_t s) {
  LAUNCH(k_finalize, ll.nitems, kBlock, ll, partial, nvals, op_sum, out, ostride);
}
void launch_from_fortran(const DevLevel* L, const LaunchList& ll, const PtrList& ptr, int which,
                         int add_fold, hipStream_t s) {
  LAUNCH(k_from_fortran, ll.nblocks, kBlock, L, ll, ptr, which, add_fold);
}
void launch_wsum(const DevLevel* L, const LaunchList& ll, int which, double* partial, hipStream_t s) {
  LAUNCH(k_wsum, ll.nblocks, kBlock, L, ll, which, partial);
}
void launch_shift(const DevLevel* L, const LaunchList& ll, int which, const double* c, int with_rho,
                  hipStream_t s) {
  LAUNCH(k_shift, ll.nblocks, kBlock, L, ll, which, c, with_rho);
}
void launch_to_fortran(const DevLevel* L, const LaunchList& ll, const PtrList& ptr, hipStream_t s) {
  LAUNCH(k_to_fortran, ll.nblocks, kBlock, L, ll, ptr);
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
void launch_ghost_face(double* a, int nx, int ny, int nz, int face, int type, double v,
                       double dedge, hipStream_t s) {
  const size_t n = (face == 1 || face == 3) ? (size_t)(nx + 2) * (nz + 2)
                 : (face == 2 || face == 4) ? (size_t)(nx + 2) * (ny + 2)
                                            : (size_t)(ny + 2) * (nz + 2);
  LAUNCH(k_ghost_face, nb(n), kBlock, a, nx, ny, nz, face, type, v, dedge);
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
int divergence_blocks(int nx, int ny, int nz) { return (int)nb((size_t)nx * ny * nz); }
void launch_divergence(const double* U, const double* V, const double* W, double* div,
                       const double* rdxc, const double* rdyc, const double* rdzc, int nx, int ny,
                       int nz, double* partial, hipStream_t s) {
  LAUNCH(k_divergence, divergence_blocks(nx, ny, nz), kBlock, U, V, W, div, rdxc, rdyc, rdzc, nx, ny,
         nz, partial);
}
void launch_divergence_finalize(const double* partial, int nblocks, double* out, hipStream_t s) {
  LAUNCH(k_divergence_finalize, 1, 64, partial, nblocks, out);
}

}  // namespace amg
