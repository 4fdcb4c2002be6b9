"""Multigrid V-cycle restatement (oracle), interior arrays ``[k, j, i]``.

Restates ``mgMK/mg_cycle.F90`` with the smoothers of ``mgMK/mg_relax.F90``,
the transfers of ``mgMK/mg_restr.F90``/``mgMK/mg_cor.F90`` and the driver of
``mgMK/mg_solver.F90``, on the metric-consistent operator of ``grid``/``ops``
(1-D coefficient tables).  Deliberate differences from the reference, each
listed in the design notes: coarse operators are re-discretised from
agglomerated metrics instead of ``mgdphpde``; the coarsest level is solved
exactly instead of by ``ipre+ipost`` sweeps (``mg_cycle.F90:117-148``);
restriction is volume-weighted (equal to the reference's ``1/8`` sum on
uniform grids); the stop test is the FP64 relative residual (the reference's
``relmax`` is reported beside it).
"""
import numpy as np

from .ops import apply_A, diag_of, volumes, compat_project, null_project, fold_bc_rhs
from .grid import NEUMANN, DIRICHLET, is_singular

RBPGS, LJ = "RBPGS", "LJ"


def _coefs(lev, dt):
    out = {}
    for key in ("cx", "cy", "cz"):
        out[key] = tuple(np.asarray(a, dtype=dt) for a in lev[key][:3])
    return out


def _off(u, lev, co):
    """Off-diagonal part of ``A u`` (all six neighbours)."""
    from .ops import _shift
    wxm, wxp, _ = co["cx"]
    wym, wyp, _ = co["cy"]
    wzm, wzp, _ = co["cz"]
    px, py, pz = lev["per"]
    out = wxm[None, None, :] * _shift(u, 2, -1, px) + wxp[None, None, :] * _shift(u, 2, 1, px)
    out = out + wym[None, :, None] * _shift(u, 1, -1, py) + wyp[None, :, None] * _shift(u, 1, 1, py)
    out = out + wzm[:, None, None] * _shift(u, 0, -1, pz) + wzp[:, None, None] * _shift(u, 0, 1, pz)
    return out


def _diag(co):
    return co["cz"][2][:, None, None] + co["cy"][2][None, :, None] + co["cx"][2][None, None, :]


def colour_mask(shape, parity, line):
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    s = (i + j) if line else (i + j + k)
    return (s % 2) == parity


def rb_half_sweep(u, f, lev, parity, dt=np.float64):
    """One colour of point red-black Gauss-Seidel, in place.

    ``mgMK/mg_relax.F90:146-168``: ``phi = (src - sum c_n phi_n)/c7`` on the
    points with ``(i+j+k) % 2 == parity`` (0-based interior indices; the
    reference's first pass has parity 0).  All their neighbours have the
    other colour, so the update is order-independent within the colour.
    """
    co = _coefs(lev, dt)
    m = colour_mask(u.shape, parity, False)
    new = (f - _off(u, lev, co)) / _diag(co)
    u[m] = new[m].astype(dt)
    return u


def zline_half_sweep(u, f, lev, parity, dt=np.float64):
    """One colour of z-line red-black Gauss-Seidel, in place.

    ``mgMK/mg_relax.F90:423-538`` (``mg_relaxLz``, smoother ``LJ``): on the
    columns with ``(i+j) % 2 == parity`` solve the z-tridiagonal system
    ``wzm u_{k-1} + (dx+dy+dz) u_k + wzp u_{k+1} = f - (x/y neighbour terms)``
    exactly (Thomas, ``:541-595``).  A periodic z wrap couples the column's
    ends; that coupling is lagged (old values), as the reference lags the
    ghost values (``:498-504``).
    """
    co = _coefs(lev, dt)
    nz, ny, nx = u.shape
    wxm, wxp, dx = co["cx"]
    wym, wyp, dy = co["cy"]
    wzm, wzp, dz = co["cz"]
    px, py, pz = lev["per"]
    from .ops import _shift
    g = (f - (wxm[None, None, :] * _shift(u, 2, -1, px) + wxp[None, None, :] * _shift(u, 2, 1, px)
              + wym[None, :, None] * _shift(u, 1, -1, py) + wyp[None, :, None] * _shift(u, 1, 1, py)))
    g = g.astype(dt)
    if pz:
        g[0] = g[0] - wzm[0] * u[nz - 1]
        g[nz - 1] = g[nz - 1] - wzp[nz - 1] * u[0]
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    cm = ((ii + jj) % 2) == parity
    gc = g[:, cm]
    ncol = gc.shape[1]
    a = np.repeat(wzm[:, None], ncol, axis=1).astype(dt)
    c = np.repeat(wzp[:, None], ncol, axis=1).astype(dt)
    if pz:
        a[0] = 0
        c[nz - 1] = 0
    b = (dz[:, None] + (dx[None, :] + dy[:, None])[cm][None, :]).astype(dt)
    from .direct import _thomas
    u[:, cm] = _thomas(a, b, c, gc).astype(dt)
    return u


def smooth(u, f, lev, sweeps, smoother, dt=np.float64):
    hs = zline_half_sweep if smoother == LJ else rb_half_sweep
    for _ in range(sweeps):
        hs(u, f, lev, 0, dt)
        hs(u, f, lev, 1, dt)
    return u


def residual(u, f, lev, dt=np.float64):
    co = _coefs(lev, dt)
    return (f - (_off(u, lev, co) + _diag(co) * u)).astype(dt)


def restrict(r, fine, coarse, factor, dt=np.float64):
    """Volume-weighted restriction of a fine residual to the coarse RHS.

    ``mgMK/mg_restr.F90:151-164`` averages the 8 children
    ``i in {2ic-2, 2ic-1}`` (Fortran, ghost-offset indices) with weight
    ``1/8``; here each child is weighted by its volume fraction, which is
    exactly ``1/8`` on uniform grids.  A direction that does not coarsen
    contributes one child (the reference's 2- and 4-point variants,
    ``:166-240``).
    """
    fx, fy, fz = factor
    V = volumes(fine)
    nxc, nyc, nzc = coarse["n"]
    acc = np.zeros((nzc, nyc, nxc))
    vc = np.zeros((nzc, nyc, nxc))
    for c in range(fz):
        for b in range(fy):
            for a in range(fx):
                sl = (slice(c, None, fz), slice(b, None, fy), slice(a, None, fx))
                acc += (V[sl] * r[sl].astype(np.float64))
                vc += V[sl]
    return (acc / vc).astype(dt)


def _prolong_1d(n_f, n_c, fac, bc_lo, bc_hi, per):
    """Per fine index: list of (coarse index or -1 for zero, weight)."""
    out = []
    for i in range(n_f):
        if fac == 1:
            out.append([(i, 1.0)])
            continue
        I = i // 2
        far = I - 1 if i % 2 == 0 else I + 1
        if far < 0 or far >= n_c:
            wall = bc_lo if far < 0 else bc_hi
            if per:
                far = far % n_c
            elif wall == DIRICHLET:
                far = -1
            else:
                far = I
        out.append([(I, 0.75), (far, 0.25)])
    return out


def prolong_add(u, e_c, fine, coarse, factor, dt=np.float64):
    """Trilinear prolongation-add ``mgMK/mg_cor.F90:102-159``.

    Each fine cell takes ``27,9,9,3,9,3,3,1 /64`` of its 8 nearest coarse
    cells (the product of 1-D weights 3/4 and 1/4).  Out-of-domain coarse
    neighbours are the coarse ghosts after ``mg_bdry`` with ``phibc = 0``:
    periodic wrap, Neumann copy of the interior, Dirichlet zero.
    """
    nx, ny, nz = fine["n"]
    nxc, nyc, nzc = coarse["n"]
    (bxl, bxh), (byl, byh), (bzl, bzh) = fine["bc"]
    px, py, pz = fine["per"]
    wx = _prolong_1d(nx, nxc, factor[0], bxl, bxh, px)
    wy = _prolong_1d(ny, nyc, factor[1], byl, byh, py)
    wz = _prolong_1d(nz, nzc, factor[2], bzl, bzh, pz)

    def mat(w, n_c):
        M = np.zeros((len(w), n_c))
        for i, lst in enumerate(w):
            for I, a in lst:
                if I >= 0:
                    M[i, I] += a
        return M
    Px, Py, Pz = mat(wx, nxc), mat(wy, nyc), mat(wz, nzc)
    ec = e_c.astype(np.float64)
    t = np.einsum("kji,xi->kjx", ec, Px)
    t = np.einsum("kjx,yj->kyx", t, Py)
    t = np.einsum("kyx,zk->zyx", t, Pz)
    return (u.astype(np.float64) + t).astype(dt) if dt == np.float64 else (u + t.astype(dt)).astype(dt)


def coarse_operator_inverse(lev, singular):
    """Dense FP64 inverse used by the coarsest-level solve.

    Singular operators use the bordered matrix ``[[A, 1], [V^T, 0]]``; the
    top-left block of its inverse maps a right-hand side to the solution
    with ``V.x = 0`` (any incompatible part of the RHS is absorbed).
    """
    from .direct import assemble
    A = assemble(lev).toarray()
    N = A.shape[0]
    if singular:
        V = volumes(lev).ravel()
        M = np.zeros((N + 1, N + 1))
        M[:N, :N] = A
        M[:N, N] = 1.0
        M[N, :N] = V
        return np.linalg.inv(M)[:N, :N]
    return np.linalg.inv(A)


def separable_solve(lev, f, singular):
    """Exact solve of ``A x = f`` on one level with non-periodic z (FP64).

    ``A`` is a sum of 1-D operators; ``Ax = Dx^-1 Sx`` (``Dx`` the cell
    widths, ``Sx`` symmetric), so ``Dx^1/2 Ax Dx^-1/2 = Q diag(l) Q^T`` and
    ``Ax = Vx diag(l) Vx^-1`` with ``Vx = Dx^-1/2 Q``; the same in y.  In
    the eigenvector bases every horizontal mode ``(e, g)`` is one tridiagonal
    system ``(Az + lx_e + ly_g) x = f^`` in z.  A singular operator's
    constant mode has its last value pinned to zero, and the result is made
    orthogonal to the null vector ``rho(z)`` (volume weights), as
    ``ops.null_project`` does for the whole level.  This is the coarse
    solve of a direct level (the product's ``k_direct_*``).
    """
    nx, ny, nz = lev["n"]
    px, py, pz = lev["per"]
    assert not pz, "separable_solve needs walls in z"

    def axis(c, dxc, n, per):
        wm, wp, d = (np.asarray(a, dtype=np.float64) for a in c[:3])
        A = np.diag(d)
        for i in range(n):
            if per or i > 0:
                A[i, (i - 1) % n] += wm[i]
            if per or i < n - 1:
                A[i, (i + 1) % n] += wp[i]
        s = np.sqrt(dxc)
        B = s[:, None] * A / s[None, :]
        lam, Q = np.linalg.eigh(0.5 * (B + B.T))
        return lam, Q / s[:, None], Q.T * s[None, :]

    dxc = np.asarray(lev["ax"]["dxc"][1:nx + 1])
    dyc = np.asarray(lev["ay"]["dxc"][1:ny + 1])
    dzc = np.asarray(lev["az"]["dxc"][1:nz + 1])
    lx, Vx, Vxi = axis(lev["cx"], dxc, nx, px)
    ly, Vy, Vyi = axis(lev["cy"], dyc, ny, py)
    sing = None
    if singular:
        ix, iy = int(np.argmin(np.abs(lx))), int(np.argmin(np.abs(ly)))
        lx[ix] = 0.0
        ly[iy] = 0.0
        sing = (iy, ix)
    fh = np.einsum("ei,kji->kje", Vxi, np.asarray(f, dtype=np.float64))
    fh = np.einsum("gj,kje->kge", Vyi, fh)
    wzm, wzp, dz = (np.asarray(a, dtype=np.float64) for a in lev["cz"][:3])
    lam = ly[:, None] + lx[None, :]

    def thomas(g, lm, rows):
        x = np.zeros_like(g)
        cp = np.zeros_like(g)
        y = np.zeros_like(g)
        for k in range(rows):
            a = wzm[k] if k > 0 else 0.0
            c = wzp[k] if k < nz - 1 else 0.0
            r = 1.0 / ((dz[k] + lm) - a * (cp[k - 1] if k > 0 else 0.0))
            y[k] = (g[k] - a * (y[k - 1] if k > 0 else 0.0)) * r
            cp[k] = c * r
        for k in range(rows - 1, -1, -1):
            x[k] = y[k] - (cp[k] * x[k + 1] if k + 1 < nz else 0.0)
        return x

    xh = thomas(fh, lam[None, :, :], nz)
    if sing is not None:
        g0 = fh[:, sing[0], sing[1]]
        x0 = thomas(g0, 0.0, nz - 1)
        rho = np.asarray(lev["rho"][1:nz + 1], dtype=np.float64)
        xh[:, sing[0], sing[1]] = x0 - ((dzc * x0).sum() / (dzc * rho).sum()) * rho
    x = np.einsum("jg,kge->kje", Vy, xh)
    return np.einsum("ie,kje->kji", Vx, x)


def vcycle(levels, factors, u, f, pre, post, smoother, dt, Minv, l=0, zero_guess=False):
    """One V-cycle on level ``l`` (``mgMK/mg_cycle.F90``): pre-smooth, restrict,
    recurse, prolongation-add, post-smooth; exact solve on the coarsest level
    (``Minv`` a dense inverse, or a function solving that level)."""
    lev = levels[l]
    if l == len(levels) - 1:
        if callable(Minv):
            x = Minv(f.astype(np.float64))
        else:
            x = (Minv @ f.astype(np.float64).ravel()).reshape(f.shape)
        u[...] = x.astype(dt)
        return u
    if zero_guess:
        u[...] = 0
    smooth(u, f, lev, pre[l], smoother, dt)
    r = residual(u, f, lev, dt)
    fc = restrict(r, lev, levels[l + 1], factors[l], dt)
    uc = np.zeros_like(fc)
    vcycle(levels, factors, uc, fc, pre, post, smoother, dt, Minv, l + 1, zero_guess=True)
    u[...] = prolong_add(u, uc, lev, levels[l + 1], factors[l], dt)
    smooth(u, f, lev, post[l], smoother, dt)
    return u


def maxabs(a):
    return float(np.max(np.abs(a))) if a.size else 0.0


PAIR_RHO0 = 1e-3


def pair_next(r, rp, m, tol):
    """Whether the next FP64 correction should close two FP32 V-cycles.

    ``r``: the relative residual after the last correction, ``rp`` the one
    before it, ``m`` V-cycles earlier.  The per-cycle rate is ``rho = (r /
    rp)^(1/m)``; two cycles are paired when one cannot reach the tolerance,
    ``r * rho > tol``, decided with multiplications only (``m = 2``:
    ``r^2 (r / rp) > tol^2``).  This is the product's paired-cycle rule
    (its device gate and host replay); it has no counterpart in the
    reference, whose cycles are all FP64.
    """
    if not rp > 0.0:
        return False
    q = r / rp
    return r * q > tol if m == 1 else (r * r) * q > tol * tol


def solve(levels, factors, tbc, vbc, p0, b_in, *, pre=2, post=2, smoother=LJ,
          mixed=False, tol=1e-10, maxcy=200, remove_null=True, direct_level=None, pair=False):
    """Solve to ``||b - A p||_inf / ||b||_inf <= tol`` (driver of ``mgMK/mg_solver.F90``).

    ``mixed=False``: V-cycles on ``p`` in FP64.  ``mixed=True``: FP64 defect
    correction ``r = b - A p`` (FP64), one FP32 V-cycle on ``A e = r`` from
    ``e = 0``, ``p += e`` (FP64).  Returns ``(p, info)``; ``info`` holds the
    cycle count, the residual history and the reference's ``relmax =
    max|p_new - p_old| / max|p_new|`` history (``mgMK/mg_residual.F90:60-94``).
    ``direct_level``: cycle the levels above it and solve it exactly
    (``separable_solve``) instead of going down to the coarsest level.
    ``pair`` (mixed): a correction may close two V-cycles, the second
    continuing from the first's ``e`` (``pair_next``; the first correction
    pairs when ``res0 * PAIR_RHO0 > tol``); the histories then hold one entry
    per correction and ``info["cycles"]`` the V-cycles each closes.
    """
    singular = is_singular(tbc)
    if direct_level is not None:
        levels = levels[:direct_level + 1]
        factors = factors[:direct_level]
    L = len(levels)
    fine = levels[0]
    pre_l = [pre] * L if np.isscalar(pre) else list(pre)
    post_l = [post] * L if np.isscalar(post) else list(post)
    b = fold_bc_rhs(b_in, fine, tbc, vbc)
    if singular:
        b = compat_project(b, fine)
    if direct_level is not None:
        Minv = lambda g: separable_solve(levels[-1], g, singular)  # noqa: E731
    else:
        Minv = coarse_operator_inverse(levels[-1], singular)
    p = np.array(p0, dtype=np.float64, copy=True)
    bn = maxabs(b)
    res_hist, rel_hist, cyc_hist = [], [], []
    r = residual(p, b, fine)
    it = 0
    conv = maxabs(r) <= tol * bn
    rprev = maxabs(r) / bn if bn > 0 else maxabs(r)
    two = pair and mixed and maxcy >= 2 and rprev * PAIR_RHO0 > tol
    while not conv and it < maxcy:
        p_old = p.copy()
        m = 2 if two else 1
        if mixed:
            e = np.zeros(p.shape, dtype=np.float32)
            r32 = r.astype(np.float32)
            for q in range(m):
                vcycle(levels, factors, e, r32, pre_l, post_l, smoother, np.float32, Minv, 0,
                       zero_guess=(q == 0))
            p = p + e.astype(np.float64)
        else:
            vcycle(levels, factors, p, b, pre_l, post_l, smoother, np.float64, Minv, 0)
        it += m
        r = residual(p, b, fine)
        res_hist.append(maxabs(r) / bn if bn > 0 else 0.0)
        cyc_hist.append(m)
        pm = maxabs(p)
        rel_hist.append(maxabs(p - p_old) / pm if pm > 0 else 0.0)
        conv = res_hist[-1] <= tol
        two = (pair and mixed and not conv and it + 2 <= maxcy
               and pair_next(res_hist[-1], rprev, m, tol))
        rprev = res_hist[-1]
    if singular and remove_null:
        p = null_project(p, fine)
    return p, {"iters": it, "residual": res_hist, "relmax": rel_hist, "converged": conv,
               "cycles": cyc_hist, "b": b}
