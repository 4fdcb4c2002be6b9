"""Pins for ``oracle.dns`` (the time stepper's ``ghost``/``bound_update`` and
the implicit vertical viscous step ``var_advance_CN_PL``).

The reference holds no golden vectors for these routines, so they are pinned
by what the mathematics fixes:

* ``ghost``: idempotence for every type (a plane rule that read its own ghost
  would fail it), each rule's meaning (Dirichlet ghost == value, Neumann
  difference == value x metric, Wall6 mean == value, periodic images), and a
  scalar-loop restatement of ``dns/BC.F90`` on tiny grids that checks the
  edge and corner ordering of the numpy version.
* ``cn_advance``: one Crank-Nicolson step multiplies each discrete
  eigenmode of the vertical diffusion operator by the closed-form factor
  ``(1 + h lam)/(1 - h lam)`` (h = rkdt/2) -- Neumann and Dirichlet rows,
  u/v, w, rho (x rPr) and the scalar (x rSc); a linear profile whose wall
  values are its own extension is a steady state on a *stretched* grid
  (pins which of dze/dzc sits where); the explicit tendency and the
  predicted-pressure term enter linearly; and a scalar-loop restatement of
  the reference lines on a tiny grid with a varying rho(z) checks the
  rho-weighted coefficients, including the reference's ordering of the two
  averaged densities (``dns/NSsolver.F90:3142-3145``).
"""
import itertools
import math

import numpy as np
import pytest

from oracle import dns
from oracle import grid as G


def _axes(nx, ny, nz, zaxis=None):
    ax = G.uniform_axis(nx, 2 * math.pi)
    ay = G.uniform_axis(ny, 2 * math.pi)
    az = zaxis if zaxis is not None else G.uniform_axis(nz, 1.0)
    return ax, ay, az


# ---- ghost ----------------------------------------------------------------

def _literal_ghost(var, vtype, TB, VB, axes):
    """dns/BC.F90 ghost + bound_update as scalar loops over Fortran indices."""
    a = var.copy()
    nz, ny, nx = (s - 2 for s in a.shape)
    sx, ex, sy, ey, sz, ez = 2, nx + 1, 2, ny + 1, 2, nz + 1

    def get(i, j, k):
        return a[k - 1, j - 1, i - 1]

    def put(i, j, k, v):
        a[k - 1, j - 1, i - 1] = v

    I, J, K = range(1, nx + 3), range(1, ny + 3), range(1, nz + 3)
    for j, k in itertools.product(J, K):
        put(sx - 1, j, k, get(ex, j, k))
        put(ex + 1, j, k, get(sx, j, k))
    for i, k in itertools.product(I, K):
        put(i, sy - 1, k, get(i, ey, k))
        put(i, ey + 1, k, get(i, sy, k))
    for i, j in itertools.product(I, J):
        put(i, j, sz - 1, get(i, j, ez))
        put(i, j, ez + 1, get(i, j, sz))
    l, rt = dns.GHOST_CLASS[vtype]
    if vtype not in dns.BOUND_METRIC:
        return a
    mx, my, mz, of1, of2 = dns.BOUND_METRIC[vtype]
    met = [axes[0][mx], axes[1][my], axes[2][mz]]
    lo = [sx, sy, sz]
    hi = [ex, ey, ez]

    def at(d, n, p, q):
        """index triple with position n along direction d, p/q along the others"""
        idx = [None, None, None]
        idx[d] = n
        others = [x for x in range(3) if x != d]
        idx[others[0]], idx[others[1]] = p, q
        return idx

    ranges = [I, J, K]
    for d in range(3):
        for m in (1, 2):
            if not (TB[d, 0, l - 1] != 3 or m == 1):
                continue
            bt, val = TB[d, m - 1, l - 1], VB[d, m - 1, l - 1]
            dt = lambda f: met[d][f - 1]
            s, e = lo[d], hi[d]
            others = [x for x in range(3) if x != d]
            for p, q in itertools.product(ranges[others[0]], ranges[others[1]]):
                g = lambda n: get(*at(d, n, p, q))
                w = lambda n, v: put(*at(d, n, p, q), v)
                if bt == 1:
                    w(s - 1 if m == 1 else e + 1, val)
                elif bt == 2:
                    if m == 1:
                        if rt == 1:
                            w(s, g(s + 1) - val * dt(s + 1 + of1))
                        w(s - 1, g(s) - val * dt(s + of1))
                    else:
                        if rt == 1:
                            w(e, g(e - 1) + val * dt(e - 1 + of2))
                        w(e + 1, g(e) + val * dt(e + of2))
                elif bt == 3:
                    w(s - 1, g(e))
                    w(e + 1, g(s))
                elif bt == 4:
                    if m == 1:
                        w(s - 1, g(s) - val * dt(s + of1))
                    else:
                        w(e + 1, g(e) + val * dt(e + of2))
                elif bt == 8 and d == 0:
                    if m == 1:
                        w(s - 1, 2.0 * g(s) - g(s + 1))
                    else:
                        w(e + 1, 2.0 * g(e) - g(e - 1))
                elif bt == 6 and d == 2:
                    w(s - 1 if m == 1 else e + 1, 2.0 * val - 1.0 * g(s if m == 1 else e))
                elif bt == 7 and d == 2:
                    if m == 1:
                        w(s, 0.0)
                        w(s - 1, -1.0 * g(s + 1))
                    else:
                        w(e, 0.0)
                        w(e + 1, -1.0 * g(e - 1))
    return a


def _bc_cases():
    """(TB, VB) tables covering every rule of bound_update."""
    rng = np.random.Generator(np.random.PCG64(11))
    out = []
    for xt, yt, zt in [(3, 3, 2), (3, 3, 1), (2, 3, 6), (8, 1, 7), (1, 4, 4), (4, 2, 3), (8, 2, 2)]:
        TB = np.zeros((3, 2, 12), dtype=np.int32)
        VB = rng.standard_normal((3, 2, 12))
        for d, t in enumerate((xt, yt, zt)):
            TB[d, :, :] = t
        if zt == 7:
            TB[2, 0, :] = 2          # Wall7 is a top-only rule in var_advance_CN; ghost knows both faces
        out.append((TB, VB))
    # one face periodic and the other not: face 1's type decides for both
    TB = np.full((3, 2, 12), 3, dtype=np.int32)
    TB[0, 1, :] = 1
    TB[2, 0, :] = 2
    TB[2, 1, :] = 3
    out.append((TB, rng.standard_normal((3, 2, 12))))
    return out


@pytest.mark.parametrize("vtype", sorted(dns.GHOST_CLASS))
def test_ghost_matches_literal_loops_and_is_idempotent(vtype):
    axes = _axes(5, 4, 6, G.tanh_axis(6, 1.0, 1.5))
    rng = np.random.Generator(np.random.PCG64(3))
    for TB, VB in _bc_cases():
        v = rng.standard_normal((8, 6, 7))
        got = dns.ghost(v.copy(), vtype, TB, VB, axes)
        assert np.array_equal(got, _literal_ghost(v, vtype, TB, VB, axes)), (vtype, TB[:, 0, 0])
        assert np.array_equal(dns.ghost(got.copy(), vtype, TB, VB, axes), got), vtype


def test_ghost_rule_meanings():
    axes = _axes(6, 5, 7, G.tanh_axis(7, 1.0, 2.0))
    az = axes[2]
    rng = np.random.Generator(np.random.PCG64(4))
    v = rng.standard_normal((9, 7, 8))
    TB, VB = dns.flow_bc((True, True, False), {(1, 1): (2, 0.3), (2, 1): (1, -0.7),
                                               (1, 3): (6, 0.25), (2, 3): (4, 1.5)})
    u = dns.ghost(v.copy(), "u", TB, VB, axes)
    # periodic images in x and y (the z ghosts of every column follow the z rules)
    assert np.array_equal(u[:, :, 0], u[:, :, 6]) and np.array_equal(u[:, :, 7], u[:, :, 1])
    assert np.array_equal(u[:, 0, :], u[:, 5, :]) and np.array_equal(u[:, 6, :], u[:, 1, :])
    # Neumann bottom: (u_sz - u_sz-1) = value * dze(sz) for u; Dirichlet top: ghost = value
    np.testing.assert_allclose(u[1] - u[0], 0.3 * az["dxe"][1], rtol=0, atol=1e-15)
    assert np.all(u[8] == -0.7)
    w = dns.ghost(v.copy(), "w", TB, VB, axes)
    # Wall6 bottom: the mean of ghost and first value is the wall value; type 4 top uses dzc for w
    np.testing.assert_allclose(0.5 * (w[0] + w[1]), 0.25, rtol=0, atol=1e-15)
    np.testing.assert_allclose(w[8] - w[7], 1.5 * az["dxc"][8], rtol=0, atol=1e-14)
    # the interior is never touched by non-rtype rules
    assert np.array_equal(u[1:-1, 1:-1, 1:-1], v[1:-1, 1:-1, 1:-1])
    # face 1 periodic, face 2 Dirichlet: only face 1's rule runs (dns/BC.F90:100-105),
    # so the upper ghost keeps its periodic image instead of the value
    TB2, VB2 = dns.flow_bc((False, True, False), {(1, 1): (2, 0.0), (2, 1): (2, 0.0)})
    TB2[0, 0, 0], TB2[0, 1, 0], VB2[0, 1, 0] = 3, 1, 5.0
    x = dns.ghost(v.copy(), "u", TB2, VB2, axes)
    assert np.array_equal(x[:, :, 7], x[:, :, 1]) and not np.any(x[:, :, 7] == 5.0)


# ---- var_advance_CN_PL ----------------------------------------------------------

def _cn(v1, v2, var, TB, VB, axes, rho=None, rkdt=0.01, f1=1.0, f2=0.0, f3=1.0, rRe=0.5,
        rPr=0.7, rSc=1.3, p=None):
    nz = v2.shape[0] - 2
    r = np.ones(nz + 2) if rho is None else rho
    return dns.cn_advance(v1, v2, f1, f2, f3, rkdt, var, r, axes, TB, VB, rRe, rPr, rSc, p=p)


@pytest.mark.parametrize("var,bc,factor", [(1, 2, 1.0), (1, 1, 1.0), (2, 1, 1.0), (3, 1, 1.0),
                                           (5, 1, 0.7), (8, 1, 1.3)])
def test_cn_eigenmode_amplification(var, bc, factor):
    nx, ny, nz = 4, 3, 12
    axes = _axes(nx, ny, nz)
    dz = 1.0 / nz
    TB, VB = dns.flow_bc((True, True, False), {(f, var): (bc, 0.0) for f in (1, 2)})
    rkdt, rRe = 0.02, 0.5
    k = np.arange(nz + 2)
    amp = np.random.Generator(np.random.PCG64(1)).standard_normal((ny + 2, nx + 2))
    for m in (1, 2, 5):
        if bc == 2:      # Neumann rows: ghost == first value; cos modes at cell centres
            phi = np.cos(math.pi * m * (k - 0.5) / nz)
            lam = -4.0 / dz ** 2 * math.sin(math.pi * m / (2 * nz)) ** 2
        else:            # Dirichlet rows: the ghost value is 0; sin modes over nz+1 intervals
            phi = np.sin(math.pi * m * k / (nz + 1))
            lam = -4.0 / dz ** 2 * math.sin(math.pi * m / (2 * (nz + 1))) ** 2
        lam *= rRe * factor
        v2 = phi[:, None, None] * amp[None, :, :]
        out = _cn(np.zeros_like(v2), v2, var, TB, VB, axes, rkdt=rkdt, rRe=rRe)
        h = 0.5 * rkdt
        g = (1.0 + h * lam) / (1.0 - h * lam)
        np.testing.assert_allclose(out[1:-1, 1:-1, 1:-1], g * v2[1:-1, 1:-1, 1:-1], rtol=0, atol=1e-12)


@pytest.mark.parametrize("var", [1, 3])
def test_cn_linear_profile_is_steady_on_a_stretched_grid(var):
    nx, ny, nz = 3, 3, 10
    az = G.tanh_axis(nz, 1.0, 2.0)
    axes = _axes(nx, ny, nz, az)
    z = az["xc"] if var == 1 else az["xe"]       # u at centres, w at upper faces
    a, b = 0.4, -1.7
    prof = a + b * z
    TB, VB = dns.flow_bc((True, True, False), {(1, var): (1, prof[0]), (2, var): (1, prof[nz + 1])})
    v2 = np.broadcast_to(prof[:, None, None], (nz + 2, ny + 2, nx + 2)).copy()
    out = _cn(np.zeros_like(v2), v2, var, TB, VB, axes, rkdt=0.05)
    np.testing.assert_allclose(out[1:-1, 1:-1, 1:-1], v2[1:-1, 1:-1, 1:-1], rtol=0, atol=1e-13)


def test_cn_explicit_and_pressure_terms_enter_linearly():
    nx, ny, nz = 4, 5, 6
    axes = _axes(nx, ny, nz, G.tanh_axis(nz, 1.0, 1.0))
    TB, VB = dns.flow_bc((True, True, False), {(f, 1): (2, 0.0) for f in (1, 2)})
    rng = np.random.Generator(np.random.PCG64(2))
    v1, v2, p = (rng.standard_normal((nz + 2, ny + 2, nx + 2)) for _ in range(3))
    base = _cn(v1, v2, 1, TB, VB, axes, f2=0.0)
    a = _cn(v1, np.zeros_like(v2), 1, TB, VB, axes, f1=0.0, f2=0.6, f3=2.0)
    np.testing.assert_allclose(_cn(v1, v2, 1, TB, VB, axes, f2=0.6, f3=2.0), base + a, rtol=0, atol=1e-13)
    q = _cn(np.zeros_like(v1), np.zeros_like(v2), 1, TB, VB, axes, f1=0.0, p=p)
    np.testing.assert_allclose(_cn(v1, v2, 1, TB, VB, axes, p=p), base + q, rtol=0, atol=1e-13)


def _literal_cn(v1, v2, f1, f2, f3, rkdt, var, r, axes, TB, VB, rRe, rPr, rSc):
    """dns/NSsolver.F90:3136-3436 for one column at a time, Fortran indices."""
    ax, ay, az = axes
    nz, ny, nx = (s - 2 for s in v2.shape)
    sz, ez = 2, nz + 1
    rdze = lambda k: 1.0 / az["dxe"][k - 1]
    rdzc = lambda k: 1.0 / az["dxc"][k - 1]
    R = lambda k: r[k - 1]
    out = v2.copy()
    for j in range(2, ny + 2):
        for i in range(2, nx + 2):
            V2 = lambda k: v2[k - 1, j - 1, i - 1]
            n = nz + 2
            sub, diag, sup, rhs = (np.zeros(n + 1) for _ in range(4))     # 1-based
            for k in range(sz, ez + 1):
                if var in (1, 2, 5, 8):
                    raver1 = 0.5 * (R(k + 1) + R(k))
                    raver2 = 0.5 * (R(k) + R(k - 1))
                    if var in (1, 2):
                        tmp1, tmp2 = rRe * raver1 / R(k), rRe * raver2 / R(k)
                    elif var == 5:
                        tmp1, tmp2 = rRe * rPr * raver1 / R(k), rRe * rPr * raver2 / R(k)
                    else:
                        tmp1 = tmp2 = rRe * rSc
                    sub[k] = -0.5 * rkdt * tmp1 * rdze(k - 1) * rdzc(k)
                    sup[k] = -0.5 * rkdt * tmp2 * rdze(k) * rdzc(k)
                    diag[k] = 1.0 + 0.5 * rkdt * (tmp1 * rdze(k - 1) + tmp2 * rdze(k)) * rdzc(k)
                    rhs[k] = f1 * V2(k) + f2 * v1[k - 1, j - 1, i - 1] * f3
                    rhs[k] = rhs[k] + rkdt * 0.5 * (tmp2 * (V2(k + 1) - V2(k)) * rdze(k)
                                                    - tmp1 * (V2(k) - V2(k - 1)) * rdze(k - 1)) * rdzc(k)
                else:
                    tmp1 = rRe * R(k) / (0.5 * (R(k + 1) + R(k)))
                    tmp2 = rRe * R(k + 1) / (0.5 * (R(k + 1) + R(k)))
                    sub[k] = -0.5 * rkdt * tmp1 * rdze(k) * rdzc(k)
                    sup[k] = -0.5 * rkdt * tmp2 * rdze(k) * rdzc(k + 1)
                    diag[k] = 1.0 + 0.5 * rkdt * (tmp1 * rdzc(k) + tmp2 * rdzc(k + 1)) * rdze(k)
                    rhs[k] = f1 * V2(k) + f2 * v1[k - 1, j - 1, i - 1] * f3
                    rhs[k] = rhs[k] + rkdt * 0.5 * (tmp2 * (V2(k + 1) - V2(k)) * rdzc(k + 1)
                                                    - tmp1 * (V2(k) - V2(k - 1)) * rdzc(k)) * rdze(k)
            # Dirichlet walls (the only rows this restatement covers)
            assert TB[2, 0, var - 1] == 1 and TB[2, 1, var - 1] == 1
            sub[sz - 1], diag[sz - 1], sup[sz - 1], rhs[sz - 1] = 0.0, 1.0, 0.0, VB[2, 0, var - 1]
            sub[ez + 1], diag[ez + 1], sup[ez + 1], rhs[ez + 1] = 0.0, 1.0, 0.0, VB[2, 1, var - 1]
            a, b, c, g = sub[1:].copy(), diag[1:].copy(), sup[1:].copy(), rhs[1:].copy()
            for q in range(n - 1):
                a[q + 1] = -a[q + 1] / b[q]
                b[q + 1] = b[q + 1] + a[q + 1] * c[q]
                g[q + 1] = g[q + 1] + a[q + 1] * g[q]
            g[n - 1] = g[n - 1] / b[n - 1]
            for q in range(n - 2, -1, -1):
                g[q] = (g[q] - c[q] * g[q + 1]) / b[q]
            out[:, j - 1, i - 1] = g
    return out


@pytest.mark.parametrize("var", [1, 3, 5, 8])
def test_cn_matches_literal_columns_with_varying_rho(var):
    nx, ny, nz = 3, 2, 7
    axes = _axes(nx, ny, nz, G.tanh_axis(nz, 1.0, 1.8))
    rng = np.random.Generator(np.random.PCG64(7))
    rho = 1.0 + 0.3 * rng.random(nz + 2)
    v1, v2 = (rng.standard_normal((nz + 2, ny + 2, nx + 2)) for _ in range(2))
    TB, VB = dns.flow_bc((True, True, False), {(1, var): (1, 0.2), (2, var): (1, -0.4)})
    args = (v1, v2, 1.0, 0.45, 1.0, 0.03, var, rho, axes, TB, VB, 0.8, 0.6, 1.4)
    got = dns.cn_advance(*args)
    want = _literal_cn(*args)
    # interior columns before the final ghost are the solve; the ghost then refills
    assert np.array_equal(got[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1])


def test_cn_rejects_bcs_the_reference_stops_on():
    axes = _axes(3, 3, 4)
    TB, VB = dns.flow_bc((True, True, False), {(1, 1): (4, 0.0), (2, 1): (1, 0.0)})
    v = np.zeros((6, 5, 5))
    with pytest.raises(ValueError, match="bottom"):
        _cn(v, v, 1, TB, VB, axes)
