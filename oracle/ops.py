"""Operator, projection step and boundary handling (oracle).

Arrays are numpy ``float64`` in the reference's Fortran memory order:
a ghosted field ``phi(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1)`` is a C-order array
of shape ``(nz+2, ny+2, nx+2)`` indexed ``[k, j, i]`` (x fastest), and an
interior field is shape ``(nz, ny, nx)``.
"""
import numpy as np

from .grid import PERIODIC, NEUMANN, DIRICHLET


def _shift(u, axis, step, per):
    """u at index (i+step) along ``axis`` with periodic wrap, or zero-padded."""
    if per:
        return np.roll(u, -step, axis=axis)
    out = np.zeros_like(u)
    sl_dst = [slice(None)] * 3
    sl_src = [slice(None)] * 3
    if step == 1:
        sl_dst[axis] = slice(0, -1)
        sl_src[axis] = slice(1, None)
    else:
        sl_dst[axis] = slice(1, None)
        sl_src[axis] = slice(0, -1)
    out[tuple(sl_dst)] = u[tuple(sl_src)]
    return out


def apply_A(u, lev):
    """``A u`` on interior cells from the 1-D coefficient tables.

    Equivalent to the stencil sum of ``mgMK/mg_restr.F90:130-138`` with the
    pcf metrics; wall couplings are zero in the tables, periodic couplings wrap.
    """
    wxm, wxp, dx = lev["cx"][:3]
    wym, wyp, dy = lev["cy"][:3]
    wzm, wzp, dz = lev["cz"][:3]
    px, py, pz = lev["per"]
    diag = dz[:, None, None] + dy[None, :, None] + dx[None, None, :]
    out = diag * u
    out += wxm[None, None, :] * _shift(u, 2, -1, px) + wxp[None, None, :] * _shift(u, 2, 1, px)
    out += wym[None, :, None] * _shift(u, 1, -1, py) + wyp[None, :, None] * _shift(u, 1, 1, py)
    out += wzm[:, None, None] * _shift(u, 0, -1, pz) + wzp[:, None, None] * _shift(u, 0, 1, pz)
    return out


def diag_of(lev):
    return (lev["cz"][2][:, None, None] + lev["cy"][2][None, :, None]
            + lev["cx"][2][None, None, :])


def fill_ghosts(phi, tbc, vbc, ax, ay, az):
    """Ghost layers of a ghosted field, as the serial path of ``mgMK/mg_bdry.F90``.

    First every face is wrapped periodically over full planes (``:71-83``);
    then the non-periodic faces are overwritten in the reference's order
    X2max, X3min, X2min, X3max, X1min, X1max (``:103-190``): Neumann ghost =
    interior +/- ``vbc * dxc(edge cell)``, Dirichlet ghost = ``vbc``.
    Works in place and returns ``phi``.
    """
    nz, ny, nx = (s - 2 for s in phi.shape)
    phi[:, :, 0] = phi[:, :, nx]
    phi[:, :, nx + 1] = phi[:, :, 1]
    phi[:, 0, :] = phi[:, ny, :]
    phi[:, ny + 1, :] = phi[:, 1, :]
    phi[0, :, :] = phi[nz, :, :]
    phi[nz + 1, :, :] = phi[1, :, :]
    dxc, dyc, dzc = ax["dxc"], ay["dxc"], az["dxc"]
    # face ids 1..6 of the reference: 1 X2max, 2 X3min, 3 X2min, 4 X3max, 5 X1min, 6 X1max
    for f in (1, 2, 3, 4, 5, 6):
        t, v = tbc[f - 1], vbc[f - 1]
        if t == PERIODIC:
            continue
        if f == 1:
            phi[:, ny + 1, :] = phi[:, ny, :] + v * dyc[ny] if t == NEUMANN else v
        elif f == 2:
            phi[0, :, :] = phi[1, :, :] - v * dzc[1] if t == NEUMANN else v
        elif f == 3:
            phi[:, 0, :] = phi[:, 1, :] - v * dyc[1] if t == NEUMANN else v
        elif f == 4:
            phi[nz + 1, :, :] = phi[nz, :, :] + v * dzc[nz] if t == NEUMANN else v
        elif f == 5:
            phi[:, :, 0] = phi[:, :, 1] - v * dxc[1] if t == NEUMANN else v
        elif f == 6:
            phi[:, :, nx + 1] = phi[:, :, nx] + v * dxc[nx] if t == NEUMANN else v
    return phi


def apply_A_ghosted(phi, ax, ay, az, rho):
    """Reference-form stencil on a ghost-filled field.

    The seven-coefficient sum of ``mgMK/mg_restr.F90:130-138`` with the
    stretched pcf metrics (``mgMK/mg_coeff.F90:78-99``) and the density
    factors of ``mgMK/mg_pfpde.F90:49-57``; every neighbour (ghosts
    included) enters.  Used to pin the table form ``apply_A``.
    """
    nz, ny, nx = (s - 2 for s in phi.shape)
    c = phi[1:-1, 1:-1, 1:-1]
    dxc, dxe = ax["dxc"], ax["dxe"]
    dyc, dye = ay["dxc"], ay["dxe"]
    dzc, dze = az["dxc"], az["dxe"]
    p1 = (1.0 / (dxc[1:nx + 1] * dxe[0:nx]))[None, None, :]
    p2 = (1.0 / (dxc[1:nx + 1] * dxe[1:nx + 1]))[None, None, :]
    p3 = (1.0 / (dyc[1:ny + 1] * dye[0:ny]))[None, :, None]
    p4 = (1.0 / (dyc[1:ny + 1] * dye[1:ny + 1]))[None, :, None]
    p5 = (1.0 / (dzc[1:nz + 1] * dze[0:nz]))
    p6 = (1.0 / (dzc[1:nz + 1] * dze[1:nz + 1]))
    r = np.asarray(rho)
    rk, rm, rp = r[1:nz + 1], r[0:nz], r[2:nz + 2]
    c5 = (p5 * 0.5 * (rk + rm) / rm)[:, None, None]
    c6 = (p6 * 0.5 * (rk + rp) / rp)[:, None, None]
    c7z = (-(p5 * 0.5 * (rk + rm) / rk) - (p6 * 0.5 * (rk + rp) / rk))[:, None, None]
    out = (p1 * phi[1:-1, 1:-1, 0:-2] + p2 * phi[1:-1, 1:-1, 2:]
           + p3 * phi[1:-1, 0:-2, 1:-1] + p4 * phi[1:-1, 2:, 1:-1]
           + c5 * phi[0:-2, 1:-1, 1:-1] + c6 * phi[2:, 1:-1, 1:-1]
           + (c7z - p1 - p2 - p3 - p4) * c)
    return out


def fold_bc_rhs(b, lev, tbc, vbc):
    """Move the fine-level boundary values into the right-hand side.

    With ghost = interior -/+ ``g*dxc`` (Neumann) or ghost = ``v``
    (Dirichlet) (``mgMK/mg_bdry.F90:103-190``, ``phibc`` = ``vbc`` on the
    finest level only, ``mgMK/mg_init.F90:307-315``), the ghost's constant
    part leaves ``A``.  Returns a new interior array.
    """
    out = np.array(b, dtype=np.float64, copy=True)
    ax, ay, az = lev["ax"], lev["ay"], lev["az"]
    nx, ny, nz = lev["n"]
    amx, apx = lev["cx"][3], lev["cx"][4]
    amy, apy = lev["cy"][3], lev["cy"][4]
    amz, apz = lev["cz"][3], lev["cz"][4]
    t = tbc
    v = vbc
    # X1min (face 5), X1max (6)
    if t[4] == NEUMANN:
        out[:, :, 0] += amx[0] * v[4] * ax["dxc"][1]
    elif t[4] == DIRICHLET:
        out[:, :, 0] -= amx[0] * v[4]
    if t[5] == NEUMANN:
        out[:, :, nx - 1] -= apx[nx - 1] * v[5] * ax["dxc"][nx]
    elif t[5] == DIRICHLET:
        out[:, :, nx - 1] -= apx[nx - 1] * v[5]
    # X2min (3), X2max (1)
    if t[2] == NEUMANN:
        out[:, 0, :] += amy[0] * v[2] * ay["dxc"][1]
    elif t[2] == DIRICHLET:
        out[:, 0, :] -= amy[0] * v[2]
    if t[0] == NEUMANN:
        out[:, ny - 1, :] -= apy[ny - 1] * v[0] * ay["dxc"][ny]
    elif t[0] == DIRICHLET:
        out[:, ny - 1, :] -= apy[ny - 1] * v[0]
    # X3min (2), X3max (4)
    if t[1] == NEUMANN:
        out[0, :, :] += amz[0] * v[1] * az["dxc"][1]
    elif t[1] == DIRICHLET:
        out[0, :, :] -= amz[0] * v[1]
    if t[3] == NEUMANN:
        out[nz - 1, :, :] -= apz[nz - 1] * v[3] * az["dxc"][nz]
    elif t[3] == DIRICHLET:
        out[nz - 1, :, :] -= apz[nz - 1] * v[3]
    return out


def volumes(lev):
    nx, ny, nz = lev["n"]
    return (lev["az"]["dxc"][1:nz + 1][:, None, None]
            * lev["ay"]["dxc"][1:ny + 1][None, :, None]
            * lev["ax"]["dxc"][1:nx + 1][None, None, :])


def compat_project(b, lev):
    """Remove the part of ``b`` outside the range of a singular ``A``.

    ``A``'s left null vector is the cell volume (its rows are flux
    differences times ``1/dxc``), so ``b - (sum V b / sum V)`` is compatible.
    """
    V = volumes(lev)
    return b - (V * b).sum() / V.sum()


def null_project(p, lev):
    """Remove the null-space component ``c*rho(z)`` of ``p``.

    ``A (rho) = 0`` because the fluxes act on ``p/rho``; ``c`` makes the
    volume integral of ``p`` zero (the volume-weighted mean for rho = 1).
    """
    V = volumes(lev)
    nz = lev["n"][2]
    r = lev["rho"][1:nz + 1][:, None, None]
    c = (V * p).sum() / (V * r * np.ones_like(p)).sum()
    return p - c * r


def psource(U, V, W, rho, ax, ay, az, dt):
    """Pressure source ``dns/NSsolver.F90:2204-2219`` (``Psource``).

    ``SRC = [rho_k (U_i - U_{i-1})/dxc_i + rho_k (V_j - V_{j-1})/dyc_j
    + ((rho_{k+1}+rho_k)/2 W_k - (rho_k+rho_{k-1})/2 W_{k-1})/dzc_k] / dt``
    on interior cells; the ghost layer is returned zero.  ``rho`` is 1-D
    (length ``nz+2``); velocities are ghosted arrays.
    """
    nz, ny, nx = (s - 2 for s in U.shape)
    r = np.asarray(rho)
    rk = r[1:nz + 1][:, None, None]
    rp = r[2:nz + 2][:, None, None]
    rm = r[0:nz][:, None, None]
    rdxc = (1.0 / ax["dxc"][1:nx + 1])[None, None, :]
    rdyc = (1.0 / ay["dxc"][1:ny + 1])[None, :, None]
    rdzc = (1.0 / az["dxc"][1:nz + 1])[:, None, None]
    src = np.zeros_like(U)
    src[1:-1, 1:-1, 1:-1] = (
        (rk * U[1:-1, 1:-1, 1:-1] - rk * U[1:-1, 1:-1, 0:-2]) * rdxc
        + (rk * V[1:-1, 1:-1, 1:-1] - rk * V[1:-1, 0:-2, 1:-1]) * rdyc
        + (0.5 * (rp + rk) * W[1:-1, 1:-1, 1:-1] - 0.5 * (rk + rm) * W[0:-2, 1:-1, 1:-1]) * rdzc
    ) * (1.0 / dt)
    return src


def pgrad(p, U, V, W, rho, ax, ay, az, dt, walls=(False, False, False)):
    """Velocity correction ``dns/NSsolver.F90:2057-2082`` (``Pgrad``, EU=EV=EW=0).

    ``U_i -= (p_{i+1}/rho_k - p_i/rho_k) rdxe_i dt`` (same for V), and
    ``W_k -= (p_{k+1}/rho_{k+1} - p_k/rho_k) rdze_k dt`` on every interior
    index; reads the p ghost at ``ex+1``.  ``walls[d]`` marks a direction
    whose upper face is a pressure-Neumann wall: that face's normal velocity
    keeps its value, as the velocity boundary condition that ``ghost()``
    re-imposes after each component (``dns/NSsolver.F90:2064,2073,2082``,
    ``dns/BC.F90:394-447``).  Returns new arrays.
    """
    nz, ny, nx = (s - 2 for s in p.shape)
    r = np.asarray(rho)
    rk = r[1:nz + 1][:, None, None]
    rp = r[2:nz + 2][:, None, None]
    rdxe = (1.0 / ax["dxe"][1:nx + 1])[None, None, :]
    rdye = (1.0 / ay["dxe"][1:ny + 1])[None, :, None]
    rdze = (1.0 / az["dxe"][1:nz + 1])[:, None, None]
    U0, V0, W0 = U, V, W
    U, V, W = U.copy(), V.copy(), W.copy()
    pc = p[1:-1, 1:-1, 1:-1]
    U[1:-1, 1:-1, 1:-1] -= (p[1:-1, 1:-1, 2:] / rk - pc / rk) * rdxe * dt
    V[1:-1, 1:-1, 1:-1] -= (p[1:-1, 2:, 1:-1] / rk - pc / rk) * rdye * dt
    W[1:-1, 1:-1, 1:-1] -= (p[2:, 1:-1, 1:-1] / rp - pc / rk) * rdze * dt
    if walls[0]:
        U[1:-1, 1:-1, nx] = U0[1:-1, 1:-1, nx]
    if walls[1]:
        V[1:-1, ny, 1:-1] = V0[1:-1, ny, 1:-1]
    if walls[2]:
        W[nz, 1:-1, 1:-1] = W0[nz, 1:-1, 1:-1]
    return U, V, W


def wrap_periodic(F, per):
    """Periodic ghost wrap of a ghosted field along the periodic axes (x, y, z)."""
    nz, ny, nx = (s - 2 for s in F.shape)
    if per[0]:
        F[:, :, 0] = F[:, :, nx]
        F[:, :, nx + 1] = F[:, :, 1]
    if per[1]:
        F[:, 0, :] = F[:, ny, :]
        F[:, ny + 1, :] = F[:, 1, :]
    if per[2]:
        F[0, :, :] = F[nz, :, :]
        F[nz + 1, :, :] = F[1, :, :]
    return F


def divergence(U, V, W, ax, ay, az):
    """Divergence diagnostics of ``dns/Routines.F90:123-144``.

    Returns ``(div, avg, maxdiv, loc)``: the cell divergence
    ``dU/dxc + dV/dyc + dW/dzc``; the volume-weighted mean of ``|div|``;
    ``max |div|``; and the 1-based Fortran indices ``(i, j, k)`` (interior
    starts at 2) of the first cell in the reference's k, j, i loop order
    where it occurs.  The reference keeps a *signed* running maximum
    (``MaxDiv = rdiv`` after ``dabs(rdiv) > MaxDiv``, ``:131-137``), so once
    that value is negative every later cell replaces it; the true maximum
    of ``|div|`` is reported here instead.
    """
    nz, ny, nx = (s - 2 for s in U.shape)
    rdxc = (1.0 / ax["dxc"][1:nx + 1])[None, None, :]
    rdyc = (1.0 / ay["dxc"][1:ny + 1])[None, :, None]
    rdzc = (1.0 / az["dxc"][1:nz + 1])[:, None, None]
    div = ((U[1:-1, 1:-1, 1:-1] - U[1:-1, 1:-1, 0:-2]) * rdxc
           + (V[1:-1, 1:-1, 1:-1] - V[1:-1, 0:-2, 1:-1]) * rdyc
           + (W[1:-1, 1:-1, 1:-1] - W[0:-2, 1:-1, 1:-1]) * rdzc)
    vol = 1.0 / (rdxc * rdyc * rdzc)
    avg = float((np.abs(div) * vol).sum() / (vol * np.ones_like(div)).sum())
    a = np.abs(div)
    flat = int(np.argmax(a))
    k, j, i = np.unravel_index(flat, div.shape)
    maxdiv = float(a[k, j, i])
    return div, avg, maxdiv, (int(i) + 2, int(j) + 2, int(k) + 2)
