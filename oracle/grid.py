"""Grid metrics, 1-D stencil coefficients and the level hierarchy (oracle).

Conventions follow the reference's ``module Grid`` (``include/modules.F90:17-39``)
with 0-based, ghost-inclusive arrays of length ``n+2``:

* ``xe[i]``  upper face of cell ``i``  (``xe[0]`` is the domain start),
* ``xc[i]``  cell centre, ``(xe[i]+xe[i-1])/2`` for interior cells,
* ``dxc[i]`` cell width ``xe[i]-xe[i-1]``,
* ``dxe[i]`` centre distance ``xc[i+1]-xc[i]``,

and the ghost fix-ups of ``utils/loadGrid.F90:79-90`` (ghost cells have the
width of their interior neighbour, ``dxe[0] = dxe[1]``).

Boundary codes per face are the reference's ``tbc`` values
(``include/mg_modules.F90:14-54``, ``mgMK/mg_read_params.F90:105-214``):
0 periodic, 1 Neumann, 2 Dirichlet.
"""
import numpy as np

PERIODIC, NEUMANN, DIRICHLET = 0, 1, 2


def axis_from_faces(faces):
    """Metric arrays from the ``n+1`` cell faces, with loadGrid's ghost fix-ups.

    ``utils/loadGrid.F90:79-90`` sets ``dxc(1)=dxc(2)``, ``dxe(1)=dxe(2)``,
    ``xe(1)=xe(2)-dxc(2)``, ``xc(1)=xc(2)-dxe(1)`` and the mirror at ``nxp2``.
    """
    F = np.asarray(faces, dtype=np.float64)
    n = F.size - 1
    dxc = np.empty(n + 2)
    dxc[1:n + 1] = F[1:] - F[:-1]
    dxc[0] = dxc[1]
    dxc[n + 1] = dxc[n]
    xe = np.empty(n + 2)
    xe[0:n + 1] = F
    xe[n + 1] = F[n] + dxc[n + 1]
    xc = np.empty(n + 2)
    xc[1:n + 1] = 0.5 * (F[1:] + F[:-1])
    xc[n + 1] = F[n] + 0.5 * dxc[n + 1]
    dxe = np.empty(n + 2)
    dxe[1:n + 1] = xc[2:n + 2] - xc[1:n + 1]
    dxe[0] = dxe[1]
    dxe[n + 1] = dxe[n]
    xc[0] = xc[1] - dxe[0]
    return {"n": n, "xe": xe, "xc": xc, "dxe": dxe, "dxc": dxc}


def uniform_axis(n, L, x0=0.0):
    """Uniform spacing ``L/n`` (the spacing mgdpfpde assumes: ``mgMK/mg_pfpde.F90:35-40``)."""
    return axis_from_faces(x0 + L * np.arange(n + 1) / n)


def tanh_axis(n, L, beta, x0=0.0):
    """Mid-plane clustered stretching ``z = L*(1/2 + atanh((2s-1) tanh b)/(2b))``.

    ``s = k/n`` are uniform face parameters.  Spacing is smallest at the
    mid-plane and largest at the walls; ``dz_max/dz_min -> cosh(b)^2``
    (14.1 at ``b = 2``).  ``beta <= 0`` gives a uniform axis.
    """
    if beta <= 0.0:
        return uniform_axis(n, L, x0)
    s = np.arange(n + 1) / n
    t = np.tanh(beta)
    z = 0.5 + np.arctanh(np.clip((2.0 * s - 1.0) * t, -t, t)) / (2.0 * beta)
    z[0], z[-1] = 0.0, 1.0
    return axis_from_faces(x0 + L * z)


def coarsen_axis(ax, factor=2, wall=False):
    """Coarse axis by 2:1 agglomeration of cells (every other face).

    Replaces ``mgMK/mg_grids.F90:112-198``, whose periodic coarse spacing
    divides a ghost-inclusive length (``:64-66,121``); here the coarse faces
    are exactly the even fine faces, so a periodic uniform axis coarsens to
    exactly ``2 dx``.
    At a wall (``wall``) the coarse ghost centres stay at the finest level's:
    a Dirichlet value sits at that point, and the coarse corrections must
    vanish at the same place.  A ghost half a coarse cell out would move the
    wall at every coarsening, and after several z-only coarsenings the
    coarse-grid correction no longer matches the fine problem (the V-cycle
    diverges).  Neumann walls carry no flux, so the distance is unused there.
    """
    if factor == 1:
        return ax
    n = ax["n"]
    assert n % 2 == 0
    faces = ax["xe"][0:n + 1:2]
    c = axis_from_faces(faces)
    if wall:
        m = c["n"]
        c["xc"][0] = ax["xc"][0]
        c["dxe"][0] = c["xc"][1] - c["xc"][0]
        c["xc"][m + 1] = ax["xc"][n + 1]
        c["dxe"][m] = c["xc"][m + 1] - c["xc"][m]
    return c


def axis_coefs(ax, bc_lo, bc_hi):
    """Off-diagonal and diagonal 1-D coefficients of one horizontal direction.

    ``mgMK/mg_coeff.F90:78-91``: ``p1 = 1/(dxc_i dxe_{i-1})``,
    ``p2 = 1/(dxc_i dxe_i)``; Neumann walls zero the wall coefficient
    (``:119-145``) which also removes it from the diagonal (as the line
    smoother's ``b = -(p1+...+p6)`` does, ``mgMK/mg_relax.F90:484-485``).
    Dirichlet walls keep the diagonal term and move ``coef*value`` to the
    right-hand side (ghost = value, ``mgMK/mg_bdry.F90:109-114``).

    Returns ``(wm, wp, d)`` of length ``n``: ``(A u)_i = wm_i u_{i-1} +
    wp_i u_{i+1} + d_i u_i`` (plus the other directions' terms).
    """
    n = ax["n"]
    dxc, dxe = ax["dxc"], ax["dxe"]
    am = 1.0 / (dxc[1:n + 1] * dxe[0:n])
    ap = 1.0 / (dxc[1:n + 1] * dxe[1:n + 1])
    wm, wp = am.copy(), ap.copy()
    d = -(am + ap)
    if bc_lo != PERIODIC:
        wm[0] = 0.0
        if bc_lo == NEUMANN:
            d[0] += am[0]
    if bc_hi != PERIODIC:
        wp[n - 1] = 0.0
        if bc_hi == NEUMANN:
            d[n - 1] += ap[n - 1]
    return wm, wp, d, am, ap


def z_coefs(ax, rho, bc_lo, bc_hi):
    """Vertical coefficients with the anelastic density weights.

    The operator is ``D(rho G(p/rho))`` of ``Psource``/``Pgrad``
    (``dns/NSsolver.F90:2212-2214`` and ``:2076-2078``): the coefficient of
    ``p_{k+1}`` is ``rdzc_k rdze_k (rho_k+rho_{k+1})/2 / rho_{k+1}`` -- the
    density factors of ``mgMK/mg_pfpde.F90:49-57`` on the stretched metrics
    of ``mgMK/mg_coeff.F90:94-99``.  ``rho`` has ghosts (length ``nz+2``).
    Neumann walls carry zero flux (both the off-diagonal and its diagonal
    share vanish); Dirichlet keeps the diagonal share.
    """
    n = ax["n"]
    dzc, dze = ax["dxc"], ax["dxe"]
    r = np.asarray(rho, dtype=np.float64)
    rk, rm, rp = r[1:n + 1], r[0:n], r[2:n + 2]
    gm = 1.0 / (dzc[1:n + 1] * dze[0:n]) * 0.5 * (rk + rm)
    gp = 1.0 / (dzc[1:n + 1] * dze[1:n + 1]) * 0.5 * (rk + rp)
    am, ap = gm / rm, gp / rp
    wm, wp = am.copy(), ap.copy()
    d = -(gm / rk + gp / rk)
    if bc_lo != PERIODIC:
        wm[0] = 0.0
        if bc_lo == NEUMANN:
            d[0] += gm[0] / rk[0]
    if bc_hi != PERIODIC:
        wp[n - 1] = 0.0
        if bc_hi == NEUMANN:
            d[n - 1] += gp[n - 1] / rk[n - 1]
    return wm, wp, d, am, ap


def coarsen_rho(rho, az, per, factor=2):
    """Volume-weighted coarse density with ghosts (periodic wrap or copy)."""
    if factor == 1:
        return np.asarray(rho, dtype=np.float64).copy()
    n = az["n"]
    dzc = az["dxc"]
    r = np.asarray(rho, dtype=np.float64)
    a, b = r[1:n + 1:2], r[2:n + 2:2]
    wa, wb = dzc[1:n + 1:2], dzc[2:n + 2:2]
    rc = (wa * a + wb * b) / (wa + wb)
    out = np.empty(rc.size + 2)
    out[1:-1] = rc
    if per:
        out[0], out[-1] = rc[-1], rc[0]
    else:
        out[0], out[-1] = rc[0], rc[-1]
    return out


def auto_factors(nx, ny, nz, min_size=4):
    """Full 2:1 coarsening while every size stays even and the smallest exceeds ``min_size``.

    x must stay even on every level (the chequer layout splits rows by
    column colour).  Powers of two end at ``4^3`` as ``mgMK`` does with
    ``nx1Cp = 4`` (``mgMK/mg_init.F90:86-89``).
    """
    f = []
    while (nx % 4 == 0 and ny % 2 == 0 and nz % 2 == 0
           and min(nx, ny, nz) > min_size):
        f.append((2, 2, 2))
        nx, ny, nz = nx // 2, ny // 2, nz // 2
    return f


def reference_factors(nlev):
    """Per-direction level counts with the reference's rule.

    ``mgMK/mg_init.F90:125-129``: ``nxk(k) = nx1Cp*2**(max(k+nx1Levels-ngrid,1)-1)+1``,
    so a direction with fewer levels coarsens at the fine end and stops at
    the coarse end.  Returns the factor triple of each transition, finest
    first.
    """
    ngrid = max(nlev)
    out = []
    for t in range(ngrid - 1):
        out.append(tuple(2 if t < nlev[d] - 1 else 1 for d in range(3)))
    return out


def hierarchy(ax, ay, az, rho, tbc, factors):
    """Axes, density and 1-D coefficient tables of every level, finest first.

    ``tbc`` is the reference's 6-face map (1-based in the reference, here a
    list indexed 0..5): X2max, X3min, X2min, X3max, X1min, X1max.
    """
    bx = (tbc[4], tbc[5])
    by = (tbc[2], tbc[0])
    bz = (tbc[1], tbc[3])
    levels = []
    cur = (ax, ay, az, np.asarray(rho, dtype=np.float64))
    for li in range(len(factors) + 1):
        a, b, c, r = cur
        cx = axis_coefs(a, *bx)
        cy = axis_coefs(b, *by)
        cz = z_coefs(c, r, *bz)
        levels.append({"ax": a, "ay": b, "az": c, "rho": r,
                       "n": (a["n"], b["n"], c["n"]),
                       "cx": cx, "cy": cy, "cz": cz,
                       "per": (bx[0] == PERIODIC, by[0] == PERIODIC, bz[0] == PERIODIC),
                       "bc": (bx, by, bz)})
        if li < len(factors):
            fx, fy, fz = factors[li]
            cur = (coarsen_axis(a, fx, bx[0] != PERIODIC), coarsen_axis(b, fy, by[0] != PERIODIC),
                   coarsen_axis(c, fz, bz[0] != PERIODIC), coarsen_rho(r, c, bz[0] == PERIODIC, fz))
    return levels


def is_singular(tbc):
    """No Dirichlet face: the operator has a one-dimensional null space."""
    return all(t != DIRICHLET for t in tbc)
