"""Time-stepper boundary handling and the implicit vertical viscous step (oracle).

* ``ghost`` / ``bound_update``: the serial branch of ``dns/BC.F90:1-187``
  and ``:334-584`` -- every ghost made periodic first (x, then y, then z),
  then each face's boundary rule, in the order X1min, X1max, X2min, X2max,
  X3min, X3max.
* ``cn_advance``: ``var_advance_CN_PL`` (``dns/NSsolver.F90:3054-3465``),
  the DNS branch (``Deddy`` is zero, ``:3102``), serial (``sizeX3 == 1``:
  the mgMK ``Thomas`` of ``mgMK/mg_relax.F90:541-595`` on every interior
  column, ``:3431-3436``).

Conventions are those of ``oracle.ops``: a ghosted field is a C-order array
``(nz+2, ny+2, nx+2)`` indexed ``[k, j, i]``; the Fortran index ``f`` of the
serial build (``sx = 2``, ``ex = nx+1``, ``dns/Sturb.F90:99-104``) is the
numpy index ``f-1``.  ``TB``/``VB`` are arrays of shape ``(3, 2, 12)`` indexed
``[dir-1, face-1, variable-1]`` (``include/modules.F90:82-85``).  The serial
build sets ``EU = EV = EW = 0`` (``dns/Sturb.F90:110-112``), so the
staggered extra point of ``bound_update`` never moves; the serial build is
also its own boundary on every face (``BND`` true).
"""
import numpy as np

# ghost's variable classes (dns/BC.F90:66-95): vtype -> (l, rt)
GHOST_CLASS = {
    "u": (1, 0), "uc": (1, 0), "v": (2, 0), "vc": (2, 0), "w": (3, 0), "wc": (3, 0),
    "p": (4, 0), "r": (5, 1), "rho": (5, 1), "rf": (5, 1), "cfluc": (6, 0), "fluc": (6, 0),
    "rp": (6, 1), "psource": (7, 0), "utemp": (7, 0), "vtemp": (7, 0), "wtemp": (7, 0),
    "rtemp": (7, 0), "s1temp": (7, 0), "scal1": (8, 1), "scal1temp": (8, 0),
}
# bound_update's metric choice (dns/BC.F90:384-447): vtype -> (dxt, dyt, dzt, of1, of2)
_SCALAR = ("dxe", "dxe", "dxe", -1, 0)
BOUND_METRIC = {
    "u": ("dxc", "dxe", "dxe", 0, 1), "utemp": ("dxc", "dxe", "dxe", 0, 1),
    "v": ("dxe", "dxc", "dxe", 0, 1), "vtemp": ("dxe", "dxc", "dxe", 0, 1),
    "w": ("dxe", "dxe", "dxc", 0, 1), "wtemp": ("dxe", "dxe", "dxc", 0, 1),
}
for _t in ("c", "p", "psource", "fluc", "rho", "rp", "rf", "uc", "vc", "wc", "rtemp", "cfluc",
           "scal1", "scal1temp", "s1temp"):
    BOUND_METRIC[_t] = _SCALAR

# var_advance_CN_PL's variable number -> the vtype of its final ghost (:3438-3452)
CN_GHOST = {1: "u", 2: "v", 3: "w", 5: "rho", 8: "scal1"}


def _plane(var, d, idx):
    """The index tuple of plane ``idx`` (numpy index) normal to direction d (1..3)."""
    s = [slice(None)] * 3
    s[3 - d] = idx
    return tuple(s)


def bound_update(var, Btype, face, d, value, vtype, rtype, axes):
    """One face's rule (``dns/BC.F90:334-584``), in place.  ``axes`` holds the
    x, y, z metric dicts (``dxc``, ``dxe`` ghost-inclusive).  A type a
    direction does not know leaves the field as it is (the reference prints
    an error and goes on); so does a vtype without metrics (``'r'``)."""
    if vtype not in BOUND_METRIC:
        return var
    mx, my, mz, of1, of2 = BOUND_METRIC[vtype]
    dt = axes[d - 1][(mx, my, mz)[d - 1]]
    n = var.shape[3 - d] - 2
    s, e = 1, n                     # sx, ex as numpy indices
    P = lambda idx: _plane(var, d, idx)
    if Btype == 1:
        if face == 1:
            var[P(s - 1)] = value
        else:
            var[P(e + 1)] = value
    elif Btype == 2:
        if face == 1:
            if rtype == 1:
                var[P(s)] = var[P(s + 1)] - value * dt[s + 1 + of1]
            var[P(s - 1)] = var[P(s)] - value * dt[s + of1]
        else:
            if rtype == 1:
                var[P(e)] = var[P(e - 1)] + value * dt[e - 1 + of2]
            var[P(e + 1)] = var[P(e)] + value * dt[e + of2]
    elif Btype == 3:
        var[P(s - 1)] = var[P(e)]
        var[P(e + 1)] = var[P(s)]
    elif Btype == 4:
        if face == 1:
            var[P(s - 1)] = var[P(s)] - value * dt[s + of1]
        else:
            var[P(e + 1)] = var[P(e)] + value * dt[e + of2]
    elif Btype == 8 and d == 1:
        if face == 1:
            var[P(s - 1)] = 2.0 * var[P(s)] - var[P(s + 1)]
        else:
            var[P(e + 1)] = 2.0 * var[P(e)] - var[P(e - 1)]
    elif Btype == 6 and d == 3:
        if face == 1:
            var[P(s - 1)] = 2.0 * value - 1.0 * var[P(s)]
        else:
            var[P(e + 1)] = 2.0 * value - 1.0 * var[P(e)]
    elif Btype == 7 and d == 3:
        if face == 1:
            var[P(s)] = 0.0
            var[P(s - 1)] = -1.0 * var[P(s + 1)]
        else:
            var[P(e)] = 0.0
            var[P(e + 1)] = -1.0 * var[P(e - 1)]
    return var


def ghost(var, vtype, TB, VB, axes):
    """``ghost(var, vtype, stat)`` (``dns/BC.F90:1-187``), serial, in place."""
    nz, ny, nx = (s - 2 for s in var.shape)
    # every ghost periodic first (:51-61)
    var[:, :, 0] = var[:, :, nx]
    var[:, :, nx + 1] = var[:, :, 1]
    var[:, 0, :] = var[:, ny, :]
    var[:, ny + 1, :] = var[:, 1, :]
    var[0, :, :] = var[nz, :, :]
    var[nz + 1, :, :] = var[1, :, :]
    l, rt = GHOST_CLASS[vtype]
    for d in (1, 2, 3):
        for m in (1, 2):
            # face 1's type decides for both faces (:100-176)
            if TB[d - 1, 0, l - 1] != 3 or m == 1:
                bound_update(var, TB[d - 1, m - 1, l - 1], m, d, VB[d - 1, m - 1, l - 1], vtype, rt, axes)
    return var


def cn_coefficients(var, rkdt, rho, az, rRe, rPr, rSc):
    """The column-independent rows of the CN system for interior k
    (``dns/NSsolver.F90:3136-3216``): ``sub, diag, sup`` of length nz+2 (rows
    0 and nz+1 left zero here) and the viscous factors ``tmp1, tmp2``."""
    nz = az["n"]
    r = np.asarray(rho, dtype=np.float64)
    rdze = 1.0 / az["dxe"]
    rdzc = 1.0 / az["dxc"]
    sub = np.zeros(nz + 2)
    diag = np.zeros(nz + 2)
    sup = np.zeros(nz + 2)
    t1 = np.zeros(nz + 2)
    t2 = np.zeros(nz + 2)
    for k in range(1, nz + 1):
        if var in (1, 2, 5):
            raver1 = 0.5 * (r[k + 1] + r[k])
            raver2 = 0.5 * (r[k] + r[k - 1])
            c = rRe if var != 5 else rRe * rPr
            tmp1 = c * raver1 / r[k]
            tmp2 = c * raver2 / r[k]
        elif var == 8:
            tmp1 = tmp2 = rRe * rSc
        elif var == 3:
            tmp1 = rRe * r[k] / (0.5 * (r[k + 1] + r[k]))
            tmp2 = rRe * r[k + 1] / (0.5 * (r[k + 1] + r[k]))
        else:
            raise ValueError("var_advance_CN_PL: variable %d is not implemented" % var)
        if var == 3:
            sub[k] = -0.5 * rkdt * tmp1 * rdze[k] * rdzc[k]
            sup[k] = -0.5 * rkdt * tmp2 * rdze[k] * rdzc[k + 1]
            diag[k] = 1.0 + 0.5 * rkdt * (tmp1 * rdzc[k] + tmp2 * rdzc[k + 1]) * rdze[k]
        else:
            sub[k] = -0.5 * rkdt * tmp1 * rdze[k - 1] * rdzc[k]
            sup[k] = -0.5 * rkdt * tmp2 * rdze[k] * rdzc[k]
            diag[k] = 1.0 + 0.5 * rkdt * (tmp1 * rdze[k - 1] + tmp2 * rdze[k]) * rdzc[k]
        t1[k], t2[k] = tmp1, tmp2
    return sub, diag, sup, t1, t2


def cn_bc_rows(sub, diag, sup, rhs, var, TB, VB, az):
    """The wall rows of ``var_advance_CN_PL`` (``:3337-3428``), in place on
    ``sub, diag, sup`` (1-D, every column alike) and ``rhs`` (3-D, the
    column's rows).  Raises on a type the reference stops on."""
    nz = az["n"]
    dzc, dze = az["dxc"], az["dxe"]
    s, e = 1, nz

    def row(k, a, b, c, val):
        sub[k], diag[k], sup[k] = a, b, c
        rhs[k] = val

    tb, vb = TB[2, 0, var - 1], VB[2, 0, var - 1]
    if tb == 1:
        row(s - 1, 0.0, 1.0, 0.0, vb)
    elif tb == 2:
        if var == 3:
            row(s - 1, 0.0, -1.0, 1.0, dzc[s - 1] * vb)
        elif var in (5, 8):
            row(s - 1, 0.0, -1.0, 1.0, dze[s - 1] * vb)
            row(s, 0.0, -1.0, 1.0, dze[s] * vb)
        else:
            row(s - 1, 0.0, -1.0, 1.0, dze[s - 1] * vb)
    elif tb == 6:
        row(s - 1, 0.0, 1.0, 1.0, 2.0 * vb)
    else:
        raise ValueError("Incorrect BC for bottom for variable %d in var_advance_CN" % var)
    tb, vb = TB[2, 1, var - 1], VB[2, 1, var - 1]
    if tb == 1:
        row(e + 1, 0.0, 1.0, 0.0, vb)
    elif tb == 2:
        if var == 3:
            row(e + 1, -1.0, 1.0, 0.0, dzc[e + 1] * vb)
        elif var in (5, 8):
            row(e + 1, -1.0, 1.0, 0.0, dze[e + 1] * vb)
            row(e, -1.0, 1.0, 0.0, dze[e] * vb)
        else:
            row(e + 1, -1.0, 1.0, 0.0, dze[e + 1] * vb)
    elif tb == 6:
        row(e + 1, 1.0, 1.0, 0.0, 2.0 * vb)
    elif tb == 7:
        row(e + 1, 0.0, 1.0, 0.0, 0.0)
        row(e, 0.0, 1.0, 0.0, 0.0)
    else:
        raise ValueError("Incorrect BC for top for variable %d in var_advance_CN" % var)


def thomas(a, b, c, G):
    """``Thomas(a,b,c,G,n)`` (``mgMK/mg_relax.F90:541-595``): a sub-, b main-,
    c super-diagonal; G (first axis = the system) is overwritten by the
    solution.  ``a`` and ``b`` are 1-D here (the same for every column) and
    are modified as in the reference."""
    n = G.shape[0]
    for j in range(n - 1):
        a[j + 1] = -a[j + 1] / b[j]
        b[j + 1] = b[j + 1] + a[j + 1] * c[j]
        G[j + 1] = G[j + 1] + a[j + 1] * G[j]
    G[n - 1] = G[n - 1] / b[n - 1]
    for j in range(n - 2, -1, -1):
        G[j] = (G[j] - c[j] * G[j + 1]) / b[j]
    return G


def cn_advance(v1, v2, f1, f2, f3, rkdt, var, rho, axes, TB, VB, rRe, rPr, rSc, p=None):
    """``var_advance_CN_PL(veltemp1, veltemp2, f1, f2, f3, rkdt, var)``
    (``dns/NSsolver.F90:3054-3465``): returns the new ``veltemp2``.

    ``p`` given means ``pcorrect`` (the predicted-pressure term, ``:3253-3258``,
    ``:3296-3297``).  Steps: right-hand side on interior cells
    (``:3232-3327``), ``ghost(..,'cfluc')`` of the four arrays (``:3331-3335``),
    the wall rows (``:3337-3428``), ``Thomas`` on every interior column
    (``:3430-3436``), ``ghost`` of the result with the variable's type
    (``:3438-3452``).  Only interior columns are solved, so the right-hand
    side is formed on them only (the reference also forms it on ghost
    columns, whose values the final ghost overwrites)."""
    ax, ay, az = axes
    nz, ny, nx = (s - 2 for s in v2.shape)
    r = np.asarray(rho, dtype=np.float64)
    sub1, diag1, sup1, t1, t2 = cn_coefficients(var, rkdt, r, az, rRe, rPr, rSc)
    rdze = 1.0 / az["dxe"]
    rdzc = 1.0 / az["dxc"]
    rhs = np.zeros_like(v2)
    I = (slice(1, nz + 1), slice(1, ny + 1), slice(1, nx + 1))
    for k in range(1, nz + 1):
        vc = v2[k, 1:-1, 1:-1]
        vm = v2[k - 1, 1:-1, 1:-1]
        vp = v2[k + 1, 1:-1, 1:-1]
        h = f1 * vc + f2 * v1[k, 1:-1, 1:-1] * f3
        if p is not None:
            if var == 1:
                h = h - rkdt * (p[k, 1:-1, 2:] - p[k, 1:-1, 1:-1]) * (1.0 / ax["dxe"][1:nx + 1])[None, :]
            elif var == 2:
                h = h - rkdt * (p[k, 2:, 1:-1] - p[k, 1:-1, 1:-1]) * (1.0 / ay["dxe"][1:ny + 1])[:, None]
            elif var == 3:
                h = h - rkdt * (p[k + 1, 1:-1, 1:-1] - p[k, 1:-1, 1:-1]) * rdze[k]
        if var == 3:
            h = h + rkdt * 0.5 * (t2[k] * (vp - vc) * rdzc[k + 1] - t1[k] * (vc - vm) * rdzc[k]) * rdze[k]
        else:
            h = h + rkdt * 0.5 * (t2[k] * (vp - vc) * rdze[k] - t1[k] * (vc - vm) * rdze[k - 1]) * rdzc[k]
        rhs[k, 1:-1, 1:-1] = h
    # ghost(sub/diag/sup/rhs, 'cfluc'): every column holds the same rows
    full = lambda a: np.broadcast_to(a[:, None, None], v2.shape).copy()
    sub, diag, sup = full(sub1), full(diag1), full(sup1)
    for a in (sub, diag, sup, rhs):
        ghost(a, "cfluc", TB, VB, axes)
    sub1, diag1, sup1 = sub[:, 1, 1].copy(), diag[:, 1, 1].copy(), sup[:, 1, 1].copy()
    cn_bc_rows(sub1, diag1, sup1, rhs, var, TB, VB, az)
    G = rhs[:, 1:-1, 1:-1].copy()
    thomas(sub1, diag1, sup1, G)
    out = v2.copy()
    out[:, 1:-1, 1:-1] = G
    ghost(out, CN_GHOST[var], TB, VB, axes)
    return out


def flow_bc(per=(True, True, False), z=None):
    """``TB``/``VB`` tables as ``bc()`` fills them (``dns/BC.F90:691-1432``):
    ``per[d]`` makes direction d periodic for every variable (type 3, value
    -1, ``:761-763``); ``z`` is ``{(face, variable): (type, value)}`` for the
    wall directions; the pressure-source class 7 is Dirichlet 0 on walls
    (``:725-726``)."""
    TB = np.zeros((3, 2, 12), dtype=np.int32)
    VB = np.zeros((3, 2, 12))
    TB[:, :, 6] = 1
    VB[:, :, 6] = 0.0
    for d in range(3):
        if per[d]:
            TB[d, :, :] = 3
            VB[d, :, :] = -1.0
    for (face, v), (t, val) in (z or {}).items():
        TB[2, face - 1, v - 1] = t
        VB[2, face - 1, v - 1] = val
    return TB, VB
