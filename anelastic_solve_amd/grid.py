"""Grid metrics for the solver: builders and the reference's grid-file format.

A direction is described as the reference's ``module Grid`` holds it
(``include/modules.F90:17-39``): ghost-inclusive arrays of length ``n+2``
(``xe`` upper cell face, ``xc`` centre, ``dxe`` centre distance, ``dxc``
cell width), with the ghost fix-ups of ``utils/loadGrid.F90:79-90``.
"""
import ctypes as C
import math

import numpy as np

from . import _capi


class Axis:
    """One direction's metric arrays (numpy float64, length ``n+2``)."""

    def __init__(self, xe, xc, dxe, dxc):
        self.xe = np.ascontiguousarray(xe, dtype=np.float64)
        self.xc = np.ascontiguousarray(xc, dtype=np.float64)
        self.dxe = np.ascontiguousarray(dxe, dtype=np.float64)
        self.dxc = np.ascontiguousarray(dxc, dtype=np.float64)
        self.n = self.xe.size - 2

    @property
    def rdxe(self):
        return 1.0 / self.dxe

    @property
    def rdxc(self):
        return 1.0 / self.dxc


def axis_from_faces(faces):
    """Metrics from the ``n+1`` cell faces, with loadGrid's ghost fix-ups."""
    F = np.asarray(faces, dtype=np.float64)
    n = F.size - 1
    if n < 1 or np.any(np.diff(F) <= 0):
        raise ValueError("faces must be strictly increasing")
    dxc = np.empty(n + 2)
    dxc[1:n + 1] = np.diff(F)
    dxc[0], dxc[n + 1] = dxc[1], dxc[n]
    xe = np.empty(n + 2)
    xe[:n + 1] = F
    xe[n + 1] = F[n] + dxc[n + 1]
    xc = np.empty(n + 2)
    xc[1:n + 1] = 0.5 * (F[1:] + F[:-1])
    xc[n + 1] = F[n] + 0.5 * dxc[n + 1]
    dxe = np.empty(n + 2)
    dxe[1:n + 1] = xc[2:] - xc[1:n + 1]
    dxe[0], dxe[n + 1] = dxe[1], dxe[n]
    xc[0] = xc[1] - dxe[0]
    return Axis(xe, xc, dxe, dxc)


def uniform_axis(n, L, x0=0.0):
    """``n`` cells of width ``L/n``."""
    return axis_from_faces(x0 + L * np.arange(n + 1) / n)


def tanh_axis(n, L, beta, x0=0.0):
    """Stretching clustered at the mid-plane: ``z = L (1/2 + atanh((2s-1) tanh b)/(2b))``.

    ``s = k/n``.  The ratio of largest to smallest spacing tends to
    ``cosh(b)^2`` (14.1 for ``b = 2``).  ``beta <= 0`` is uniform.
    """
    if beta <= 0.0:
        return uniform_axis(n, L, x0)
    s = np.arange(n + 1) / n
    t = math.tanh(beta)
    z = 0.5 + np.arctanh(np.clip((2.0 * s - 1.0) * t, -t, t)) / (2.0 * beta)
    z[0], z[-1] = 0.0, 1.0
    return axis_from_faces(x0 + L * z)


def load_grid_file(path):
    """Read one ``x{1,2,3}_grid.in`` file of the reference (``utils/loadGrid.F90:38-70``).

    Format: two header lines, then ``nxp2``, ``L``, ``x0``, the grid type,
    and ``nxp2`` rows ``idum xe xc dxe dxc rdxe rdxc`` (Fortran ``D``
    exponents allowed).  The end-cell fix-ups of ``:79-90`` are applied.
    Returns ``(Axis, L, x0)``.
    """
    with open(path) as fh:
        lines = [l for l in fh.read().splitlines()]
    def num(s):
        return float(s.strip().replace("D", "E").replace("d", "e"))
    nxp2 = int(lines[2].split()[0])
    L = num(lines[3].split()[0])
    x0 = num(lines[4].split()[0])
    rows = []
    for l in lines[6:6 + nxp2]:
        parts = l.split()
        rows.append([num(p) for p in parts[1:7]])
    a = np.array(rows, dtype=np.float64)
    if a.shape != (nxp2, 6):
        raise ValueError(f"{path}: expected {nxp2} rows of 7 columns")
    xe, xc, dxe, dxc = (a[:, q].copy() for q in range(4))
    n = nxp2 - 2
    dxc[n + 1] = dxc[n]
    dxe[n + 1] = dxe[n]
    dxc[0] = dxc[1]
    dxe[0] = dxe[1]
    xe[n + 1] = xe[n] + dxc[n + 1]
    xc[n + 1] = xc[n] + dxe[n]
    xe[0] = xe[1] - dxc[1]
    xc[0] = xc[1] - dxe[0]
    return Axis(xe, xc, dxe, dxc), L, x0


def write_grid_file(path, axis, L, x0, label="X1", gridtype="custom"):
    """Write one direction in the format ``load_grid_file`` reads
    (``utils/loadGrid.F90:44-58``): two comment lines, ``nxp2``, ``L``,
    ``x0``, the grid type, then ``nxp2`` rows ``i xe xc dxe dxc rdxe rdxc``
    with Fortran ``D`` exponents and a trailing spacing-ratio column."""
    def d(v):
        return ("%24.16E" % v).replace("E", "D")
    n2 = axis.n + 2
    out = [f"#{label}-GRID",
           "#   i" + "".join(f"{h:>24}" for h in ("xe(i)", "xc(i)", "dxe(i)", "dxc(i)",
                                                  "rdxe(i)", "rdxc(i)")) + " dxc(i)/dxc(i-1)",
           f"{n2:6d}", d(L), d(x0), f" {gridtype}"]
    for i in range(n2):
        ratio = "n/a" if i == 0 else "%.8f" % (axis.dxc[i] / axis.dxc[i - 1])
        out.append(f"{i + 1:5d}" + "".join(d(v) for v in (axis.xe[i], axis.xc[i], axis.dxe[i],
                                                           axis.dxc[i], axis.rdxe[i],
                                                           axis.rdxc[i])) + f"{ratio:>17}")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def load_grid_dir(directory, rho=None):
    """``Grid`` from ``x1_grid.in``, ``x2_grid.in`` and ``x3_grid.in`` in ``directory``
    (the three files ``utils/loadGrid.F90`` reads)."""
    import os
    axes = [load_grid_file(os.path.join(directory, f"x{d}_grid.in"))[0] for d in (1, 2, 3)]
    return Grid(*axes, rho=rho)


class Grid:
    """Three axes plus the background density ``rho(z)`` (length ``nz+2``, or None = 1)."""

    def __init__(self, ax, ay, az, rho=None):
        self.ax, self.ay, self.az = ax, ay, az
        self.nx, self.ny, self.nz = ax.n, ay.n, az.n
        self.rho = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        if self.rho is not None and self.rho.size != self.nz + 2:
            raise ValueError("rho must have nz+2 values (ghosts included)")

    @property
    def shape(self):
        """Shape of a ghosted field in the reference's memory order: (nz+2, ny+2, nx+2)."""
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    def rho_or_ones(self):
        return np.ones(self.nz + 2) if self.rho is None else self.rho

    def to_c(self):
        """``amg_grid`` view; keep the Grid alive while the struct is in use."""
        def p(a):
            return a.ctypes.data_as(C.POINTER(C.c_double))
        g = _capi.amg_grid()
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.xe, g.xc, g.dxe, g.dxc = p(self.ax.xe), p(self.ax.xc), p(self.ax.dxe), p(self.ax.dxc)
        g.ye, g.yc, g.dye, g.dyc = p(self.ay.xe), p(self.ay.xc), p(self.ay.dxe), p(self.ay.dxc)
        g.ze, g.zc, g.dze, g.dzc = p(self.az.xe), p(self.az.xc), p(self.az.dxe), p(self.az.dxc)
        g.rho = p(self.rho) if self.rho is not None else C.POINTER(C.c_double)()
        return g


def box_grid(nx, ny, nz, Lx=2 * math.pi, Ly=2 * math.pi, Lz=1.0, beta=0.0, rho=None):
    """Periodic-box style grid: uniform x/y, uniform or tanh-stretched z."""
    return Grid(uniform_axis(nx, Lx), uniform_axis(ny, Ly), tanh_axis(nz, Lz, beta), rho)
