"""Boundary conditions of the flow fields: the reference's ``TB``/``VB`` tables.

``TB(1:3,1:2,1:12)`` (type) and ``VB(1:3,1:2,1:12)`` (value) per direction,
face and variable class (module ``boundC``, ``include/modules.F90:82-85``),
as ``bc()`` fills them (``dns/BC.F90:691-1432``).  Variable classes: 1 u,
2 v, 3 w, 4 p, 5 rho, 6 fluctuations, 7 pressure source / tendencies,
8 scal1.  Types: 1 Dirichlet, 2 Neumann, 3 periodic, 4 Neumann on the ghost
only, 6 wall (ghost/first-value mean = value, z), 7 zero wall (z),
8 linear extrapolation (x).

``FlowBC.to_c()`` gives the ``amg_flowbc`` struct of ``include/amg.h``
(Fortran order: index ``(d-1) + 3*(f-1) + 6*(v-1)``).
"""
import numpy as np

from . import _capi

U, V, W, P, RHO, FLUC, SRC, SCAL1 = 1, 2, 3, 4, 5, 6, 7, 8
DIRICHLET, NEUMANN, PERIODIC, GHOST_NEUMANN, WALL6, WALL7, EXTRAP8 = 1, 2, 3, 4, 6, 7, 8


class FlowBC:
    """``TB``/``VB`` as numpy arrays of shape ``(3, 2, 12)`` indexed
    ``[direction-1, face-1, class-1]``."""

    def __init__(self, TB=None, VB=None):
        self.TB = np.zeros((3, 2, 12), dtype=np.int32) if TB is None else np.asarray(TB, np.int32)
        self.VB = np.zeros((3, 2, 12)) if VB is None else np.asarray(VB, np.float64)
        if self.TB.shape != (3, 2, 12) or self.VB.shape != (3, 2, 12):
            raise ValueError("TB and VB are (3, 2, 12) arrays")

    @classmethod
    def make(cls, periodic=(True, True, False), walls=None):
        """As ``bc()`` builds them: the pressure-source class 7 is Dirichlet 0
        (``dns/BC.F90:725-726``); a periodic direction is type 3, value -1 for
        every class (``:761-763``); ``walls`` maps ``(direction, face, class)``
        (1-based) to ``(type, value)``."""
        b = cls()
        b.TB[:, :, SRC - 1] = DIRICHLET
        b.VB[:, :, SRC - 1] = 0.0
        for d in range(3):
            if periodic[d]:
                b.TB[d, :, :] = PERIODIC
                b.VB[d, :, :] = -1.0
        for (d, f, v), (t, val) in (walls or {}).items():
            b.TB[d - 1, f - 1, v - 1] = t
            b.VB[d - 1, f - 1, v - 1] = val
        return b

    def to_c(self):
        s = _capi.amg_flowbc()
        s.tb[:] = [int(x) for x in self.TB.flatten(order="F")]
        s.vb[:] = [float(x) for x in self.VB.flatten(order="F")]
        return s
