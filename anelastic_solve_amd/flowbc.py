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

# type names as bc()'s select-case lists spell them (case-sensitive)
_NAMES = {1: ("Dirichlet", "dirichlet", "dir"), 2: ("Neumann", "neumann", "neu"),
          8: ("Etpl8", "etpl8"), 6: ("Wall6", "wall6"), 7: ("Wall7", "wall7"),
          4: ("Calc", "calc", "c")}

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


# bc() reads, for each face block and variable key, a type (names allowed on
# that face) and a value (dns/BC.F90:759-1393).  The table keeps the reader's
# own value targets: on X1max, U1VALUE's key reads U2VALUE into VB(1,1,2) and
# U2VALUE's key reads U1VALUE into VB(1,2,2) (:843-866), and the X1max U3
# extrapolation type is spelled 'Etpl8' or 'expl8' (:870).
_KEYS = (("U1", U), ("U2", V), ("U3", W), ("P", P), ("RHO", RHO), ("FLUC", FLUC), ("SCAL1", SCAL1))


def _allowed(d, f, v):
    t = [1, 2]
    if d == 1 and f == 2 and v in (U, V, W, RHO):
        t.append(8)
    if d == 3 and v in (U, V, W, RHO, SCAL1):
        t += [6, 7]
    if d == 3 and v in (RHO, SCAL1):
        t.append(4)
    return t


def _value_target(d, f, key):
    if d == 1 and f == 2 and key == "U1":
        return "U2VALUE", (1, 1, V)
    if d == 1 and f == 2 and key == "U2":
        return "U1VALUE", (1, 2, V)
    return key + "VALUE", None


def _read_ini(path):
    """The reference's INI syntax (utils/readINI.F90): [block], key=value,
    comments after #, ; or !, case-insensitive names."""
    ini, block = {}, ""
    with open(path) as fh:
        for line in fh:
            for c in "#;!":
                line = line.split(c, 1)[0]
            line = line.strip()
            if not line:
                continue
            if line.startswith("["):
                block = line[1:].split("]", 1)[0].strip().lower()
                continue
            if "=" in line:
                k, v = line.split("=", 1)
                ini.setdefault(block, {})[k.strip().lower()] = v.strip()
    return ini


def _real(s):
    return float(s.lower().replace("d", "e"))


def read_bc(infile):
    """``bc(infile, stat)`` (``dns/BC.F90:691-1432``): the ``[BCs]`` block's
    ``X1``/``X2``/``X3`` (``periodic`` or ``read``) and, for read directions,
    the ``X<d>min``/``X<d>max`` blocks' ``<VAR>TYPE``/``<VAR>VALUE`` keys.
    The pressure-source class is Dirichlet 0 (``:725-726``).  A missing key
    or a type the face does not allow raises ``ValueError`` (the reference
    stops)."""
    ini = _read_ini(infile)
    b = FlowBC()
    b.TB[:, :, SRC - 1] = DIRICHLET
    b.VB[:, :, SRC - 1] = 0.0
    blk = ini.get("bcs", {})
    for d in (1, 2, 3):
        mode = blk.get("x%d" % d, "ERROR")
        if mode in ("Periodic", "periodic", "per"):
            b.TB[d - 1, :, :] = PERIODIC
            b.VB[d - 1, :, :] = -1.0
            continue
        if mode not in ("Read", "read"):
            raise ValueError("NO VALID BC TYPE FOR X%d: %s" % (d, mode))
        for f, side in ((1, "min"), (2, "max")):
            sec = ini.get("x%d%s" % (d, side), {})
            for key, v in _KEYS:
                name = sec.get(key.lower() + "type", "ERROR")
                t = None
                for cand in _allowed(d, f, v):
                    names = _NAMES[cand]
                    if cand == 8 and key == "U3":
                        names = ("Etpl8", "expl8")
                    if name in names:
                        t = cand
                if t is None:
                    raise ValueError("NO VALID BC'S FOUND FOR %s AT X%d%s" % (key, d, side))
                b.TB[d - 1, f - 1, v - 1] = t
                vkey, target = _value_target(d, f, key)
                if vkey.lower() not in sec:
                    raise ValueError("missing %s in [X%d%s]" % (vkey, d, side))
                dd, ff, vv = target if target else (d, f, v)
                b.VB[dd - 1, ff - 1, vv - 1] = _real(sec[vkey.lower()])
    return b
