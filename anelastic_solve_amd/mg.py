"""Python mirror of the reference's solver and projection interface.

The reference's DNS/LES drivers see the pressure solver only through four
external subroutines (``mgMK/mg_read_params.F90:1``, ``mgMK/mg_init.F90:1``,
``mgMK/mg_solver.F90:1``, ``mgMK/mg_end.F90:1``) and the projection step
through ``Psource``/``Pgrad`` (``dns/NSsolver.F90:2167``, ``:2023``) and
``divergence`` (``dns/Routines.F90:54``).  This module offers the same
names, argument meaning and error behaviour on top of the C ABI
(``include/amg.h``), with fields in the reference's layout: a ghosted FP64
array of shape ``(nz+2, ny+2, nx+2)`` in C order (= the Fortran array
``phi(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1)``).

Device fields are ``torch`` CUDA tensors (HBM-resident, no copies); host
fields are numpy arrays (copied to the device and back, as a Fortran caller
would hand them over).
"""
import ctypes as C
import sys

import numpy as np

from . import _capi
from ._capi import AMGError, check


def default_config(**kw):
    """``amg_config`` with the library defaults, then keyword overrides.

    Keywords: smoother ('RBPGS'|'LJ'), maxcy, tol, stop_rule ('residual'|
    'relmax'), pre, post (int or coarsest-first list), tbc (6 ints, face
    order X2max, X3min, X2min, X3max, X1min, X1max), vbc (6 floats),
    precision ('fp64'|'mixed'), nlevels (3 ints), report_relmax, verbose,
    use_graph, coarse_direct, pair_cycles.
    """
    lib = _capi.load()
    c = _capi.amg_config()
    lib.amg_config_default(C.byref(c))
    for k, v in kw.items():
        if k == "smoother":
            c.smoother = {"RBPGS": _capi.SMOOTHER_RBPGS, "LJ": _capi.SMOOTHER_LJ}[v] if isinstance(v, str) else v
        elif k == "precision":
            c.precision = {"fp64": _capi.FP64, "mixed": _capi.MIXED}[v] if isinstance(v, str) else v
        elif k == "stop_rule":
            c.stop_rule = {"residual": _capi.STOP_RESIDUAL, "relmax": _capi.STOP_RELMAX}[v] if isinstance(v, str) else v
        elif k in ("pre", "post"):
            arr = c.pre_relax if k == "pre" else c.post_relax
            vals = [v] * _capi.MAX_LEVELS if np.isscalar(v) else list(v)
            for i in range(_capi.MAX_LEVELS):
                arr[i] = int(vals[i] if i < len(vals) else vals[-1])
        elif k == "tbc":
            for i in range(6):
                c.tbc[i] = int(v[i])
        elif k == "vbc":
            for i in range(6):
                c.vbc[i] = float(v[i])
        elif k == "nlevels":
            for i in range(3):
                c.nlevels[i] = int(v[i])
        elif k in ("maxcy", "report_relmax", "verbose", "use_graph", "strict", "coarse_direct",
                   "pair_cycles"):
            setattr(c, k, int(v))
        elif k == "tol":
            c.tol = float(v)
        else:
            raise KeyError(k)
    return c


def mg_read_params(infile, cfg=None):
    """Read ``[MGsolver]``/``[BCs]`` (``mgMK/mg_read_params.F90:44-214``).

    Returns ``(cfg, stat)``; ``stat`` is 0 on success, else the AMG error
    code (the reference ``stop``s on a missing key; here the caller decides).
    """
    lib = _capi.load()
    c = cfg if cfg is not None else default_config()
    stat = lib.amg_read_params(str(infile).encode(), C.byref(c))
    return c, stat


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _dev_ptr(t, grid_shape, name):
    if tuple(t.shape) != tuple(grid_shape) or not t.is_contiguous() or str(t.dtype) != "torch.float64":
        raise ValueError(f"{name}: need a contiguous float64 tensor of shape {grid_shape}")
    if not t.is_cuda:
        raise ValueError(f"{name}: expected a device tensor")
    return C.c_void_p(t.data_ptr())


def _sync_torch():
    import torch
    torch.cuda.current_stream().synchronize()


class PressureSolver:
    """One resident Poisson problem: ``MG_init`` builds it, ``mg_solver`` solves it."""

    def __init__(self, grid, cfg=None):
        self._lib = _capi.load()
        self.grid = grid
        self.cfg = cfg if cfg is not None else default_config()
        self._gc = grid.to_c()
        h = C.c_void_p()
        check(self._lib.amg_create(C.byref(self.cfg), C.byref(self._gc), C.byref(h)), "MG_init")
        self._h = h
        self.stats = _capi.amg_stats()

    # -- mg_solver ------------------------------------------------------------
    def mg_solver(self, phif, rhsf):
        """Solve ``A p = rhs``; ``phif`` is the warm start in and the solution out.

        Returns ``(iter, relmax)`` like ``mgMK/mg_solver.F90``.  On
        non-convergence the reference prints a WARNING and returns with
        ``iter = maxcy+1``; so does this (the WARNING goes to stderr).  Other
        failures raise ``AMGError``.  ``rhsf`` is not modified.
        """
        it = C.c_int(0)
        rel = C.c_double(0.0)
        shape = self.grid.shape
        if _is_torch(phif):
            _sync_torch()
            rc = self._lib.amg_solve_device(self._h, _dev_ptr(phif, shape, "phif"),
                                            _dev_ptr(rhsf, shape, "rhsf"), C.byref(it),
                                            C.byref(rel), C.byref(self.stats))
        else:
            if phif.shape != shape or phif.dtype != np.float64 or not phif.flags.c_contiguous:
                raise ValueError(f"phif: need a C-contiguous float64 array of shape {shape}")
            rhs = np.ascontiguousarray(rhsf, dtype=np.float64)
            if rhs.shape != shape:
                raise ValueError(f"rhsf: need shape {shape}")
            dp = C.POINTER(C.c_double)
            rc = self._lib.amg_solve_host(self._h, phif.ctypes.data_as(dp), rhs.ctypes.data_as(dp),
                                          C.byref(it), C.byref(rel), C.byref(self.stats))
        if rc not in (_capi.AMG_OK, _capi.AMG_ENOTCONV):
            raise AMGError(rc, "mg_solver")
        return it.value, rel.value

    # -- projection -------------------------------------------------------------
    def psource(self, U, V, W, SRC, dt):
        """``SRC = D(rho u*)/dt`` (``dns/NSsolver.F90:2204-2219``); device tensors."""
        s = self.grid.shape
        _sync_torch()
        check(self._lib.amg_psource_device(self._h, _dev_ptr(U, s, "U"), _dev_ptr(V, s, "V"),
                                           _dev_ptr(W, s, "W"), _dev_ptr(SRC, s, "SRC"), float(dt)),
              "Psource")
        return 0

    def pgrad(self, p, U, V, W, dt):
        """``u -= dt G(p/rho)`` (``dns/NSsolver.F90:2057-2082``); device tensors, in place."""
        s = self.grid.shape
        _sync_torch()
        check(self._lib.amg_pgrad_device(self._h, _dev_ptr(p, s, "p"), _dev_ptr(U, s, "U"),
                                         _dev_ptr(V, s, "V"), _dev_ptr(W, s, "W"), float(dt)),
              "Pgrad")

    def pgrad_ghost(self, p, U, V, W, dt, bc):
        """``Pgrad`` complete (``dns/NSsolver.F90:2023-2085``): the correction of
        ``pgrad`` then ``ghost(U,'utemp')``, ``ghost(V,'vtemp')``,
        ``ghost(W,'wtemp')`` with the flow BCs ``bc`` (a ``FlowBC``)."""
        s = self.grid.shape
        _sync_torch()
        cbc = bc.to_c()
        check(self._lib.amg_pgrad_ghost_device(self._h, _dev_ptr(p, s, "p"), _dev_ptr(U, s, "U"),
                                               _dev_ptr(V, s, "V"), _dev_ptr(W, s, "W"), float(dt),
                                               C.byref(cbc)), "Pgrad")

    def ghost(self, var, vtype, bc):
        """``ghost(var, vtype, stat)`` (``dns/BC.F90:1-187``) on a device field."""
        s = self.grid.shape
        _sync_torch()
        cbc = bc.to_c()
        check(self._lib.amg_ghost_device(self._h, _dev_ptr(var, s, "var"), vtype.encode(),
                                         C.byref(cbc)), "ghost")

    def var_advance_CN_PL(self, veltemp1, veltemp2, f1, f2, f3, rkdt, var, bc, rRe, rPr=1.0,
                          rSc=1.0, p=None):
        """``var_advance_CN_PL(veltemp1,veltemp2,f1,f2,f3,rkdt,var)``
        (``dns/NSsolver.F90:3054-3465``): Crank-Nicolson in z of ``veltemp2``
        (in place).  ``p`` given is the reference's ``pcorrect``; ``rRe``,
        ``rPr``, ``rSc`` are 1/Re, 1/Pr, 1/Sc."""
        s = self.grid.shape
        _sync_torch()
        cbc = bc.to_c()
        pp = _dev_ptr(p, s, "p") if p is not None else C.c_void_p()
        check(self._lib.amg_cn_advance_device(self._h, _dev_ptr(veltemp1, s, "veltemp1"),
                                              _dev_ptr(veltemp2, s, "veltemp2"), float(f1), float(f2),
                                              float(f3), float(rkdt), int(var), pp, C.byref(cbc),
                                              float(rRe), float(rPr), float(rSc)),
              "var_advance_CN_PL")

    def divergence(self, U, V, W, tempField=None):
        """Divergence report (``dns/Routines.F90:123-173``).

        Returns ``(avg, maxabs, (i, j, k))`` with 1-based Fortran indices;
        fills ``tempField`` with the cell divergence when given.
        """
        s = self.grid.shape
        _sync_torch()
        avg = C.c_double()
        mx = C.c_double()
        loc = (C.c_int * 3)()
        tf = _dev_ptr(tempField, s, "tempField") if tempField is not None else C.c_void_p()
        check(self._lib.amg_divergence_device(self._h, _dev_ptr(U, s, "U"), _dev_ptr(V, s, "V"),
                                              _dev_ptr(W, s, "W"), tf, C.byref(avg), C.byref(mx),
                                              loc), "divergence")
        return avg.value, mx.value, (loc[0], loc[1], loc[2])

    # -- info / measurement --------------------------------------------------------
    @property
    def levels(self):
        n = C.c_int()
        sizes = ((C.c_int * 3) * _capi.MAX_LEVELS)()
        check(self._lib.amg_levels(self._h, C.byref(n), sizes), "amg_levels")
        return [tuple(sizes[l]) for l in range(n.value)]

    @property
    def direct_level(self):
        """The level solved by the separable direct solve, or -1."""
        return self._lib.amg_direct_level(self._h)

    @property
    def stream(self):
        return self._lib.amg_stream(self._h)

    def run_kernel(self, which, level, reps):
        """Launch one hierarchy kernel ``reps`` times (no sync); returns its bytes per launch."""
        b = C.c_double()
        check(self._lib.amg_run_kernel(self._h, which, level, self.cfg.precision, reps, C.byref(b)),
              "amg_run_kernel")
        return b.value

    def level_field(self, level, field, values=None):
        """Read (values None) or write a level field in natural (k, j, i) order, FP64."""
        if level == -1:
            n = (self.grid.nz, self.grid.ny, self.grid.nx)
        else:
            lv = self.levels[level]
            n = (lv[2], lv[1], lv[0])
        dp = C.POINTER(C.c_double)
        if values is None:
            buf = np.empty(n, dtype=np.float64)
            check(self._lib.amg_level_field(self._h, level, field, 0, buf.ctypes.data_as(dp)),
                  "amg_level_field")
            return buf
        buf = np.ascontiguousarray(values, dtype=np.float64).reshape(n)
        check(self._lib.amg_level_field(self._h, level, field, 1, buf.ctypes.data_as(dp)),
              "amg_level_field")
        return None

    def sync(self):
        """Wait for the handle's stream."""
        import torch
        torch.cuda.ExternalStream(self.stream).synchronize()

    def close(self):
        """``mg_end``: free the device state."""
        if getattr(self, "_h", None):
            self._lib.amg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- the reference's subroutine names ------------------------------------------
def mg_init(grid, cfg):
    """``MG_init(stat)``: returns ``(solver, stat)``; stat -1 for bad boundary codes."""
    try:
        return PressureSolver(grid, cfg), 0
    except AMGError as e:
        if e.code == _capi.AMG_EBADBC:
            return None, -1
        return None, e.code


def mg_solver(solver, phif, rhsf):
    """``mg_solver(phif, rhsf, iter, relmax)``: returns ``(iter, relmax)``."""
    return solver.mg_solver(phif, rhsf)


def mg_end(solver):
    """``mg_end(ierr)``: returns 0."""
    solver.close()
    return 0


def Psource(solver, U, V, W, SRC, dt):
    return solver.psource(U, V, W, SRC, dt)


def Pgrad(solver, p, U, V, W, dt, bc=None):
    """``Pgrad(p,U,V,W,dt)``; with flow BCs ``bc`` also its three ghost calls."""
    if bc is None:
        return solver.pgrad(p, U, V, W, dt)
    return solver.pgrad_ghost(p, U, V, W, dt, bc)


def divergence(solver, U, V, W, tempField=None):
    return solver.divergence(U, V, W, tempField)


def ghost(solver, var, vtype, bc):
    """``ghost(var, vtype, stat)`` (``dns/BC.F90:1-187``)."""
    return solver.ghost(var, vtype, bc)


def var_advance_CN_PL(solver, veltemp1, veltemp2, f1, f2, f3, rkdt, var, bc, rRe, rPr=1.0, rSc=1.0,
                      p=None):
    """``var_advance_CN_PL`` (``dns/NSsolver.F90:3054-3465``)."""
    return solver.var_advance_CN_PL(veltemp1, veltemp2, f1, f2, f3, rkdt, var, bc, rRe, rPr, rSc, p)


# ---- batched and served solves ------------------------------------------------------
class BatchSolver:
    """Independent problems of uneven sizes advanced together (one launch per level step)."""

    def __init__(self, grids, cfg=None):
        self._lib = _capi.load()
        self.grids = list(grids)
        self.cfg = cfg if cfg is not None else default_config()
        arr = (_capi.amg_grid * len(self.grids))(*[g.to_c() for g in self.grids])
        self._arr = arr
        h = C.c_void_p()
        check(self._lib.amg_batch_create(len(self.grids), C.byref(self.cfg), arr, C.byref(h)),
              "amg_batch_create")
        self._h = h
        self.stats = _capi.amg_stats()

    def solve(self, phis, rhss):
        """Solve every instance (device tensors); returns (iters, relres) lists."""
        n = len(self.grids)
        for q in range(n):
            _dev_ptr(phis[q], self.grids[q].shape, f"phi[{q}]")
            _dev_ptr(rhss[q], self.grids[q].shape, f"rhs[{q}]")
        _sync_torch()
        P = (C.c_void_p * n)(*[t.data_ptr() for t in phis])
        R = (C.c_void_p * n)(*[t.data_ptr() for t in rhss])
        it = (C.c_int * n)()
        rr = (C.c_double * n)()
        rc = self._lib.amg_batch_solve_device(self._h, P, R, it, rr, C.byref(self.stats))
        if rc not in (_capi.AMG_OK, _capi.AMG_ENOTCONV):
            raise AMGError(rc, "amg_batch_solve_device")
        return list(it), list(rr)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.amg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching solve server (see ``amg_server_*`` in ``include/amg.h``)."""

    def __init__(self, cfg=None, hbm_budget=200e9, max_slots=64, max_batch=64):
        self._lib = _capi.load()
        self.cfg = cfg if cfg is not None else default_config()
        h = C.c_void_p()
        check(self._lib.amg_server_create(C.byref(self.cfg), float(hbm_budget), int(max_slots),
                                          int(max_batch), C.byref(h)), "amg_server_create")
        self._h = h
        self._keep = []

    def add_member(self, grid):
        g = grid.to_c()
        self._keep.append((grid, g))
        m = self._lib.amg_server_add_member(self._h, C.byref(g))
        if m < 0:
            raise AMGError(m, "amg_server_add_member")
        return m

    def submit(self, member, rhs, out):
        """Queue a solve; ``rhs``/``out`` are device tensors kept alive until done."""
        self._keep.append((rhs, out))
        r = self._lib.amg_server_submit(self._h, int(member), C.c_void_p(rhs.data_ptr()),
                                        C.c_void_p(out.data_ptr()))
        if r < 0:
            raise AMGError(r, "amg_server_submit")
        return r

    def poll(self, req):
        """None while pending, else ``(status, iters, relres, latency_ms)``."""
        it = C.c_int()
        rr = C.c_double()
        lat = C.c_double()
        rc = self._lib.amg_server_poll(self._h, int(req), C.byref(it), C.byref(rr), C.byref(lat))
        if rc == 0:
            return None
        return (0 if rc == 1 else rc), it.value, rr.value, lat.value

    def drain(self):
        check(self._lib.amg_server_drain(self._h), "amg_server_drain")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.amg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
