"""The Fortran drop-in (ISO_C_BINDING shim with the reference's subroutine
names and argument lists) driven by a Fortran test program.

CPU: the bind(C) derived types have the C structs' sizes, and an INI file
read through the Fortran ``mg_read_params`` gives the same controls as the
C reader.  GPU: the closed-form case ``dirichletBCs`` of the reference's
test program (mgMK/test/test.F90:150-189) solved through
``mg_read_params -> MG_init -> mg_solver -> mg_end`` from Fortran.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from anelastic_solve_amd import _capi
from anelastic_solve_amd import build as B

from test_capi import INI


def _driver():
    assert os.path.exists(B.FDRIVER), "Fortran driver not built (run __graft_entry__.build())"
    return B.FDRIVER


def _run(*args, cwd=None):
    return subprocess.run([_driver(), *args], capture_output=True, text=True, timeout=120, cwd=cwd)


def test_bind_c_layouts_match_c():
    r = _run("layout")
    assert r.returncode == 0, r.stderr
    rows = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}
    assert [int(x) for x in rows["LAYOUT"]] == [
        ctypes.sizeof(_capi.amg_config), ctypes.sizeof(_capi.amg_grid),
        ctypes.sizeof(_capi.amg_stats), ctypes.sizeof(_capi.amg_flowbc)]
    # TB(2,2,5) and TB(3,2,12) at the flat indexes the C side reads
    # ((d-1) + 3*(f-1) + 6*(v-1)), and the values beside them
    from anelastic_solve_amd.flowbc import FlowBC
    TB = np.zeros((3, 2, 12), dtype=np.int32)
    for d, f, v in np.ndindex(3, 2, 12):
        TB[d, f, v] = 100 * (d + 1) + 10 * (f + 1) + v + 1
    c = FlowBC(TB, -TB.astype(float)).to_c()
    assert [int(x) for x in rows["FLOWBC"][:3]] == [c.tb[0], c.tb[1 + 3 * 1 + 6 * 4], c.tb[71]] == [111, 225, 332]
    assert float(rows["FLOWBC"][3]) == c.vb[71]


def test_read_params_through_fortran(tmp_path):
    import anelastic_solve_amd as A
    p = tmp_path / "Sturb.ini"
    p.write_text(INI)
    r = _run("params", str(p))
    assert r.returncode == 0, r.stderr
    rows = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}
    assert rows["STAT"] == ["0"]
    c, stat = A.mg_read_params(p)
    assert stat == 0
    assert [int(x) for x in rows["CTRL"]] == [c.smoother, c.maxcy, c.stop_rule, c.precision,
                                              c.strict]
    assert [int(x) for x in rows["LEVELS"]] == list(c.nlevels)
    assert float(rows["TOL"][0]) == c.tol
    assert [int(x) for x in rows["TBC"]] == list(c.tbc)
    assert [float(x) for x in rows["VBC"]] == list(c.vbc)
    assert [int(x) for x in rows["PRE"]] == list(c.pre_relax)
    assert [int(x) for x in rows["POST"]] == list(c.post_relax)


def test_read_params_error_through_fortran(tmp_path):
    p = tmp_path / "bad.ini"
    p.write_text(INI.replace("Smoother=RBPGS", "Smoother=XYZ"))
    r = _run("params", str(p))
    assert r.returncode == 0
    assert r.stdout.split()[:2] == ["STAT", str(_capi.AMG_EPARSE)]
    assert "UNKNOWN" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("direction", ["x1", "x3"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dirichlet_closed_form_through_fortran(cuda, tmp_path, direction, precision):
    r = _run("dirichlet", direction, precision, str(tmp_path / "ptest.ini"))
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split()[:3] == ["PASS", direction, precision]


def test_horizontally_varying_density_is_refused(tmp_path):
    """The reference's Flow::r is 3-D (include/modules.F90:45); the shim
    passes one column of it and refuses a density that varies horizontally
    (AMG_EUNSUPPORTED from MG_init, before any device work)."""
    r = _run("rho3d", str(tmp_path / "rho3d.ini"))
    assert r.returncode == 0, r.stderr
    assert r.stdout.split()[:2] == ["STAT", str(_capi.AMG_EUNSUPPORTED)]
    assert "varies horizontally" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_relmax_stop_rule_through_fortran(cuda, tmp_path, precision):
    """An INI with only tolmax (the reference's own stop rule,
    mgMK/mg_solver.F90:114-128) converges on the relative solution change in
    both precisions: FP64 computes the change although the Fortran caller
    cannot ask for it."""
    r = _run("relmax", precision, str(tmp_path / "relmax.ini"))
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split()[:3] == ["PASS", "x3", precision]
