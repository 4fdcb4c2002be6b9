"""Host-side checks of the C ABI (no GPU needed).

The library loads, exports every entry point that ``include/amg.h``
declares, and its host logic (defaults, the reference-format INI reader,
argument validation) behaves like the reference's ``mg_read_params``.
"""
import os
import re
import subprocess

import pytest

import anelastic_solve_amd as A
from anelastic_solve_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "amg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(amg_[a-z_0-9]+)\s*\(", src)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _capi.load()
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_capi.EXPORTED) == declared
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = sorted(set(re.findall(r" T (amg_\w+)", out)))
    assert exported == declared


def test_library_targets_gfx950():
    """The library built here embeds a gfx950 code object (offload bundle target id)."""
    blob = open(_capi.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_config_defaults():
    c = A.default_config()
    assert c.smoother == _capi.SMOOTHER_LJ and c.maxcy == 100 and c.tol == 1e-10
    assert list(c.tbc) == [0, 1, 0, 1, 0, 0]
    assert c.precision == _capi.MIXED and list(c.pre_relax)[:4] == [2, 2, 2, 2]
    assert c.coarse_direct == 1 and c.pair_cycles == 1


INI = """
[MGsolver]
nxlevels=4
nylevels=3
nzlevels=5
Smoother=RBPGS
maxcy=99
tolmax=1.0d-4
PreRelax=14,10,4,4,2
PstRelax=12,10,4,4,2
Strict=yes
Verbose=no

[BCs]
X1=periodic
X2=read
X3=read

[X2min]
PTYPE=Neumann
PVALUE=0.5d0
[X2max]
PTYPE=Dirichlet
PVALUE=-1.5
[X3min]
PTYPE=Dirichlet
PVALUE=0.d0
[X3max]
PTYPE=neu
PVALUE=2.0d-1
"""


def test_read_params_reference_format(tmp_path):
    p = tmp_path / "Sturb.ini"
    p.write_text(INI)
    c, stat = A.mg_read_params(p)
    assert stat == 0
    assert c.smoother == _capi.SMOOTHER_RBPGS and c.maxcy == 99
    assert c.tol == 1e-4 and c.stop_rule == _capi.STOP_RELMAX
    assert list(c.nlevels) == [4, 3, 5]
    assert list(c.pre_relax)[:5] == [14, 10, 4, 4, 2] and list(c.post_relax)[:5] == [12, 10, 4, 4, 2]
    assert c.strict == 1 and c.verbose == 0
    # face map of mgMK/mg_read_params.F90: [X2max, X3min, X2min, X3max, X1min, X1max]
    assert list(c.tbc) == [2, 2, 1, 1, 0, 0]
    assert list(c.vbc) == [-1.5, 0.0, 0.5, 0.2, -1.0, -1.0]


@pytest.mark.parametrize("edit,code", [
    (lambda s: s.replace("Smoother=RBPGS", "Smoother=XYZ"), _capi.AMG_EPARSE),
    (lambda s: s.replace("PreRelax=14,10,4,4,2", "PreRelax=14,10"), _capi.AMG_EPARSE),
    (lambda s: s.replace("maxcy=99\n", ""), _capi.AMG_EPARSE),
    (lambda s: s.replace("Strict=yes", "Strict=maybe"), _capi.AMG_EPARSE),
    (lambda s: s.replace("PTYPE=Neumann", "PTYPE=Robin"), _capi.AMG_EPARSE),
])
def test_read_params_errors(tmp_path, edit, code):
    p = tmp_path / "bad.ini"
    p.write_text(edit(INI))
    _, stat = A.mg_read_params(p)
    assert stat == code
    assert _capi.load().amg_last_error()


def test_read_params_missing_file():
    _, stat = A.mg_read_params("/nonexistent/Sturb.ini")
    assert stat == _capi.AMG_EPARSE


def test_tolres_selects_residual_rule(tmp_path):
    p = tmp_path / "r.ini"
    p.write_text(INI.replace("tolmax=1.0d-4", "tolmax=1.0d-4\ntolres=1d-10\nPrecision=fp64"))
    c, stat = A.mg_read_params(p)
    assert stat == 0 and c.stop_rule == _capi.STOP_RESIDUAL and c.tol == 1e-10
    assert c.precision == _capi.FP64


def test_coarse_direct_and_pair_cycles_keys(tmp_path):
    """The two optional [MGsolver] switches of this build: yes/no as the
    reference's logical keys (Strict, Verbose), anything else a parse error."""
    p = tmp_path / "k.ini"
    p.write_text(INI.replace("tolmax=1.0d-4", "tolmax=1.0d-4\nCoarseDirect=no\nPairCycles=no"))
    c, stat = A.mg_read_params(p)
    assert stat == 0 and c.coarse_direct == 0 and c.pair_cycles == 0
    p.write_text(INI)
    c, stat = A.mg_read_params(p)
    assert stat == 0 and c.coarse_direct == 1 and c.pair_cycles == 1
    p.write_text(INI.replace("tolmax=1.0d-4", "tolmax=1.0d-4\nPairCycles=sometimes"))
    _, stat = A.mg_read_params(p)
    assert stat == _capi.AMG_EPARSE


def test_errors_have_names():
    lib = _capi.load()
    for code in range(-11, 1):
        assert lib.amg_strerror(code) not in (b"", b"unknown error")
