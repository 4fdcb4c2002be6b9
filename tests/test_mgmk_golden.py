"""Parity with the reference's own solver (mgMK), through fixtures it wrote.

``tests/golden/mgmk_uniform_{16,32}.npz`` hold the reference's ``mg_solver``
result on the uniform config-1-style problem (2*pi x 2*pi x 1, periodic x/y,
Neumann z; built from its Fortran sources and run by
``tests/golden/make_mgmk_golden.py``).  On a uniform grid with rho = 1 the
reference's fine operator (``mgMK/mg_pfpde.F90:35-57`` with dl = L/n) is
this build's, so after removing the null space (the reference lets the
pressure's mean drift) the three answers must agree:

* CPU: the oracle's exact direct solve == the reference's solution (pins
  the oracle's operator and solver against the real reference);
* GPU: the HIP solver (FP64 and mixed, RBPGS and LJ) == the reference's
  solution to the 1e-9 bar of the pressure.
"""
import os

import numpy as np
import pytest

from conftest import oracle_levels, relerr

HERE = os.path.dirname(os.path.abspath(__file__))
TBC = (0, 1, 0, 1, 0, 0)
SIZES = (16, 32)


def _fixture(n):
    d = np.load(os.path.join(HERE, "golden", f"mgmk_uniform_{n}.npz"))
    return {k: d[k] for k in d.files}


def _grid(n):
    from anelastic_solve_amd import box_grid
    return box_grid(n, n, n, beta=0.0)


def _null(x, g):
    from oracle import ops
    levs, _ = oracle_levels(g, TBC)
    return ops.null_project(x, levs[0]), levs


@pytest.mark.parametrize("n", SIZES)
def test_reference_converged(n):
    f = _fixture(n)
    assert f["relres"] <= 1e-12 and f["rhs"].shape == (n, n, n)
    # the reference's V-cycle contracts, slowly (its coarse operators are not
    # the re-discretised ones): the residual falls over its first cycles
    tr = f["trace_relres"]
    assert tr[-1] < tr[0] and np.all(np.isfinite(tr))


@pytest.mark.parametrize("n", SIZES)
def test_oracle_direct_equals_reference(n):
    from oracle import direct
    f = _fixture(n)
    g = _grid(n)
    ref, levs = _null(f["phi"], g)
    x = direct.solve(levs[0], f["rhs"], list(TBC))
    err = relerr(x, ref)
    print(f"{n}^3: oracle vs reference {err:.2e}")
    assert err <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("precision,smoother", [("fp64", "RBPGS"), ("fp64", "LJ"),
                                                ("mixed", "RBPGS"), ("mixed", "LJ")])
def test_gpu_equals_reference(cuda, n, precision, smoother):
    torch = cuda
    from anelastic_solve_amd import PressureSolver, default_config
    f = _fixture(n)
    g = _grid(n)
    ref, _ = _null(f["phi"], g)
    b = np.zeros(g.shape)
    b[1:-1, 1:-1, 1:-1] = f["rhs"]
    s = PressureSolver(g, default_config(precision=precision, smoother=smoother, tol=1e-12,
                                         maxcy=300))
    p = torch.zeros(g.shape, dtype=torch.float64, device="cuda")
    it, _ = s.mg_solver(p, torch.from_numpy(b).cuda())
    assert s.stats.converged == 1, it
    s.close()
    err = relerr(p.cpu().numpy()[1:-1, 1:-1, 1:-1], ref)
    print(f"{n}^3 {precision} {smoother}: {it} cycles, vs reference {err:.2e}")
    assert err <= 1e-9
