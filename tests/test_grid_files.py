"""Grid-file ingestion: the reference's ``x{1,2,3}_grid.in`` format
(``utils/loadGrid.F90:38-209``) read into the solver.

CPU: the committed fixtures (``tests/golden/grid_32x16x24``, written by
``tests/golden/make_grid_files.py``) load back bit-exactly to the axes they
were written from, satisfy loadGrid's metric identities, and -- where the
reference tree is present in this container -- its own sample grid files
(``run_dir/x{1,2,3}_grid.in``) parse with the same identities.  GPU: a
solve on the ingested stretched grid agrees with the oracle's direct solve.
"""
import math
import os

import numpy as np
import pytest

from conftest import oracle_levels, relerr

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "grid_32x16x24")
REF_RUN_DIR = "/root/reference/ANELASTIC_SL/run_dir"


def _identities(ax):
    """loadGrid's definitions on the interior, its end-cell fix-ups on the ghosts."""
    n = ax.n
    i = np.arange(1, n + 1)
    np.testing.assert_allclose(ax.xc[i], 0.5 * (ax.xe[i] + ax.xe[i - 1]), rtol=0, atol=1e-12 * ax.xe[n])
    np.testing.assert_allclose(ax.dxc[i], ax.xe[i] - ax.xe[i - 1], rtol=1e-12)
    np.testing.assert_allclose(ax.dxe[:n], ax.xc[1:n + 1] - ax.xc[:n], rtol=1e-12)
    assert ax.dxc[0] == ax.dxc[1] and ax.dxc[n + 1] == ax.dxc[n]
    assert ax.dxe[0] == ax.dxe[1] and ax.dxe[n + 1] == ax.dxe[n]


def test_fixture_roundtrip_exact():
    from anelastic_solve_amd.grid import load_grid_file, tanh_axis, uniform_axis
    want = [(uniform_axis(32, 2 * math.pi), 2 * math.pi), (uniform_axis(16, math.pi), math.pi),
            (tanh_axis(24, 1.0, 1.5), 1.0)]
    for d, (ax, L) in enumerate(want, start=1):
        got, gL, x0 = load_grid_file(os.path.join(GOLDEN, f"x{d}_grid.in"))
        assert gL == L and x0 == 0.0 and got.n == ax.n
        for f in ("xe", "xc", "dxe", "dxc"):
            assert np.array_equal(getattr(got, f), getattr(ax, f)), (d, f)
        _identities(got)


def test_writer_reader_roundtrip(tmp_path):
    from anelastic_solve_amd.grid import load_grid_file, tanh_axis, write_grid_file
    ax = tanh_axis(40, 3.0, 2.0, x0=-1.5)
    p = tmp_path / "x3_grid.in"
    write_grid_file(p, ax, 3.0, -1.5, "X3")
    got, L, x0 = load_grid_file(p)
    assert (L, x0) == (3.0, -1.5)
    for f in ("xe", "xc", "dxe", "dxc"):
        assert np.array_equal(getattr(got, f), getattr(ax, f)), f


@pytest.mark.skipif(not os.path.isdir(REF_RUN_DIR), reason="reference tree not present")
def test_reference_sample_grid_files():
    from anelastic_solve_amd.grid import load_grid_dir, load_grid_file
    sizes = []
    for d in (1, 2, 3):
        ax, L, x0 = load_grid_file(os.path.join(REF_RUN_DIR, f"x{d}_grid.in"))
        _identities(ax)
        assert np.all(ax.dxc > 0) and L > 0
        sizes.append(ax.n)
    assert sizes == [256, 64, 256]
    g = load_grid_dir(REF_RUN_DIR)
    assert (g.nx, g.ny, g.nz) == (256, 64, 256)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_on_ingested_grid(cuda, precision):
    torch = cuda
    from anelastic_solve_amd import PressureSolver, default_config, workloads
    from anelastic_solve_amd.grid import load_grid_dir
    from oracle import direct, ops
    g = load_grid_dir(GOLDEN)
    tbc = (0, 1, 0, 1, 0, 0)
    b = workloads.psource_host(*workloads.velocity_white(g, 7), g)
    s = PressureSolver(g, default_config(precision=precision, tol=1e-12, maxcy=80))
    p = torch.zeros(g.shape, dtype=torch.float64, device="cuda")
    it, _ = s.mg_solver(p, torch.from_numpy(b).cuda())
    assert s.stats.converged == 1 and s.stats.relres <= 1e-12, (it, s.stats.relres)
    levs, _ = oracle_levels(g, tbc, s.levels)
    bb = ops.compat_project(b[1:-1, 1:-1, 1:-1], levs[0])
    x = direct.solve(levs[0], bb, list(tbc))
    assert relerr(p.cpu().numpy()[1:-1, 1:-1, 1:-1], x) <= 1e-9
    s.close()
