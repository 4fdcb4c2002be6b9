"""The anelastic stratified-layer production configuration
(``ANELASTIC_SL/run_dir``): its INI settings, grid and background density.

CPU: the package's ``configs/anelastic_sl.ini`` gives the same flow BC tables
(``bc()``, ``dns/BC.F90:691-1432``) and the same solver settings
(``mg_read_params``) as the reference's own ``run_dir/Sturb.ini``; the
generated grid equals the run's grid files; the BC reader keeps the
reference reader's value slips and case-sensitive type names; the density
is a decreasing moist-air profile whose constants carry their float32 values.
GPU: the production solve (Dirichlet pressure on the z walls, rho(z),
4/3/8 levels per direction, RBPGS with the run's per-level sweeps) against
the oracle's direct solve, and a CN step with the run's BCs to rounding.
"""
import math
import os

import numpy as np
import pytest

from conftest import oracle_levels, relerr

REF = "/root/reference/ANELASTIC_SL/run_dir"
needs_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")


@needs_ref
def test_config_matches_reference_run():
    from anelastic_solve_amd import workloads
    from anelastic_solve_amd.flowbc import read_bc
    from anelastic_solve_amd.mg import mg_read_params
    ours = read_bc(workloads.ANELASTIC_SL_INI)
    theirs = read_bc(os.path.join(REF, "Sturb.ini"))
    assert np.array_equal(ours.TB, theirs.TB) and np.array_equal(ours.VB, theirs.VB)
    a, sa = mg_read_params(workloads.ANELASTIC_SL_INI)
    b, sb = mg_read_params(os.path.join(REF, "Sturb.ini"))
    assert sa == 0 and sb == 0
    for f, _ in a._fields_:
        va, vb = getattr(a, f), getattr(b, f)
        if hasattr(va, "__len__"):
            va, vb = list(va), list(vb)
        assert va == vb, f


@needs_ref
def test_grid_matches_run_grid_files():
    from anelastic_solve_amd import workloads
    from anelastic_solve_amd.grid import load_grid_file
    g = workloads.anelastic_sl_grid()
    for d, a in ((1, g.ax), (2, g.ay), (3, g.az)):
        ref, _, _ = load_grid_file(os.path.join(REF, "x%d_grid.in" % d))
        for f in ("xe", "xc", "dxe", "dxc"):
            np.testing.assert_allclose(getattr(a, f), getattr(ref, f), rtol=1e-12, atol=1e-9)


def test_production_bc_values():
    from anelastic_solve_amd import workloads
    from anelastic_solve_amd.flowbc import read_bc
    b = read_bc(workloads.ANELASTIC_SL_INI)
    assert np.all(b.TB[0] == 3) and np.all(b.TB[1] == 3)          # periodic x, y
    assert b.TB[2, 0, 3] == 1 and b.TB[2, 1, 3] == 1              # Dirichlet pressure
    assert (b.VB[2, 0, 0], b.VB[2, 1, 0]) == (10.0, -10.0)        # U1 walls
    assert b.VB[2, 0, 4] == 0.0032074536561066 and b.TB[2, 0, 4] == 2
    assert b.TB[2, 0, 6] == 1 and b.VB[2, 0, 6] == 0.0            # pressure-source class


def _ini(tmp_path, x1min, x1max):
    keys = ("U1", "U2", "U3", "P", "RHO", "FLUC", "SCAL1")
    body = ["[BCs]", "X1=read", "X2=periodic", "X3=periodic", "[X1min]"]
    body += ["%sTYPE=%s\n%sVALUE=%d.0d0" % (k, x1min.get(k, "Neumann"), k, 10 + i) for i, k in enumerate(keys)]
    body += ["[X1max]"]
    body += ["%sTYPE=%s\n%sVALUE=%d.0d0" % (k, x1max.get(k, "Dirichlet"), k, 20 + i) for i, k in enumerate(keys)]
    p = tmp_path / "bc.ini"
    p.write_text("\n".join(body) + "\n")
    return str(p)


def test_reader_keeps_the_reference_value_slips(tmp_path):
    from anelastic_solve_amd.flowbc import read_bc
    b = read_bc(_ini(tmp_path, {}, {"U1": "Etpl8", "U3": "expl8"}))
    # X1max: U1's block reads U2VALUE into VB(1,1,2), U2's reads U1VALUE into VB(1,2,2)
    assert b.VB[0, 0, 1] == 21.0 and b.VB[0, 1, 1] == 20.0
    assert b.VB[0, 1, 0] == 0.0                                   # VB(1,2,1) is never read
    assert b.TB[0, 1, 0] == 8 and b.TB[0, 1, 2] == 8
    assert b.VB[0, 0, 0] == 10.0 and b.TB[0, 0, 0] == 2


@pytest.mark.parametrize("face,names", [("max", {"U3": "etpl8"}), ("min", {"U1": "DIRICHLET"}),
                                        ("min", {"P": "Wall6"})])
def test_reader_rejects_what_the_reference_stops_on(tmp_path, face, names):
    from anelastic_solve_amd.flowbc import read_bc
    ini = _ini(tmp_path, names if face == "min" else {}, names if face == "max" else {})
    with pytest.raises(ValueError, match="NO VALID"):
        read_bc(ini)


def test_density_profile():
    from anelastic_solve_amd import workloads
    z = np.linspace(0.0, 5000.0, 51)
    r = workloads.density_bg(z)
    assert np.all(np.diff(r) < 0)
    # moist air at 300 K and one atmosphere: p / (R T) with the vapour correction
    assert abs(r[0] - 101325.0 / (287.1 * 300.0)) / r[0] < 0.01
    # the constants enter at their float32 values: a pure-double evaluation differs
    zz = z
    pz = 101325.0 * np.exp(5.5547 * np.log(46153.8 - zz) - 2.28088 * np.log(317030.0 - zz) - 30.7647)
    tz, qv = 300.0 - 0.0065 * zz, 0.0115 - 0.000002 * zz
    r64 = pz * (1.0 + qv) / (287.1 * tz * (1.0 + qv / (287.1 / 461.5)))
    assert 1e-9 < np.abs(r - r64).max() / r.max() < 1e-5


@pytest.mark.gpu
def test_production_solve_against_oracle(cuda):
    torch = cuda
    from anelastic_solve_amd import PressureSolver, workloads
    from anelastic_solve_amd._capi import AMG_STOP_RESIDUAL
    from oracle import direct, ops
    g, cfg, bc, U, V, W = workloads.anelastic_sl()
    b = workloads.psource_host(U, V, W, g)
    tbc = tuple(cfg.tbc)
    assert tbc == (0, 2, 0, 2, 0, 0)
    # the run's own stop rule (relmax <= 1e-4) first, then a tight residual stop
    s = PressureSolver(g, cfg)
    p = torch.zeros(g.shape, dtype=torch.float64, device="cuda")
    s.mg_solver(p, torch.from_numpy(b).cuda())
    assert s.stats.converged == 1 and s.stats.relmax <= 1e-4
    assert [tuple(l) for l in s.levels][-1] == (32, 16, 2)
    s.close()
    cfg.stop_rule = AMG_STOP_RESIDUAL
    cfg.tol = 1e-12
    s = PressureSolver(g, cfg)
    p = torch.zeros(g.shape, dtype=torch.float64, device="cuda")
    s.mg_solver(p, torch.from_numpy(b).cuda())
    assert s.stats.converged == 1 and s.stats.relres <= 1e-12
    levs, _ = oracle_levels(g, tbc)
    x = direct.solve(levs[0], b[1:-1, 1:-1, 1:-1], list(tbc))
    assert relerr(p.cpu().numpy()[1:-1, 1:-1, 1:-1], x) <= 1e-9
    s.close()


@pytest.mark.gpu
def test_production_cn_step(cuda):
    """The vertical CN step with the run's BCs equals the oracle to rounding
    (1e-13 of the field scale; the columns' recurrences run in segments)."""
    torch = cuda
    from anelastic_solve_amd import PressureSolver, workloads
    from oracle import dns
    g, cfg, bc, U, V, W = workloads.anelastic_sl()
    s = PressureSolver(g, cfg)
    ax = lambda a: {"n": a.n, "xe": a.xe, "xc": a.xc, "dxe": a.dxe, "dxc": a.dxc}
    axes = (ax(g.ax), ax(g.ay), ax(g.az))
    rng = np.random.Generator(np.random.PCG64(8))
    tend = 1e-3 * rng.standard_normal(g.shape)
    for var, field in ((1, U), (5, 1e-3 * rng.standard_normal(g.shape))):
        want = dns.cn_advance(tend, field, 1.0, -5.0 / 9.0, 1.0, 0.5, var, g.rho_or_ones(), axes,
                              bc.TB, bc.VB, 1.0 / 5000.0, 1.0 / 0.7, 1.0)
        d2 = torch.from_numpy(field.copy()).cuda()
        s.var_advance_CN_PL(torch.from_numpy(tend).cuda(), d2, 1.0, -5.0 / 9.0, 1.0, 0.5, var, bc,
                            1.0 / 5000.0, 1.0 / 0.7, 1.0)
        got = d2.cpu().numpy()
        err = float(np.abs(got - want).max() / np.abs(want).max())
        assert err <= 1e-13, (var, err)
    s.close()
