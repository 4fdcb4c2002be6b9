"""Failure behaviour and the reported convergence quantities.

CPU (host-side validation runs before any device work, so these need no
GPU): the size and boundary checks of ``MG_init`` (mgMK/mg_check_size.F90:
43-151 and mgMK/mg_init.F90:265-305) and this build's own limits.

GPU: relmax is the reference's ``max|p_new - p_old| / max|p_new|``
(mgMK/mg_residual.F90:60-94) of the last cycle, recomputed here from the
iterates; ``maxcy`` exhaustion returns ``iter = maxcy+1`` after the
reference's WARNING line (mgMK/mg_solver.F90:127-132); a NaN or Inf in the
right-hand side is AMG_ENAN; the FP64 solver with the reference's relmax
stop rule converges without being asked to report relmax; a narrow, tall
grid (one block partial per 16-plane slab) solves correctly; a served
member whose hierarchy cannot fit the HBM budget fails instead of waiting.
"""
import numpy as np
import pytest

from conftest import relerr


def _create(n, beta=0.0, **kw):
    from anelastic_solve_amd import PressureSolver, box_grid, default_config
    return PressureSolver(box_grid(*n, beta=beta), default_config(**kw))


def _code(n, **kw):
    from anelastic_solve_amd._capi import AMGError
    with pytest.raises(AMGError) as e:
        _create(n, **kw)
    return e.value.code, str(e.value)


@pytest.mark.parametrize("n,kw,code,msg", [
    ((18, 16, 16), dict(nlevels=(3, 3, 3)), -2, "not a multiple of 2**(levels-1)"),
    ((16, 16, 20), dict(nlevels=(3, 3, 4)), -2, "not a multiple of 2**(levels-1)"),
    ((64, 64, 64), dict(nlevels=(2, 2, 2)), -2, "at most 1024 are solved directly"),
    ((6, 8, 8), dict(nlevels=(2, 2, 2)), -2, "nx must be even on every level"),
    ((64, 64, 512), dict(smoother="LJ"), -2, "supports nz <= 256"),
    ((16, 16, 16), dict(tbc=(0, 1, 0, 1, 0, 3)), -3, "BAD BOUNDARY CONDITION CHECK 6"),
    ((16, 16, 16), dict(tbc=(0, 1, 0, 1, 0, 1)), -3, "periodic on both faces"),
])
def test_create_rejects(n, kw, code, msg):
    c, text = _code(n, **kw)
    assert c == code and msg in text, (c, text)


def test_create_rejects_bad_density():
    from anelastic_solve_amd import PressureSolver, box_grid, default_config
    from anelastic_solve_amd._capi import AMGError, AMG_EINVAL
    r = np.ones(18)
    r[5] = 0.0
    with pytest.raises(AMGError) as e:
        PressureSolver(box_grid(16, 16, 16, rho=r), default_config())
    assert e.value.code == AMG_EINVAL and "rho must be positive" in str(e.value)


# ---- GPU -------------------------------------------------------------------------
def _rhs(g, seed=11):
    from anelastic_solve_amd import workloads
    return workloads.psource_host(*workloads.velocity_white(g, seed), g)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_relmax_is_the_solution_change_of_the_last_cycle(cuda, precision):
    """Dirichlet walls (no null-space shift on return): one cycle from a
    warm start p1 gives p2; the reported relmax is max|p2-p1|/max|p2|."""
    torch = cuda
    tbc = (0, 2, 0, 2, 0, 0)
    vbc = (-1.0, 0.0, -1.0, 0.0, -1.0, -1.0)
    s = _create((32, 32, 32), beta=1.5, precision=precision, tbc=tbc, vbc=vbc, tol=1e-14,
                maxcy=1, report_relmax=1)
    b = torch.from_numpy(_rhs(s.grid)).cuda()
    p = torch.zeros(s.grid.shape, dtype=torch.float64, device="cuda")
    it, rel0 = s.mg_solver(p, b)
    assert it == 2 and rel0 == 1.0            # from zero: the change is the iterate
    for _ in range(3):
        p1 = p[1:-1, 1:-1, 1:-1].clone()
        it, rel = s.mg_solver(p, b)
        p2 = p[1:-1, 1:-1, 1:-1]
        ref = ((p2 - p1).abs().max() / p2.abs().max()).item()
        assert it == 2
        if precision == "fp64":
            assert rel == ref, (rel, ref)      # the same FP64 differences, the same maximum
        else:
            # mixed: max|e| of the FP32 correction e, and p2 = p1 + e rounded in FP64
            assert abs(rel - ref) <= 1e-8 * ref, (rel, ref)
    s.close()


@pytest.mark.gpu
def test_maxcy_exhaustion_warns_and_returns_maxcy_plus_one(cuda, capfd):
    torch = cuda
    s = _create((32, 32, 32), beta=2.0, tol=1e-30, maxcy=3)
    b = torch.from_numpy(_rhs(s.grid)).cuda()
    p = torch.zeros(s.grid.shape, dtype=torch.float64, device="cuda")
    it, rel = s.mg_solver(p, b)
    err = capfd.readouterr().err
    assert it == 4 and s.stats.converged == 0
    assert "WARNING: failed to achieve  convergence in       3 cycles" in err
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_nonfinite_rhs_is_enan(cuda, bad):
    torch = cuda
    from anelastic_solve_amd._capi import AMGError, AMG_ENAN
    s = _create((16, 16, 16))
    b = _rhs(s.grid)
    b[5, 6, 7] = bad
    p = torch.zeros(s.grid.shape, dtype=torch.float64, device="cuda")
    with pytest.raises(AMGError) as e:
        s.mg_solver(p, torch.from_numpy(b).cuda())
    assert e.value.code == AMG_ENAN
    # the handle stays usable
    b[5, 6, 7] = 0.0
    s.mg_solver(p, torch.from_numpy(_rhs(s.grid)).cuda())
    assert s.stats.converged == 1
    s.close()


@pytest.mark.gpu
def test_fp64_relmax_rule_without_reporting(cuda):
    """The reference's stop rule (relmax <= tolmax) in FP64 with the default
    report_relmax = 0 converges: the change is computed whenever the rule
    needs it."""
    torch = cuda
    s = _create((32, 32, 32), beta=1.0, precision="fp64", stop_rule="relmax", tol=1e-8,
                maxcy=50, report_relmax=0)
    b = torch.from_numpy(_rhs(s.grid)).cuda()
    p = torch.zeros(s.grid.shape, dtype=torch.float64, device="cuda")
    it, rel = s.mg_solver(p, b)
    assert s.stats.converged == 1 and 1 <= it < 50 and 0.0 < rel <= 1e-8, (it, rel)
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n,nlevels", [((2, 2, 256), (1, 1, 2)), ((2, 2, 64), (1, 1, 2))])
def test_narrow_tall_grid(cuda, n, nlevels):
    """2x2 columns of 256 cells: the FP64 stencil kernels write one partial
    per 16-plane slab (16 blocks for 1024 cells), more than one per 256
    cells; the solve matches the oracle's direct solve."""
    torch = cuda
    from conftest import oracle_levels
    from oracle import direct, ops
    tbc = (0, 1, 0, 1, 0, 0)
    for precision in ("fp64", "mixed"):
        s = _create(n, beta=1.0, nlevels=nlevels, precision=precision, tol=1e-12, maxcy=200)
        b = _rhs(s.grid)
        p = torch.zeros(s.grid.shape, dtype=torch.float64, device="cuda")
        s.mg_solver(p, torch.from_numpy(b).cuda())
        assert s.stats.converged == 1, precision
        levs, _ = oracle_levels(s.grid, tbc, s.levels, nlevels)
        bb = ops.compat_project(b[1:-1, 1:-1, 1:-1], levs[0])
        x = direct.solve(levs[0], bb, list(tbc))
        assert relerr(p.cpu().numpy()[1:-1, 1:-1, 1:-1], x) <= 1e-9, precision
        s.close()


@pytest.mark.gpu
def test_server_refuses_what_the_budget_cannot_hold(cuda):
    torch = cuda
    from anelastic_solve_amd import Server, box_grid, default_config
    from anelastic_solve_amd._capi import AMGError, AMG_ENOMEM
    g = box_grid(64, 64, 64, beta=1.0)
    warm = 66 ** 3 * 8
    srv = Server(default_config(), hbm_budget=warm * 1.2, max_slots=2, max_batch=2)
    m = srv.add_member(g)                     # the warm start fits ...
    with pytest.raises(AMGError) as e:        # ... a second one does not
        srv.add_member(g)
    assert e.value.code == AMG_ENOMEM
    b = torch.from_numpy(_rhs(g)).cuda()
    out = torch.zeros(g.shape, dtype=torch.float64, device="cuda")
    r = srv.submit(m, b, out)                 # its hierarchy never fits: the request fails
    srv.drain()
    st = srv.poll(r)
    assert st is not None and st[0] == AMG_ENOMEM
    srv.close()
