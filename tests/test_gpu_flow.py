"""The time stepper's boundary handling and vertical viscous step on the GPU,
through the C ABI, against ``oracle.dns``.

* ``ghost`` (``dns/BC.F90``): bit-exact for every variable type and a set of
  BC tables that uses every face rule (including mixed periodic/wall faces);
* ``Pgrad`` with its three ghost calls: equal to the oracle's ``pgrad``
  followed by ``ghost`` to 1e-14 of the field scale (the interior update
  is the existing Pgrad kernel, whose multiply-adds may be fused);
* ``var_advance_CN_PL``: equal to rounding (1e-13 of the field scale) for
  u, v, w, rho, scal1 with Neumann, Dirichlet, Wall6 and Wall7 rows, with
  and without the predicted pressure, on a stretched grid with a varying
  rho(z): the kernel evaluates the reference's expressions without
  contraction, as numpy does, and splits each column's Thomas recurrences
  into segments joined by carried values (a few more roundings per row;
  nz = 264 takes the serial kernel);
* a substep tail on the device (CN for u, v, w, then Psource, the solve and
  Pgrad with ghosts) ends divergence-free.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(nx=32, ny=24, nz=20, beta=1.7, rho_amp=0.2):
    from anelastic_solve_amd import box_grid
    z = np.linspace(-0.1, 1.1, nz + 2)
    rho = 1.0 + rho_amp * np.cos(2.0 * z)
    return box_grid(nx, ny, nz, beta=beta, rho=rho)


def _axes(g):
    ax = lambda a: {"n": a.n, "xe": a.xe, "xc": a.xc, "dxe": a.dxe, "dxc": a.dxc}
    return ax(g.ax), ax(g.ay), ax(g.az)


def _solver(g):
    from anelastic_solve_amd import PressureSolver, default_config
    # the z-line smoother takes nz <= 256; the flow steps do not depend on it
    sm = "LJ" if g.nz <= 256 else "RBPGS"
    return PressureSolver(g, default_config(precision="fp64", tol=1e-12, smoother=sm))


def _tables():
    from anelastic_solve_amd import FlowBC
    rng = np.random.Generator(np.random.PCG64(21))
    out = []
    for xt, yt, zt in [(3, 3, 2), (3, 3, 1), (2, 3, 6), (8, 1, 7), (1, 4, 4), (4, 2, 3), (8, 2, 2)]:
        TB = np.zeros((3, 2, 12), dtype=np.int32)
        for d, t in enumerate((xt, yt, zt)):
            TB[d] = t
        if zt == 7:
            TB[2, 0] = 2
        out.append(FlowBC(TB, rng.standard_normal((3, 2, 12))))
    TB = np.full((3, 2, 12), 3, dtype=np.int32)
    TB[0, 1] = 1
    TB[2, 0] = 2
    out.append(FlowBC(TB, rng.standard_normal((3, 2, 12))))
    return out


def test_ghost_bit_exact(cuda):
    torch = cuda
    from oracle import dns
    g = _grid()
    axes = _axes(g)
    s = _solver(g)
    rng = np.random.Generator(np.random.PCG64(5))
    for bc in _tables():
        for vtype in sorted(dns.GHOST_CLASS):
            v = rng.standard_normal(g.shape)
            want = dns.ghost(v.copy(), vtype, bc.TB, bc.VB, axes)
            d = torch.from_numpy(v).cuda()
            s.ghost(d, vtype, bc)
            got = d.cpu().numpy()
            assert np.array_equal(got, want), (vtype, bc.TB[:, :, 0].tolist())
    from anelastic_solve_amd import AMGError
    with pytest.raises(AMGError):
        s.ghost(torch.zeros(g.shape, dtype=torch.float64, device="cuda"), "nope", _tables()[0])
    s.close()


def test_pgrad_with_ghosts(cuda):
    torch = cuda
    from anelastic_solve_amd import FlowBC
    from oracle import dns, ops
    g = _grid()
    axes = _axes(g)
    s = _solver(g)
    bc = FlowBC.make((True, True, False), {(3, 1, 1): (2, 0.0), (3, 2, 1): (2, 0.0),
                                            (3, 1, 3): (1, 0.0), (3, 2, 3): (1, 0.0)})
    rng = np.random.Generator(np.random.PCG64(6))
    p, U, V, W = (rng.standard_normal(g.shape) for _ in range(4))
    dt = 0.013
    walls = (False, False, True)      # Neumann pressure at the z walls
    U1, V1, W1 = ops.pgrad(p, U, V, W, g.rho_or_ones(), *axes, dt, walls=walls)
    for a, t in ((U1, "utemp"), (V1, "vtemp"), (W1, "wtemp")):
        dns.ghost(a, t, bc.TB, bc.VB, axes)
    dev = [torch.from_numpy(a).cuda() for a in (p, U, V, W)]
    s.pgrad_ghost(*dev, dt, bc)
    for got, want in zip(dev[1:], (U1, V1, W1)):
        np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=1e-14 * np.abs(want).max())
    s.close()


CN_CASES = [
    (1, {(3, 1, 1): (2, 0.0), (3, 2, 1): (2, 0.0)}, True),
    (2, {(3, 1, 2): (1, 0.3), (3, 2, 2): (6, -0.2)}, True),
    (3, {(3, 1, 3): (1, 0.0), (3, 2, 3): (2, 0.1)}, True),
    (3, {(3, 1, 3): (6, 0.2), (3, 2, 3): (7, 0.0)}, False),
    (5, {(3, 1, 5): (2, 0.5), (3, 2, 5): (2, -0.5)}, False),
    (8, {(3, 1, 8): (1, 1.0), (3, 2, 8): (2, 0.0)}, False),
    # a 'cfluc' Wall7 rule zeroes row sz (ez) of the system; the wall rows rewrite it here
    (8, {(3, 1, 8): (2, 0.0), (3, 2, 8): (1, 0.5), (3, 1, 6): (7, 0.0), (3, 2, 6): (2, 0.0)}, False),
    (1, {(3, 1, 1): (2, 0.0), (3, 2, 1): (7, 0.0), (3, 1, 6): (2, 0.0), (3, 2, 6): (7, 0.0)}, False),
]


@pytest.mark.parametrize("var,walls,pcorrect", CN_CASES)
@pytest.mark.parametrize("nz", [20, 256, 264])
def test_cn_step(cuda, var, walls, pcorrect, nz):
    torch = cuda
    from anelastic_solve_amd import FlowBC
    from oracle import dns
    g = _grid(nx=32, ny=24 if nz == 20 else 32, nz=nz)
    axes = _axes(g)
    s = _solver(g)
    bc = FlowBC.make((True, True, False), walls)
    rng = np.random.Generator(np.random.PCG64(100 + var))
    v1, v2, p = (rng.standard_normal(g.shape) for _ in range(3))
    args = dict(f1=1.0, f2=0.45, f3=1.0, rkdt=0.02, rRe=1e-2, rPr=1.4, rSc=0.9)
    want = dns.cn_advance(v1, v2, args["f1"], args["f2"], args["f3"], args["rkdt"], var,
                          g.rho_or_ones(), axes, bc.TB, bc.VB, args["rRe"], args["rPr"], args["rSc"],
                          p=p if pcorrect else None)
    d1, d2, dp = (torch.from_numpy(a).cuda() for a in (v1, v2, p))
    s.var_advance_CN_PL(d1, d2, args["f1"], args["f2"], args["f3"], args["rkdt"], var, bc,
                        args["rRe"], args["rPr"], args["rSc"], p=dp if pcorrect else None)
    got = d2.cpu().numpy()
    err = float(np.abs(got - want).max() / np.abs(want).max())
    print(f"var {var} nz {nz}: {err:.2e}")
    assert err <= 1e-13, err
    assert torch.equal(d1, torch.from_numpy(v1).cuda())      # veltemp1 is read only
    s.close()


def test_cn_rejects_what_the_reference_stops_on(cuda):
    torch = cuda
    from anelastic_solve_amd import AMGError, FlowBC
    g = _grid(8, 8, 8)
    s = _solver(g)
    z = torch.zeros(g.shape, dtype=torch.float64, device="cuda")
    bc = FlowBC.make((True, True, False), {(3, 1, 1): (4, 0.0), (3, 2, 1): (1, 0.0)})
    with pytest.raises(AMGError):
        s.var_advance_CN_PL(z.clone(), z, 1.0, 0.5, 1.0, 0.01, 1, bc, 1e-2)
    with pytest.raises(AMGError):
        s.var_advance_CN_PL(z.clone(), z, 1.0, 0.5, 1.0, 0.01, 4, bc, 1e-2)
    # a 'cfluc' Wall7 rule that zeroes a row no wall row rewrites: the reference
    # divides by zero (the oracle's result is not finite); here it is an error
    from oracle import dns
    bc = FlowBC.make((True, True, False), {(3, 1, 1): (2, 0.0), (3, 2, 1): (1, 0.0), (3, 1, 6): (7, 0.0)})
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = dns.cn_advance(np.ones(g.shape), np.ones(g.shape), 1.0, 0.5, 1.0, 0.01, 1, g.rho_or_ones(),
                             _axes(g), bc.TB, bc.VB, 1e-2, 1.0, 1.0)
    assert not np.all(np.isfinite(ref))
    with pytest.raises(AMGError, match="zero pivot"):
        s.var_advance_CN_PL(z.clone(), z, 1.0, 0.5, 1.0, 0.01, 1, bc, 1e-2)
    s.close()


def test_substep_tail_on_device(cuda):
    """CN for u, v, w, then Psource, mg_solver and Pgrad with ghosts
    (dns/NSsolver.F90:508-535, RK3CN branch): the projected field is
    divergence-free and its ghosts satisfy the flow BCs."""
    torch = cuda
    from anelastic_solve_amd import FlowBC, PressureSolver, default_config, workloads
    from oracle import dns
    g, U, V, W, R = workloads.shear_layer((64, 32, 32))
    # free-slip walls: u, v Neumann; w is the wall-normal face value itself at
    # the bottom (index sz-1, Dirichlet 0) and at the top (index ez, type 7)
    bc = FlowBC.make((True, True, False), {(3, 1, 1): (2, 0.0), (3, 2, 1): (2, 0.0),
                                            (3, 1, 2): (2, 0.0), (3, 2, 2): (2, 0.0),
                                            (3, 1, 3): (1, 0.0), (3, 2, 3): (7, 0.0)})
    s = PressureSolver(g, default_config(precision="mixed", tol=1e-12))
    rng = np.random.Generator(np.random.PCG64(9))
    dev = {n: torch.from_numpy(a).cuda() for n, a in (("u", U), ("v", V), ("w", W))}
    for n, var in (("u", 1), ("v", 2), ("w", 3)):
        tend = torch.from_numpy(1e-3 * rng.standard_normal(g.shape)).cuda()
        s.var_advance_CN_PL(tend, dev[n], 1.0, 0.5, 1.0, 0.01, var, bc, 1e-3)
    src = torch.zeros(g.shape, dtype=torch.float64, device="cuda")
    dt = 0.01
    s.psource(dev["u"], dev["v"], dev["w"], src, dt)
    div0 = s.divergence(dev["u"], dev["v"], dev["w"])[1]
    p = torch.zeros(g.shape, dtype=torch.float64, device="cuda")
    s.mg_solver(p, src)
    assert s.stats.converged == 1
    s.pgrad_ghost(p, dev["u"], dev["v"], dev["w"], dt, bc)
    div1 = s.divergence(dev["u"], dev["v"], dev["w"])[1]
    assert div1 <= 2e-10 * div0, (div0, div1)
    axes = _axes(g)
    for n, t in (("u", "utemp"), ("v", "vtemp"), ("w", "wtemp")):
        a = dev[n].cpu().numpy()
        assert np.array_equal(dns.ghost(a.copy(), t, bc.TB, bc.VB, axes), a), n
    s.close()
