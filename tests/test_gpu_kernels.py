"""Per-kernel parity of the HIP path with the oracle (through the C ABI).

Each test loads seeded fields into one level of a solver handle
(``amg_level_field``), runs exactly one kernel (``amg_run_kernel``), reads
the result back and compares it with the oracle's restatement of the same
step on the same inputs.  FP64 kernels agree to rounding (1e-12 relative,
the stencil sums are ordered identically but FMA contraction differs);
FP32 kernels to 2e-5 (1e-4 for the z-line Thomas solve, whose forward
elimination accumulates FP32 rounding along 8-32 cells).
"""
import numpy as np
import pytest

from conftest import oracle_levels, relerr

pytestmark = pytest.mark.gpu

CASES = [
    # (n, beta, tbc, nlevels)
    ((16, 16, 16), 2.0, (0, 1, 0, 1, 0, 0), None),          # periodic x/y, Neumann z, stretched
    ((16, 8, 12), 0.0, (0, 1, 0, 1, 0, 0), None),           # uneven sizes
    ((8, 8, 8), 1.0, (2, 2, 1, 1, 2, 2), None),             # Dirichlet x/y faces, Neumann z
    ((16, 16, 8), 0.0, (0, 0, 0, 0, 0, 0), None),           # fully periodic
    # z-columns longer than one 16-cell segment (partitioned line solve)
    ((8, 8, 64), 2.0, (0, 1, 0, 1, 0, 0), None),            # 4 segments
    ((16, 16, 36), 1.0, (2, 2, 1, 1, 2, 2), None),          # 3 segments, tail of 4
    ((8, 8, 48), 0.0, (0, 0, 0, 0, 0, 0), None),            # periodic z across segments
    ((8, 8, 17), 1.5, (0, 1, 0, 1, 0, 0), (2, 2, 1)),       # tail of 1 cell, z not coarsened
]


def _solver(n, beta, tbc, smoother, precision, rho=False, nlevels=None):
    from anelastic_solve_amd import PressureSolver, box_grid, default_config
    nx, ny, nz = n
    r = None
    if rho:
        z = np.linspace(-0.05, 1.05, nz + 2)
        r = 1.0 - 0.1 * z
    g = box_grid(nx, ny, nz, beta=beta, rho=r)
    vbc = [0.0 if t else -1.0 for t in tbc]
    kw = {} if nlevels is None else {"nlevels": nlevels}
    cfg = default_config(smoother=smoother, precision=precision, tbc=tbc, vbc=vbc, **kw)
    s = PressureSolver(g, cfg)
    levs, fac = oracle_levels(g, tbc, s.levels, nlevels)
    return s, g, levs, fac


def _tol(precision, smoother):
    if precision == "fp64":
        return 1e-12
    return 1e-4 if smoother == "LJ" else 2e-5


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("smoother", ["LJ", "RBPGS"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_half_sweeps(cuda, case, smoother, precision):
    from oracle import vcycle as VC
    n, beta, tbc, nlev = CASES[case]
    s, g, levs, fac = _solver(n, beta, tbc, smoother, precision, rho=(case in (0, 4)), nlevels=nlev)
    dt = np.float64 if precision == "fp64" else np.float32
    rng = np.random.default_rng(100 + case)
    L = len(levs)
    for lvl in range(L - 1):
        lev = levs[lvl]
        nx, ny, nz = lev["n"]
        u = rng.standard_normal((nz, ny, nx)).astype(dt)
        f = rng.standard_normal((nz, ny, nx)).astype(dt)
        hs = VC.zline_half_sweep if smoother == "LJ" else VC.rb_half_sweep
        runs = [(0, 0), (4, 1), (6, 0)]
        if smoother == "LJ" and nz <= (16 if precision == "fp64" else 32):
            runs += [(7, 0), (8, 1)]      # the coarse tail's short-column line solve
        for which, parity in runs:
            s.level_field(lvl, 0, u)
            s.level_field(lvl, 1, f)
            s.run_kernel(which, lvl, 1)
            got = s.level_field(lvl, 0)
            ref = (np.zeros_like(u) if which == 6 else u.copy())
            hs(ref, f, lev, parity, dt)
            if which == 6:
                # a zero-guess half-sweep leaves the other colour untouched: compare updated colour only
                m = VC.colour_mask(u.shape, parity, smoother == "LJ")
                assert relerr(got[m], ref[m]) < _tol(precision, smoother), (lvl, which)
            else:
                assert relerr(got, ref) < _tol(precision, smoother), (lvl, which)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("smoother", ["LJ", "RBPGS"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_prolong_skip_then_post_sweep(cuda, case, smoother, precision):
    """The V-cycle's prolongation leaves the cells that the level's first
    post-smoothing half-sweep (colour 0) rewrites without reading them: the
    z-line sweep's (i+j) even cells, except rows 0 and nz-1 with periodic z
    (its lagged wrap reads them); the point sweep's (i+j+k) even cells.
    Followed by a red and a black half-sweep, it gives the bits of the full
    prolongation followed by the same sweeps; on its own it corrects the
    other cells as the full one does and leaves the skipped ones.  Only
    2:2:2 transfers skip; the others correct every cell."""
    n, beta, tbc, nlev = CASES[case]
    s, g, levs, fac = _solver(n, beta, tbc, smoother, precision, rho=(case in (0, 4)), nlevels=nlev)
    dt = np.float64 if precision == "fp64" else np.float32
    rng = np.random.default_rng(300 + case)
    pz = tbc[1] == 0 and tbc[3] == 0
    for lvl in range(len(levs) - 1):
        nx, ny, nz = levs[lvl]["n"]
        cx, cy, cz = levs[lvl + 1]["n"]
        u = rng.standard_normal((nz, ny, nx)).astype(dt)
        f = rng.standard_normal((nz, ny, nx)).astype(dt)
        uc = rng.standard_normal((cz, cy, cx)).astype(dt)
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        if smoother == "LJ":
            skipped = ((i + j) & 1) == 0
            if pz:
                skipped &= (k != 0) & (k != nz - 1)
        else:
            skipped = ((i + j + k) & 1) == 0
        if (nx, ny, nz) != (2 * cx, 2 * cy, 2 * cz):
            skipped = np.zeros_like(skipped)
        out = {}
        for which in (2, 15):
            s.level_field(lvl, 0, u)
            s.level_field(lvl, 1, f)
            s.level_field(lvl + 1, 0, uc)
            s.run_kernel(which, lvl, 1)
            alone = s.level_field(lvl, 0)
            s.run_kernel(0, lvl, 1)
            s.run_kernel(4, lvl, 1)
            out[which] = (alone, s.level_field(lvl, 0))
        full, skip = out[2][0], out[15][0]
        assert np.array_equal(skip[~skipped], full[~skipped]), lvl
        assert np.array_equal(skip[skipped], u[skipped].astype(np.float64)), lvl
        assert skipped.any() != np.array_equal(skip, full), lvl
        assert np.array_equal(out[15][1], out[2][1]), lvl
    s.close()


@pytest.mark.parametrize("n,beta,tbc,nlevels", [
    # slabs of 8 planes: the production grid's finest level (Dirichlet z walls)
    ((256, 64, 256), 0.0, (0, 2, 0, 2, 0, 0), None),
    # slabs of 4 planes, the last one of 2, periodic in every direction
    ((512, 128, 42), 1.0, (0, 0, 0, 0, 0, 0), (7, 7, 1)),
])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_point_sweep_march_large_level(cuda, precision, n, beta, tbc, nlevels):
    """The point half-sweep on a level large enough to march in z (a thread
    per chequer position over a slab of planes, the z-neighbour above plane
    k reused below plane k+2) against the oracle's point sweep, both
    colours, from a guess and from zero."""
    from oracle import vcycle as VC
    s, g, levs, fac = _solver(n, beta, tbc, "RBPGS", precision, nlevels=nlevels)
    dt = np.float64 if precision == "fp64" else np.float32
    rng = np.random.default_rng(77)
    lev = levs[0]
    nx, ny, nz = lev["n"]
    u = rng.standard_normal((nz, ny, nx)).astype(dt)
    f = rng.standard_normal((nz, ny, nx)).astype(dt)
    for which, parity in ((0, 0), (4, 1), (6, 0)):
        s.level_field(0, 0, u)
        s.level_field(0, 1, f)
        s.run_kernel(which, 0, 1)
        got = s.level_field(0, 0)
        ref = np.zeros_like(u) if which == 6 else u.copy()
        VC.rb_half_sweep(ref, f, lev, parity, dt)
        if which == 6:
            m = VC.colour_mask(u.shape, parity, False)
            assert relerr(got[m], ref[m]) < _tol(precision, "RBPGS"), which
        else:
            assert relerr(got, ref) < _tol(precision, "RBPGS"), which
    s.close()


@pytest.mark.parametrize("tbc", [(0, 1, 0, 1, 0, 0), (2, 2, 1, 1, 2, 2), (0, 0, 0, 0, 0, 0)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_restriction_march_large_level(cuda, precision, tbc):
    """A factor-2 restriction onto a coarse level of 2^18 cells or more runs
    the z-marching kernel (a thread per coarse column and slab of coarse
    planes; here 128x128x35 or x36 coarse cells, slabs of 2, the last one
    partial with 35) instead of the per-cell one.  It agrees with the
    oracle's residual + restriction to the same bar, with periodic, Neumann
    and Dirichlet faces, periodic z included."""
    from oracle import vcycle as VC
    nz = 72 if tbc[4] == 0 else 70          # periodic z needs even sizes on every level
    s, g, levs, fac = _solver((256, 256, nz), 1.5, tbc, "LJ", precision, rho=True,
                              nlevels=(7, 7, 2))
    assert tuple(fac[0]) == (2, 2, 2) and tuple(levs[1]["n"]) == (128, 128, nz // 2)
    dt = np.float64 if precision == "fp64" else np.float32
    tol = 1e-12 if precision == "fp64" else 2e-5
    rng = np.random.default_rng(300)
    u = rng.standard_normal((nz, 256, 256)).astype(dt)
    f = rng.standard_normal((nz, 256, 256)).astype(dt)
    s.level_field(0, 0, u)
    s.level_field(0, 1, f)
    s.run_kernel(1, 0, 1)
    got = s.level_field(1, 1)
    ref = VC.restrict(VC.residual(u, f, levs[0], dt), levs[0], levs[1], fac[0], dt)
    assert relerr(got, ref) < tol
    s.close()


@pytest.mark.parametrize("smoother,tbc", [
    ("LJ", (0, 1, 0, 1, 0, 0)), ("LJ", (0, 2, 0, 2, 0, 0)), ("RBPGS", (0, 1, 0, 1, 0, 0)),
    ("LJ", (0, 0, 0, 0, 0, 0))])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_restriction_skips_solved_residuals(cuda, precision, smoother, tbc):
    """The V-cycle's marching restriction takes as zero the residuals of the
    cells that the level's last pre-smoothing half-sweep (colour 1) has just
    solved: (i+j) odd for the z-line sweep, (i+j+k) odd for the point sweep;
    with periodic z (the z-line sweep's lagged wrap) it forms every residual.
    On random fields it is the oracle's restriction of the residual with
    those cells zeroed; after a red and a black half-sweep it is the full
    restriction to rounding."""
    from oracle import vcycle as VC
    nz = 72
    s, g, levs, fac = _solver((256, 256, nz), 1.5, tbc, smoother, precision, rho=True,
                              nlevels=(7, 7, 2))
    assert tuple(fac[0]) == (2, 2, 2) and tuple(levs[1]["n"]) == (128, 128, nz // 2)
    dt = np.float64 if precision == "fp64" else np.float32
    tol = 1e-12 if precision == "fp64" else 2e-5
    pz = tbc[1] == 0 and tbc[3] == 0
    rng = np.random.default_rng(310)
    u = rng.standard_normal((nz, 256, 256)).astype(dt)
    f = rng.standard_normal((nz, 256, 256)).astype(dt)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(256), np.arange(256), indexing="ij")
    solved = (((i + j) if smoother == "LJ" else (i + j + k)) & 1) == 1
    if smoother == "LJ" and pz:
        solved[:] = False

    def restriction(which, sweep):
        s.level_field(0, 0, u)
        s.level_field(0, 1, f)
        if sweep:
            s.run_kernel(0, 0, 1)
            s.run_kernel(4, 0, 1)
        s.run_kernel(which, 0, 1)
        return s.level_field(1, 1)

    r = VC.residual(u, f, levs[0], dt)
    r[solved] = 0
    ref = VC.restrict(r, levs[0], levs[1], fac[0], dt)
    got = restriction(16, False)
    assert relerr(got, ref) < tol
    if not solved.any():
        assert np.array_equal(got, restriction(1, False))
    full, skip = restriction(1, True), restriction(16, True)
    err = relerr(skip, full)
    print(f"{precision} {smoother} tbc {tbc}: after the sweeps {err:.2e}", flush=True)
    assert err < (1e-12 if precision == "fp64" else 1e-5), err
    s.close()


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_transfers_and_coarse_solve(cuda, case, precision):
    from oracle import vcycle as VC
    from oracle.grid import is_singular
    n, beta, tbc, nlev = CASES[case]
    s, g, levs, fac = _solver(n, beta, tbc, "LJ", precision, rho=(case in (0, 4)), nlevels=nlev)
    dt = np.float64 if precision == "fp64" else np.float32
    tol = 1e-12 if precision == "fp64" else 2e-5
    rng = np.random.default_rng(200 + case)
    L = len(levs)
    for lvl in range(L - 1):
        fine, coarse = levs[lvl], levs[lvl + 1]
        nx, ny, nz = fine["n"]
        cx, cy, cz = coarse["n"]
        u = rng.standard_normal((nz, ny, nx)).astype(dt)
        f = rng.standard_normal((nz, ny, nx)).astype(dt)
        uc = rng.standard_normal((cz, cy, cx)).astype(dt)
        # residual + restriction
        s.level_field(lvl, 0, u)
        s.level_field(lvl, 1, f)
        s.run_kernel(1, lvl, 1)
        got = s.level_field(lvl + 1, 1)
        ref = VC.restrict(VC.residual(u, f, fine, dt), fine, coarse, fac[lvl], dt)
        assert relerr(got, ref) < tol, ("restrict", lvl)
        # prolongation-add
        s.level_field(lvl, 0, u)
        s.level_field(lvl + 1, 0, uc)
        s.run_kernel(2, lvl, 1)
        got = s.level_field(lvl, 0)
        ref = VC.prolong_add(u.copy(), uc, fine, coarse, fac[lvl], dt)
        assert relerr(got, ref) < tol, ("prolong", lvl)
    # coarsest-level solve
    cl = levs[-1]
    nx, ny, nz = cl["n"]
    fc = rng.standard_normal((nz, ny, nx)).astype(dt)
    s.level_field(L - 1, 1, fc)
    s.run_kernel(5, L - 1, 1)
    got = s.level_field(L - 1, 0)
    Minv = VC.coarse_operator_inverse(cl, is_singular(tbc))
    ref = (Minv @ fc.astype(np.float64).ravel()).reshape(fc.shape)
    assert relerr(got, ref) < (1e-11 if precision == "fp64" else 2e-6)


@pytest.mark.parametrize("tbc", [(0, 1, 0, 1, 0, 0), (0, 0, 0, 0, 0, 0)])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_zline_ragged_full_segments(cuda, precision, tbc):
    """A level with more column groups than the device has CUs (128 x 128
    columns per colour: the lean form in FP32) and ragged columns of 40
    cells (two full 16-cell segments and a tail of 8), Neumann and periodic
    in z, both colours, from a guess and from zero, against the oracle's
    Thomas solve per column."""
    from oracle import vcycle as VC
    dt = np.float64 if precision == "fp64" else np.float32
    s, g, levs, fac = _solver((256, 128, 40), 1.5, tbc, "LJ", precision, nlevels=(6, 5, 3))
    lev = levs[0]
    rng = np.random.default_rng(11)
    u = rng.standard_normal((40, 128, 256)).astype(dt)
    f = rng.standard_normal((40, 128, 256)).astype(dt)
    for which, parity in ((0, 0), (4, 1), (6, 0)):
        s.level_field(0, 0, u)
        s.level_field(0, 1, f)
        s.run_kernel(which, 0, 1)
        got = s.level_field(0, 0)
        ref = np.zeros_like(u) if which == 6 else u.copy()
        VC.zline_half_sweep(ref, f, lev, parity, dt)
        if which == 6:
            m = VC.colour_mask(u.shape, parity, True)
            got, ref = got[m], ref[m]
        err = relerr(got, ref)
        print(f"{precision} tbc {tbc} kernel {which} parity {parity}: err {err:.2e}", flush=True)
        assert err < _tol(precision, "LJ"), (which, parity, err)
    s.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_zline_lean_form_full_columns(cuda, precision):
    """The benchmarked z-line sweep: the finest level of the 256^3 stretched
    workload (256-cell columns in 16 segments, a launch of 512 blocks, more
    than the device's CUs, so it takes the lean form) against the oracle's
    Thomas solve per column, both colours.  In FP32 the short form on the
    same input gives the same bits; in either precision the lean form forced
    onto a small launch gives the bits of that launch's own choice."""
    from oracle import vcycle as VC
    dt = np.float64 if precision == "fp64" else np.float32
    s, g, levs, fac = _solver((256, 256, 256), 2.0, (0, 1, 0, 1, 0, 0), "LJ", precision)
    lev = levs[0]
    rng = np.random.default_rng(7)
    u = rng.standard_normal((256, 256, 256)).astype(dt)
    f = rng.standard_normal((256, 256, 256)).astype(dt)
    for which, parity in ((0, 0), (4, 1)):
        s.level_field(0, 0, u)
        s.level_field(0, 1, f)
        s.run_kernel(which, 0, 1)
        got = s.level_field(0, 0)
        ref = u.copy()
        VC.zline_half_sweep(ref, f, lev, parity, dt)
        err = relerr(got, ref)
        print(f"256^3 {precision} z-line parity {parity}: err {err:.2e}", flush=True)
        assert err < _tol(precision, "LJ"), (parity, err)
        if precision == "mixed":
            s.level_field(0, 0, u)
            s.run_kernel(10 + parity, 0, 1)
            assert np.array_equal(s.level_field(0, 0), got), parity
    s.close()
    # a small launch (16x16 columns of 64 cells: 2 blocks) takes the short
    # form in FP32; the lean form forced onto it gives the same bits
    s, g, levs, fac = _solver((16, 16, 64), 1.0, (0, 1, 0, 1, 0, 0), "LJ", precision)
    u = rng.standard_normal((64, 16, 16)).astype(dt)
    f = rng.standard_normal((64, 16, 16)).astype(dt)
    for parity in (0, 1):
        s.level_field(0, 0, u)
        s.level_field(0, 1, f)
        s.run_kernel(4 * parity, 0, 1)
        auto = s.level_field(0, 0)
        s.level_field(0, 0, u)
        s.run_kernel(12 + parity, 0, 1)
        assert np.array_equal(s.level_field(0, 0), auto), parity
    s.close()


@pytest.mark.parametrize("n,tbc", [((128, 128, 128), (0, 1, 0, 1, 0, 0)),
                                   ((48, 40, 40), (0, 0, 0, 0, 0, 0)),
                                   ((32, 32, 256), (0, 1, 0, 1, 0, 0)),
                                   ((48, 32, 80), (0, 0, 0, 0, 0, 0))])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_zline_half_wave_groups(cuda, precision, n, tbc):
    """A launch of at most half as many column groups as the device has CUs
    runs 32 columns per workgroup, two segments per wave: the 128^3 level of
    the benchmarked hierarchy (8 full segments per column), a ragged level
    (40-cell columns: 3 segments, rounded up to 4; 960 columns per colour, so
    the last group is partial), the longest columns (256 cells, 16 segments,
    8 waves per workgroup) and an odd count of full segments (80 cells: 5,
    rounded up to 6), periodic z included.  Against the oracle's
    Thomas solve per column, and bitwise against the forced forms, which
    keep 64 columns per workgroup."""
    from oracle import vcycle as VC
    dt = np.float64 if precision == "fp64" else np.float32
    nx, ny, nz = n
    s, g, levs, fac = _solver(n, 1.5, tbc, "LJ", precision)
    lev = levs[0]
    rng = np.random.default_rng(17)
    u = rng.standard_normal((nz, ny, nx)).astype(dt)
    f = rng.standard_normal((nz, ny, nx)).astype(dt)
    for which, parity in ((0, 0), (4, 1), (6, 0)):
        s.level_field(0, 0, u)
        s.level_field(0, 1, f)
        s.run_kernel(which, 0, 1)
        got = s.level_field(0, 0)
        ref = np.zeros_like(u) if which == 6 else u.copy()
        VC.zline_half_sweep(ref, f, lev, parity, dt)
        gm, rm = got, ref
        if which == 6:
            m = VC.colour_mask(u.shape, parity, True)
            gm, rm = got[m], ref[m]
        err = relerr(gm, rm)
        print(f"{n} {precision} kernel {which} parity {parity}: err {err:.2e}", flush=True)
        assert err < _tol(precision, "LJ"), (which, parity, err)
        if which == 6:
            continue
        forced = (12 + parity,) if precision == "fp64" else (10 + parity, 12 + parity)
        for w in forced:
            s.level_field(0, 0, u)
            s.run_kernel(w, 0, 1)
            assert np.array_equal(s.level_field(0, 0), got), (w, parity)
    s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n,beta,tbc,nlevels,direct", [
    ((64, 64, 64), 2.0, (0, 1, 0, 1, 0, 0), None, ((32, 32, 32),) * 2),     # singular, stretched z
    ((128, 64, 64), 1.0, (2, 2, 2, 2, 2, 2), None, ((64, 32, 32),) * 2),    # Dirichlet x, y and z
    ((96, 96, 48), 1.5, (1, 1, 1, 1, 1, 1), None, ((48, 48, 24),) * 2),     # Neumann everywhere
    ((256, 64, 256), 0.0, (0, 2, 0, 2, 0, 0), (4, 3, 8), ((64, 16, 64),) * 2),   # production levels
    ((256, 256, 256), 2.0, (0, 1, 0, 1, 0, 0), None, ((64, 64, 64),) * 2),      # the benchmarked levels
])
def test_direct_level_solve(cuda, precision, n, beta, tbc, nlevels, direct):
    """The separable direct solve of the coarse level (transforms with the
    eigenvectors of the x, y and z operators on the matrix cores, scaling by
    the inverse eigenvalues, transforms back) against the oracle's
    restatement (numpy eigh in x and y, Thomas along z, pinned against a
    sparse LU) on a random right-hand side.  An FP64 V-cycle's level is
    solved in FP64 arithmetic, to 1e-11; a mixed solve's FP32 V-cycle's in
    FP32 arithmetic, to 2e-5."""
    from oracle import ops, vcycle as VC
    from oracle.grid import is_singular
    s, g, levs, fac = _solver(n, beta, tbc, "LJ", precision, rho=True, nlevels=nlevels)
    ld = s.direct_level
    want = direct[0] if precision == "fp64" else direct[1]
    assert ld >= 1 and s.levels[ld] == want, (ld, s.levels)
    lev = levs[ld]
    rng = np.random.default_rng(21)
    f = rng.standard_normal(tuple(reversed(lev["n"])))
    sing = is_singular(tbc)
    if sing:
        f = ops.compat_project(f, lev)
    s.level_field(ld, 1, f)
    fd = s.level_field(ld, 1)                 # as stored (FP32 in mixed)
    s.run_kernel(14, ld, 1)
    got = s.level_field(ld, 0)
    ref = VC.separable_solve(lev, fd, sing)
    err = relerr(got, ref)
    print(f"direct {direct} {precision}: err {err:.2e}", flush=True)
    assert err <= (1e-11 if precision == "fp64" else 2e-5), err
    s.close()
