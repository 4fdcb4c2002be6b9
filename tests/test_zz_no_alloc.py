"""The allocation-free divergence report (``amg_divergence_device_noalloc``)
against the allocating one (``amg_divergence_device``), through the C ABI.

Both run the same kernels; the no-alloc entry keeps its block partials in
a static device array, so its results must equal the allocating entry's
bit for bit.  Two calls in a row, the second on a smaller grid (the array
keeps the first call's contents), then a grid beyond the array's limit,
which must be refused before any kernel runs.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(torch, g, seed):
    from anelastic_solve_amd import workloads
    U, V, W = workloads.velocity_white(g, seed)
    return [torch.from_numpy(a).cuda() for a in (U, V, W)]


def _report(s, fn, U, V, W, torch):
    div = torch.zeros(s.grid.shape, dtype=torch.float64, device="cuda")
    avg, mx = C.c_double(), C.c_double()
    loc = (C.c_int * 3)()
    torch.cuda.synchronize()
    rc = fn(s._h, C.c_void_p(U.data_ptr()), C.c_void_p(V.data_ptr()), C.c_void_p(W.data_ptr()),
            C.c_void_p(div.data_ptr()), C.byref(avg), C.byref(mx), loc)
    return rc, avg.value, mx.value, tuple(loc), div


def test_noalloc_divergence_equals_allocating(cuda):
    torch = cuda
    from anelastic_solve_amd import PressureSolver, box_grid, default_config
    from anelastic_solve_amd._capi import AMG_OK
    lib = None
    for n, seed in (((64, 48, 40), 1), ((16, 8, 8), 2)):
        g = box_grid(*n, beta=1.5)
        s = PressureSolver(g, default_config())
        lib = s._lib
        U, V, W = _fields(torch, g, seed)
        print(f"noalloc divergence: grid {n} seed {seed}", flush=True)
        rc1, a1, m1, l1, d1 = _report(s, lib.amg_divergence_device_noalloc, U, V, W, torch)
        rc0, a0, m0, l0, d0 = _report(s, lib.amg_divergence_device, U, V, W, torch)
        assert rc0 == AMG_OK and rc1 == AMG_OK, (rc0, rc1)
        assert (a1, m1, l1) == (a0, m0, l0)
        assert torch.equal(d1, d0)
        # and the allocating entry agrees with the field it wrote
        dd = d0.cpu().numpy()[1:-1, 1:-1, 1:-1]
        assert m0 == np.abs(dd).max()
        s.close()


def test_noalloc_divergence_refuses_oversize(cuda):
    torch = cuda
    from anelastic_solve_amd import PressureSolver, box_grid, default_config
    from anelastic_solve_amd._capi import AMG_EBADSIZE
    n = (512, 512, 256)          # 67.1 M cells > 33 554 432
    g = box_grid(*n, beta=0.0)
    s = PressureSolver(g, default_config(precision="fp64"))
    z = torch.zeros(g.shape, dtype=torch.float64, device="cuda")
    print(f"noalloc divergence: oversize grid {n}", flush=True)
    rc, *_ = _report(s, s._lib.amg_divergence_device_noalloc, z, z, z, torch)
    assert rc == AMG_EBADSIZE
    s.close()
