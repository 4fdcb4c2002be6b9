"""Test configuration: the ``gpu`` marker and shared builders.

``-m "not gpu"`` tests run on the CPU (oracle pinning, host logic, library
load and exports); ``-m gpu`` tests call the HIP path through the C ABI
and compare it with the oracle.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def cuda():
    """torch with a visible GPU; a gpu test without one fails loudly."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test collected but no HIP device is visible")
    return torch


def oracle_levels(grid, tbc, solver_levels=None, nlevels=None):
    """Oracle hierarchy built from the same metrics as the product grid."""
    from oracle import grid as G
    ax = G.axis_from_faces(grid.ax.xe[:grid.nx + 1])
    ay = G.axis_from_faces(grid.ay.xe[:grid.ny + 1])
    az = G.axis_from_faces(grid.az.xe[:grid.nz + 1])
    if nlevels is None:
        fac = G.auto_factors(grid.nx, grid.ny, grid.nz)
    else:
        fac = G.reference_factors(nlevels)
    levs = G.hierarchy(ax, ay, az, grid.rho_or_ones(), list(tbc), fac)
    if solver_levels is not None:
        assert [l["n"] for l in levs] == [tuple(s) for s in solver_levels]
    return levs, fac


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / s)
