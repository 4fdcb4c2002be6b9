"""Seeded synthetic inputs of the benchmark workloads (host numpy, reference layout).

A workload's pressure RHS is the discrete divergence (``Psource``) of a
seeded face-velocity field; the fields are ghosted arrays of shape
``(nz+2, ny+2, nx+2)`` with periodic ghosts wrapped in x and y and zero
wall-normal velocity on the z walls, so the RHS is compatible.
"""
import math

import numpy as np

from .grid import box_grid

# the five workloads (sizes, domain 2*pi x 2*pi x 1, periodic x/y, Neumann z)
UNIFORM_64 = dict(n=(64, 64, 64), beta=0.0, seed=0, spectrum="white")
STRETCHED_256 = dict(n=(256, 256, 256), beta=2.0, seed=1, spectrum="k53")
SHEAR_512 = dict(n=(512, 256, 256), beta=2.0, seed=2)
BATCH_SIZES = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 48), (128, 128, 128)]
SERVE_SIZES = [(64, 64, 64), (128, 128, 128), (192, 192, 96), (256, 256, 256), (384, 384, 192)]


def finish_velocity(U, V, W):
    """Periodic x/y ghost wrap; zero wall-normal velocity on both z walls."""
    for F in (U, V, W):
        F[:, :, 0] = F[:, :, -2]
        F[:, :, -1] = F[:, :, 1]
        F[:, 0, :] = F[:, -2, :]
        F[:, -1, :] = F[:, 1, :]
    nz = U.shape[0] - 2
    W[0, :, :] = 0.0
    W[nz, :, :] = 0.0
    W[nz + 1, :, :] = 0.0
    return U, V, W


def velocity_white(grid, seed):
    """i.i.d. N(0,1) face velocities (numpy PCG64 with ``seed``)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    U, V, W = (rng.standard_normal(grid.shape) for _ in range(3))
    return finish_velocity(U, V, W)


def _spectral_planes(rng, nz, ny, nx, slope):
    """Per-plane random-phase fields with |u_hat(k)|^2 k ~ k^slope in x/y."""
    ky = np.fft.fftfreq(ny, 1.0 / ny)[:, None]
    kx = np.fft.rfftfreq(nx, 1.0 / nx)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    amp = np.zeros_like(k)
    nz_k = k > 0
    # 2-D shell: E(k) ~ |u_hat|^2 * 2 pi k  =>  |u_hat| ~ k^((slope-1)/2)
    amp[nz_k] = k[nz_k] ** ((slope - 1.0) / 2.0)
    out = np.empty((nz, ny, nx))
    for q in range(nz):
        ph = rng.uniform(0.0, 2.0 * math.pi, size=k.shape)
        out[q] = np.fft.irfft2(amp * np.exp(1j * ph), s=(ny, nx))
    out /= np.sqrt(np.mean(out * out))
    return out


def velocity_spectral(grid, seed, slope=-5.0 / 3.0):
    """Random-phase velocity with a ``k^slope`` spectrum in x/y, independent per z-plane."""
    rng = np.random.Generator(np.random.PCG64(seed))
    F = []
    for _ in range(3):
        a = np.zeros(grid.shape)
        a[1:-1, 1:-1, 1:-1] = _spectral_planes(rng, grid.nz, grid.ny, grid.nx, slope)
        F.append(a)
    return finish_velocity(*F)


def make(workload, n=None):
    """(grid, U, V, W) of a named workload (optionally at another size)."""
    nx, ny, nz = n if n is not None else workload["n"]
    g = box_grid(nx, ny, nz, beta=workload.get("beta", 0.0))
    if workload.get("spectrum", "white") == "white":
        U, V, W = velocity_white(g, workload["seed"])
    else:
        U, V, W = velocity_spectral(g, workload["seed"])
    return g, U, V, W


def batch_specs(count=96, seed=3):
    """Sizes and stretching factors of the uneven batch (seeded)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    idx = rng.integers(0, len(BATCH_SIZES), size=count)
    beta = rng.uniform(1.0, 2.5, size=count)
    return [(BATCH_SIZES[i], float(b)) for i, b in zip(idx, beta)]


def psource_host(U, V, W, grid, dt=1.0):
    """Host copy of ``Psource`` (``dns/NSsolver.F90:2204-2219``) used to build inputs."""
    nz, ny, nx = grid.nz, grid.ny, grid.nx
    r = grid.rho_or_ones()
    rk = r[1:nz + 1][:, None, None]
    rp = r[2:nz + 2][:, None, None]
    rm = r[0:nz][:, None, None]
    rdxc = (1.0 / grid.ax.dxc[1:nx + 1])[None, None, :]
    rdyc = (1.0 / grid.ay.dxc[1:ny + 1])[None, :, None]
    rdzc = (1.0 / grid.az.dxc[1:nz + 1])[:, None, None]
    src = np.zeros_like(U)
    src[1:-1, 1:-1, 1:-1] = (
        (rk * U[1:-1, 1:-1, 1:-1] - rk * U[1:-1, 1:-1, 0:-2]) * rdxc
        + (rk * V[1:-1, 1:-1, 1:-1] - rk * V[1:-1, 0:-2, 1:-1]) * rdyc
        + (0.5 * (rp + rk) * W[1:-1, 1:-1, 1:-1] - 0.5 * (rk + rm) * W[0:-2, 1:-1, 1:-1]) * rdzc
    ) * (1.0 / dt)
    return src
