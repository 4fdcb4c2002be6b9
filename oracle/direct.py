"""Exact FP64 direct solves of the pressure operator (oracle).

Two independent solvers, used to pin each other and the V-cycle:

* ``solve_fft``    -- FFT in the periodic, uniform x and y directions
  diagonalises the horizontal part (eigenvalues ``-4/dx^2 sin^2(pi m/n)``),
  leaving one tridiagonal system in z per horizontal mode (Thomas).
* ``solve_sparse`` -- scipy sparse LU of the assembled 7-point matrix, any
  boundary conditions, small grids.

A singular operator (no Dirichlet face) is solved through the bordered
system ``[[A, 1], [V^T, 0]]`` (``V`` = cell volumes), whose solution
satisfies ``A x = b - lambda 1`` with ``V.x = 0``; the result is then put on
the null-space convention of ``ops.null_project``.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .grid import is_singular
from .ops import null_project, volumes


def assemble(lev):
    """Sparse matrix of ``ops.apply_A`` on one level (interior cells, x fastest)."""
    nx, ny, nz = lev["n"]
    N = nx * ny * nz
    wxm, wxp, dx = lev["cx"][:3]
    wym, wyp, dy = lev["cy"][:3]
    wzm, wzp, dz = lev["cz"][:3]
    px, py, pz = lev["per"]
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    k, j, i = k.ravel(), j.ravel(), i.ravel()
    row = (k * ny + j) * nx + i
    rows, cols, vals = [row], [row], [dz[k] + dy[j] + dx[i]]

    def add(w, ii, jj, kk, valid):
        m = valid & (w != 0.0)
        rows.append(row[m])
        cols.append(((kk[m] % nz) * ny + (jj[m] % ny)) * nx + (ii[m] % nx))
        vals.append(w[m])

    tv = np.ones_like(i, dtype=bool)
    add(wxm[i], i - 1, j, k, tv if px else i > 0)
    add(wxp[i], i + 1, j, k, tv if px else i < nx - 1)
    add(wym[j], i, j - 1, k, tv if py else j > 0)
    add(wyp[j], i, j + 1, k, tv if py else j < ny - 1)
    add(wzm[k], i, j, k - 1, tv if pz else k > 0)
    add(wzp[k], i, j, k + 1, tv if pz else k < nz - 1)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(N, N))


def solve_sparse(lev, b, singular):
    """Sparse-LU solve of ``A x = b`` (interior arrays); small grids only."""
    nx, ny, nz = lev["n"]
    N = nx * ny * nz
    if N > 1 << 21:
        raise ValueError(f"solve_sparse: {N} cells is beyond a sparse LU on the host")
    A = assemble(lev)
    rhs = np.asarray(b, dtype=np.float64).ravel()
    if singular:
        V = volumes(lev).ravel()
        ones = np.ones((N, 1))
        M = sp.bmat([[A, sp.csr_matrix(ones)], [sp.csr_matrix(V[None, :]), None]], format="csc")
        x = spla.spsolve(M, np.concatenate([rhs, [0.0]]))[:N]
    else:
        x = spla.spsolve(A.tocsc(), rhs)
    x = x.reshape(nz, ny, nx)
    return null_project(x, lev) if singular else x


def _thomas(a, b, c, d):
    """Batched tridiagonal solve; ``a`` sub, ``b`` diag, ``c`` super, leading axis = z."""
    n = b.shape[0]
    cp = np.empty_like(b)
    dp = np.empty_like(d)
    cp[0] = c[0] / b[0]
    dp[0] = d[0] / b[0]
    for k in range(1, n):
        den = b[k] - a[k] * cp[k - 1]
        cp[k] = c[k] / den
        dp[k] = (d[k] - a[k] * dp[k - 1]) / den
    x = np.empty_like(d)
    x[n - 1] = dp[n - 1]
    for k in range(n - 2, -1, -1):
        x[k] = dp[k] - cp[k] * x[k + 1]
    return x


def solve_fft(lev, b, singular):
    """FFT(x,y) + Thomas(z) solve; x and y must be periodic with uniform spacing."""
    nx, ny, nz = lev["n"]
    px, py, pz = lev["per"]
    assert px and py and not pz, "solve_fft needs periodic x, y and walls in z"
    wxm, wxp, dx = lev["cx"][:3]
    wym, wyp, dy = lev["cy"][:3]
    wzm, wzp, dz = lev["cz"][:3]
    # a uniform axis built from face positions carries rounding of a few ulps
    # per cell width (1e-13 relative at 512 cells); the modes use the mean
    for w in (wxm, wxp, dx, wym, wyp, dy):
        assert np.allclose(w, w[0], rtol=1e-11, atol=0.0), "non-uniform horizontal spacing"
    wxm0, wxp0, dx0 = wxm.mean(), wxp.mean(), dx.mean()
    wym0, wyp0, dy0 = wym.mean(), wyp.mean(), dy.mean()
    mx = np.arange(nx)
    my = np.arange(ny)
    lam_x = dx0 + wxm0 * np.exp(-2j * np.pi * mx / nx) + wxp0 * np.exp(2j * np.pi * mx / nx)
    lam_y = dy0 + wym0 * np.exp(-2j * np.pi * my / ny) + wyp0 * np.exp(2j * np.pi * my / ny)
    lam = (lam_y[:, None] + lam_x[None, :]).real
    bh = np.fft.fft2(np.asarray(b, dtype=np.float64), axes=(1, 2))
    a = np.broadcast_to(wzm[:, None, None], bh.shape).astype(np.float64)
    c = np.broadcast_to(wzp[:, None, None], bh.shape).astype(np.float64)
    dd = dz[:, None, None] + lam[None, :, :]
    xh = np.empty_like(bh)
    mask = np.ones((ny, nx), dtype=bool)
    if singular:
        mask[0, 0] = False
    xh[:, mask] = _thomas(a[:, mask], dd[:, mask], c[:, mask], bh[:, mask])
    if singular:
        # (0,0) mode: bordered dense solve of the z-problem
        T = np.diag(dz) + np.diag(wzm[1:], -1) + np.diag(wzp[:-1], 1)
        dzc = lev["az"]["dxc"][1:nz + 1]
        M = np.zeros((nz + 1, nz + 1))
        M[:nz, :nz] = T
        M[:nz, nz] = 1.0
        M[nz, :nz] = dzc
        rhs = np.concatenate([bh[:, 0, 0].real, [0.0]])
        xh[:, 0, 0] = np.linalg.solve(M, rhs)[:nz]
    x = np.fft.ifft2(xh, axes=(1, 2)).real
    return null_project(x, lev) if singular else x


def solve(lev, b, tbc):
    """Exact solve with the fastest applicable method."""
    singular = is_singular(tbc)
    px, py, pz = lev["per"]
    nx, ny, nz = lev["n"]
    if px and py and not pz:
        try:
            return solve_fft(lev, b, singular)
        except AssertionError:
            pass
    return solve_sparse(lev, b, singular)
