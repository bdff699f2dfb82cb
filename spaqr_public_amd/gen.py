"""Synthetic matrices for the spaQR workloads (the reference's matrices came
from the external MatrixGen repository and are not shipped, so they are
regenerated here from their definitions).

Grid ordering follows the reference's tensor coordinates (util.cpp:852-881):
2-D dof id = i*N + j, 3-D dof id = (i*N + j)*N + k, where coordinate 0 is the
slowest index.  h = 1/(N+1), homogeneous Dirichlet boundary.

All matrices are returned as scipy.sparse CSC with sorted row indices.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _offsets(n: int, dim: int):
    """Strides of each coordinate for the tensor ordering."""
    return [n ** (dim - 1 - d) for d in range(dim)]


def advection_diffusion(n: int, dim: int, a: float = 1.0, velocity=None, scheme: str = "upwind") -> sp.csc_matrix:
    """-a Lap(u) + v . grad(u) with central second differences for the
    diffusion.  The convection term is either first-order upwind
    (scheme="upwind": backward difference for a positive velocity component)
    or second-order central (scheme="central": (u_{i+1} - u_{i-1}) / 2h).
    (2*dim+1)-point stencil.

    scheme="central" with n=128, dim=2, a=1, v=(1000,1000) reproduces every
    integer of the reference's published log (README.md:63-83: top-separator
    sizes, nnzR/nnzH/nnzQ, 5 GMRES iterations), i.e. it is the matrix of
    MatrixGen's advdiff_2_128_1_p1_1000.mm up to a global scale."""
    if velocity is None:
        velocity = [0.0] * dim
    v = np.asarray(velocity, dtype=np.float64)
    assert v.shape == (dim,)
    h = 1.0 / (n + 1)
    N = n ** dim
    idx = np.arange(N, dtype=np.int64)
    coords = np.stack(np.unravel_index(idx, (n,) * dim), axis=0)
    strides = _offsets(n, dim)
    rows = [idx]
    cols = [idx]
    if scheme == "upwind":
        diag = np.full(N, 2.0 * dim * a / h**2 + np.abs(v).sum() / h)
    elif scheme == "central":
        diag = np.full(N, 2.0 * dim * a / h**2)
    else:
        raise ValueError(scheme)
    vals = [diag]
    for d in range(dim):
        if scheme == "upwind":
            minus = -a / h**2 - max(v[d], 0.0) / h
            plus = -a / h**2 - max(-v[d], 0.0) / h
        else:
            minus = -a / h**2 - v[d] / (2.0 * h)
            plus = -a / h**2 + v[d] / (2.0 * h)
        m = coords[d] > 0
        rows.append(idx[m]); cols.append(idx[m] - strides[d]); vals.append(np.full(m.sum(), minus))
        p = coords[d] < n - 1
        rows.append(idx[p]); cols.append(idx[p] + strides[d]); vals.append(np.full(p.sum(), plus))
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    A = A.tocsc()
    A.sort_indices()
    return A


def high_contrast_diffusion(n: int, dim: int = 3, block: int = 8, seed: int = 0,
                            log10_range=(-2.0, 2.0)) -> sp.csc_matrix:
    """-div(a grad u) by cell-centred finite volumes.  a is piecewise constant on
    blocks of block^dim cells with log10(a) ~ U(log10_range) drawn from `seed`.
    Interior faces use the harmonic mean of the two cell values, boundary faces
    the cell value."""
    h = 1.0 / (n + 1)
    N = n ** dim
    nb = (n + block - 1) // block
    rng = np.random.default_rng(seed)
    logs = rng.uniform(log10_range[0], log10_range[1], size=(nb,) * dim)
    idx = np.arange(N, dtype=np.int64)
    coords = np.stack(np.unravel_index(idx, (n,) * dim), axis=0)
    acell = 10.0 ** logs[tuple(coords[d] // block for d in range(dim))]
    strides = _offsets(n, dim)
    diag = np.zeros(N)
    rows, cols, vals = [], [], []
    for d in range(dim):
        for sgn in (-1, 1):
            has = coords[d] > 0 if sgn < 0 else coords[d] < n - 1
            nbr = idx + sgn * strides[d]
            af = np.where(has, 2.0 * acell * acell[np.where(has, nbr, idx)] /
                          (acell + acell[np.where(has, nbr, idx)]), acell)
            diag += af / h**2
            rows.append(idx[has]); cols.append(nbr[has]); vals.append(-af[has] / h**2)
    rows.append(idx); cols.append(idx); vals.append(diag)
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    A = A.tocsc()
    A.sort_indices()
    return A


def laplacian_gradient_ls(n: int, mu: float = 1e-2) -> sp.csc_matrix:
    """Least-squares operator A = [L_h; mu * D_h] on an n x n grid: L_h is the
    5-point Laplacian (N = n^2 columns), D_h the forward-difference gradient
    (2N rows, zero Dirichlet values outside the grid).  Shape 3N x N."""
    L = advection_diffusion(n, 2, 1.0, [0.0, 0.0])
    h = 1.0 / (n + 1)
    N = n * n
    idx = np.arange(N, dtype=np.int64)
    i, j = np.unravel_index(idx, (n, n))
    blocks = []
    for d, stride in ((0, n), (1, 1)):
        c = i if d == 0 else j
        r = [idx, idx[c < n - 1]]
        cc = [idx, idx[c < n - 1] + stride]
        v = [np.full(N, -1.0 / h), np.full((c < n - 1).sum(), 1.0 / h)]
        blocks.append(sp.coo_matrix((np.concatenate(v), (np.concatenate(r), np.concatenate(cc))), shape=(N, N)))
    A = sp.vstack([L, mu * blocks[0], mu * blocks[1]]).tocsc()
    A.sort_indices()
    return A


def config_matrix(name: str, **kw) -> sp.csc_matrix:
    """Matrices of the named workloads.
      advdiff2d_128 : 2-D advection-diffusion, 128^2, a=1, v=(1000,1000), upwind
      advdiff2d_128_central : same with central convection (the reference's
                      README run, README.md:48-84)
      advdiff3d_32  : 3-D advection-diffusion, 32^3,  a=1, v=(100,100,100)
      advdiff3d_64  : 3-D advection-diffusion, 64^3,  a=1, v=(100,100,100)
      hicontrast3d_64 : high-contrast 3-D diffusion, 64^3, 8^3 blocks
      ls_lapgrad_256  : [L_h; 1e-2 D_h] on 256^2
    """
    scheme = "central" if name.endswith("_central") else "upwind"
    if name.startswith("advdiff2d_"):
        n = int(name.split("_")[1])
        return advection_diffusion(n, 2, 1.0, kw.get("velocity", [1000.0, 1000.0]), scheme)
    if name.startswith("advdiff3d_"):
        n = int(name.split("_")[1])
        return advection_diffusion(n, 3, 1.0, kw.get("velocity", [100.0, 100.0, 100.0]), scheme)
    if name.startswith("hicontrast3d_"):
        n = int(name.split("_")[1])
        return high_contrast_diffusion(n, 3, 8, kw.get("seed", 0))
    if name.startswith("ls_lapgrad_"):
        n = int(name.split("_")[2])
        return laplacian_gradient_ls(n, kw.get("mu", 1e-2))
    raise KeyError(name)


def write_mm_array(path: str, X) -> None:
    """Matrix Market dense array writer (general, real): the header, the
    shape, then the entries column by column, one per line (the layout
    mmio.hpp:235-320 reads)."""
    X = np.atleast_2d(np.asarray(X, dtype=np.float64))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n")
        f.write(f"{X.shape[0]} {X.shape[1]}\n")
        np.savetxt(f, X.ravel(order="F"), fmt="%.17g")


def write_mm(path: str, A: sp.spmatrix) -> None:
    """Matrix Market coordinate writer (general, real), 1-based indices."""
    A = A.tocoo()
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{A.shape[0]} {A.shape[1]} {A.nnz}\n")
        data = np.column_stack([A.row + 1, A.col + 1, A.data])
        np.savetxt(f, data, fmt=["%d", "%d", "%.17g"])
