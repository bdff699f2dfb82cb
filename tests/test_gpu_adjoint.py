"""Adjoint of the spaQR solve: the transposed apply M^{-T}, GMRES on A^T and
the differentiable solve (torch.autograd) against FP64 references.

- M^{-T} is pinned by the adjoint identity <M^{-T} u, v> = <u, M^{-1} v>
  (random u, v) and, on a small tree, by the dense matrices of both applies.
- Gradients of L = ||x - x_t||^2, x = A^{-1} b, are checked against the
  FP64 reference gradient (scipy sparse LU on A and A^T: lambda = A^{-T} 2(x -
  x_t), dL/dA_ij = -lambda_i x_j, dL/db = lambda) to 1e-6 relative, and
  against central finite differences on 32 random nonzeros.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# torch carries its own HIP runtime next to the one libspaqr.so links; let it
# initialise the device at collection time, before any test drives the GPU
# through libspaqr.so in this process.
try:
    import torch
    if torch.cuda.is_available():
        torch.cuda.init()
except ImportError:
    pass


def _tree(name, n, dim, tol, precision=None, order=1.0, scale=1):
    import spaqr_public_amd as sq
    from spaqr_public_amd import gen
    A = gen.config_matrix(name)
    As, _ = sq.colscale(A)
    t = sq.Tree(sq.default_levels(As.shape[1]), 0, sq.FP64 if precision is None else precision)
    t.set_tol(tol)
    t.set_scale(scale)
    t.set_square(1)
    t.set_order(order)
    t.set_tensor_coordinates(n, dim)
    t.partition(As)
    t.assemble(As)
    t.factorize()
    return t, As


@pytest.mark.parametrize("name,n,dim,tol,prec,order", [
    ("advdiff2d_64", 64, 2, 1e-2, "fp64", 1.0),
    ("advdiff2d_128", 128, 2, 1e-3, "fp64", 1.0),
    ("advdiff3d_16", 16, 3, 1e-2, "fp64", 1.0),
    ("advdiff3d_32", 32, 3, 1e-2, "fp64", 1.0),
    ("advdiff3d_16", 16, 3, 1e-2, "fp64", 1.5),
    ("advdiff3d_16", 16, 3, 1e-2, "fp32", 1.0),
    ("advdiff2d_64", 64, 2, 1e-2, "fp64-scale0", 1.0),
    ("advdiff3d_16", 16, 3, 1e-2, "fp64-scale0", 1.0),
])
def test_transpose_apply_adjoint_identity(name, n, dim, tol, prec, order):
    import spaqr_public_amd as sq
    t, As = _tree(name, n, dim, tol, sq.FP32 if prec == "fp32" else sq.FP64, order,
                  0 if prec.endswith("scale0") else 1)
    rng = np.random.default_rng(7)
    N = As.shape[0]
    eps = 2e-6 if prec == "fp32" else 1e-12
    for _ in range(3):
        u, v = rng.standard_normal(N), rng.standard_normal(N)
        Mtu, Mv = t.solve_transpose(u), t.solve(v)
        lhs, rhs = Mtu @ v, u @ Mv
        scale = np.linalg.norm(Mtu) * np.linalg.norm(v) + np.linalg.norm(u) * np.linalg.norm(Mv)
        assert abs(lhs - rhs) <= eps * scale, (lhs, rhs)


def test_transpose_apply_dense():
    """Small tree: the dense matrix of M^{-T} equals the transpose of M^{-1}."""
    t, As = _tree("advdiff2d_16", 16, 2, 1e-1)
    N = As.shape[0]
    E = np.eye(N)
    Minv = np.stack([t.solve(E[:, j]) for j in range(N)], axis=1)
    MinvT = np.stack([t.solve_transpose(E[:, j]) for j in range(N)], axis=1)
    assert np.abs(MinvT - Minv.T).max() <= 1e-12 * np.abs(Minv).max()


def test_gmres_transpose():
    import spaqr_public_amd as sq
    from spaqr_public_amd import hip
    t, As = _tree("advdiff3d_16", 16, 3, 1e-2)
    M = sq.Matrix(t, As)
    g = np.random.default_rng(3).standard_normal(As.shape[0])
    db, dx = hip.DeviceArray.from_host(g), hip.DeviceArray.from_host(np.zeros(As.shape[0]))
    it, est = sq.gmres_transpose_device(t, M, db.ptr, dx.ptr, 300, 300, 1e-12)
    lam = dx.to_host()
    assert np.linalg.norm(As.T @ lam - g) / np.linalg.norm(g) <= 1e-12
    assert it <= 20


def _instance(n, vmag, seed):
    """configs[4] family: 2-D advection-diffusion, grid n^2, velocity vmag
    along (1, 1)/sqrt(2)."""
    from spaqr_public_amd import gen
    v = vmag / np.sqrt(2.0)
    A = gen.advection_diffusion(n, 2, 1.0, [v, v])
    rng = np.random.default_rng(seed)
    N = n * n
    return A.tocsc(), rng.standard_normal(N), rng.standard_normal(N)


def _reference_grad(A, b, xt):
    import scipy.sparse.linalg as spl
    lu = spl.splu(A.tocsc())
    x = lu.solve(b)
    g = 2.0 * (x - xt)
    lam = lu.solve(g, trans="T")
    col = np.repeat(np.arange(A.shape[1]), np.diff(A.indptr))
    return x, -lam[A.indices] * x[col], lam


@pytest.mark.parametrize("n,vmag,prec", [(32, 10.0, "fp64"), (48, 500.0, "fp64"), (64, 1000.0, "fp64"),
                                         (48, 200.0, "fp32")])
def test_autograd_gradients(n, vmag, prec):
    import torch
    import scipy.sparse.linalg as spl
    import spaqr_public_amd as sq
    from spaqr_public_amd import autograd as sg
    A, b, xt = _instance(n, vmag, seed=n)
    solver = sg.SparseSolver(A.indptr, A.indices, A.data, n, 2, tol=1e-3,
                             precision=sq.FP32 if prec == "fp32" else sq.FP64)
    vals = torch.tensor(A.data, dtype=torch.float64, device="cuda", requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, device="cuda", requires_grad=True)
    x = sg.solve(vals, bt, solver)
    loss = ((x - torch.tensor(xt, device="cuda")) ** 2).sum()
    loss.backward()
    x_ref, dA_ref, db_ref = _reference_grad(A, b, xt)
    assert np.linalg.norm(x.detach().cpu().numpy() - x_ref) <= 1e-10 * np.linalg.norm(x_ref)
    dA = vals.grad.cpu().numpy()
    db = bt.grad.cpu().numpy()
    assert np.linalg.norm(dA - dA_ref) <= 1e-6 * np.linalg.norm(dA_ref)
    assert np.linalg.norm(db - db_ref) <= 1e-6 * np.linalg.norm(db_ref)
    # central finite differences in FP64 on 32 random nonzeros, h = 1e-4 |a_k|.
    # L(+h) - L(-h) is formed as sum (x+ - x-)(x+ + x- - 2 x_t), the same
    # number without cancelling the O(N) constant part of L in floating point
    rng = np.random.default_rng(1)

    def x_at(k, d):
        Ad = A.copy()
        Ad.data[k] += d
        return spl.spsolve(Ad.tocsc(), b)

    for k in rng.choice(A.nnz, size=32, replace=False):
        h = 1e-4 * abs(A.data[k])
        xp, xm = x_at(k, h), x_at(k, -h)
        fd = np.sum((xp - xm) * (xp + xm - 2.0 * xt)) / (2 * h)
        assert abs(fd - dA[k]) <= 1e-4 * (np.abs(dA_ref).max() + abs(fd)), (k, fd, dA[k])
