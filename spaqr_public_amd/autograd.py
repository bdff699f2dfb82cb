"""Differentiable sparse solve x = A^{-1} b on the GPU (torch.autograd).

The forward pass is the spaQR-preconditioned FP64 GMRES of the C ABI
(`spaqr_gmres_device`).  The backward pass solves the adjoint system
A^T lambda = dL/dx by GMRES preconditioned with the transposed factors
(`spaqr_gmres_transpose_device`, M^{-T}: R^{-T} sweeps, then the Householder
products in reverse level order) and forms dL/dA = -lambda_i x_j on A's
sparsity pattern in one kernel (`spaqr_csc_outer_device`); dL/db = lambda.

The factorization is only a preconditioner: it approximates A, so it is not
differentiated.  The solver factors the column-scaled matrix As = A diag(s),
s_j = 10 / ||A(:, j)||_2 (the reference driver's scaling, spaQR.cpp:131-147),
and `solve` differentiates through As_vals = A_vals * s[col] and x = s * y with
As y = b, s held constant.  torch is used only as this thin harness: device
memory, streams and the autograd tape; every solve runs in libspaqr.so.
Each solver works on its own HIP stream and waits only for torch's current
stream before it reads torch tensors, so solvers may run in parallel threads.
"""
import numpy as np
import scipy.sparse as sp
import torch

from . import core


class SparseSolver:
    """A square sparse operator with a spaQR factorization on the GPU.

    indptr/indices/values: CSC arrays (sorted row indices, no duplicates).
    n, dim: the tensor grid of the unknowns (geometric partition).
    tol: spaQR sparsification tolerance; res: GMRES tolerance.
    precision: core.FP64, or core.FP32 (FP32 factors, FP64 refinement)."""

    def __init__(self, indptr, indices, values, n, dim, tol=1e-3, res=1e-12, precision=core.FP64, nlevels=None,
                 iters=300, device=0):
        indptr = np.asarray(indptr, dtype=np.int32)
        indices = np.asarray(indices, dtype=np.int32)
        vals = torch.as_tensor(values).detach().to("cpu", torch.float64).numpy()
        N = len(indptr) - 1
        A = sp.csc_matrix((vals, indices, indptr), shape=(N, N))
        if not A.has_sorted_indices:
            raise ValueError("SparseSolver: CSC row indices must be sorted")
        As, D = core.colscale(A)
        self.N, self.nnz = N, int(indptr[-1])
        self.values = torch.as_tensor(vals.copy())
        self.col = torch.as_tensor(np.repeat(np.arange(N), np.diff(indptr)), dtype=torch.int64)
        self.s = torch.as_tensor(10.0 * D)
        self.res, self.iters = res, iters
        self.device = torch.device("cuda", device)
        if nlevels is None:
            nlevels = core.default_levels(N)
        t = core.Tree(nlevels, 0, precision, device)
        t.set_tol(tol)
        t.set_scale(1)
        t.set_square(1)
        t.set_tensor_coordinates(n, dim)
        t.partition(As)
        t.assemble(As)
        t.factorize()
        self.tree = t
        self.M = core.Matrix(t, As)
        self.last_iters = (0, 0)

    def close(self):
        """Release the device factors and matrix now (instead of at garbage
        collection)."""
        if getattr(self, "M", None) is not None:
            self.M.close()
            self.M = None
        if getattr(self, "tree", None) is not None:
            self.tree.close()
            self.tree = None

    def _vec(self, v):
        v = v.detach().to(self.device, torch.float64).contiguous()
        if v.shape != (self.N,):
            raise ValueError(f"expected a vector of length {self.N}, got {tuple(v.shape)}")
        return v

    def solve_scaled(self, b):
        """y = As^{-1} b by GMRES (FP64 vectors on the device)."""
        b = self._vec(b)
        y = torch.zeros_like(b)
        torch.cuda.current_stream(self.device).synchronize()
        it, _ = core.gmres_device(self.tree, self.M, b.data_ptr(), y.data_ptr(), self.iters, self.iters, self.res)
        self.last_iters = (it, self.last_iters[1])
        return y

    def solve_scaled_adjoint(self, g):
        """lambda = As^{-T} g by GMRES with M^{-T}."""
        g = self._vec(g)
        lam = torch.zeros_like(g)
        torch.cuda.current_stream(self.device).synchronize()
        it, _ = core.gmres_transpose_device(self.tree, self.M, g.data_ptr(), lam.data_ptr(), self.iters,
                                            self.iters, self.res)
        self.last_iters = (self.last_iters[0], it)
        return lam

    def pattern_outer(self, u, v, alpha=-1.0):
        """alpha * u_i v_j on As's nonzeros, CSC order."""
        u, v = self._vec(u), self._vec(v)
        out = torch.empty(self.nnz, dtype=torch.float64, device=self.device)
        torch.cuda.current_stream(self.device).synchronize()
        core.csc_outer_device(self.M, alpha, u.data_ptr(), v.data_ptr(), out.data_ptr())
        return out


class _ScaledSolve(torch.autograd.Function):
    """y = As^{-1} b with gradients w.r.t. As's values and b."""

    @staticmethod
    def forward(ctx, as_values, b, solver):
        y = solver.solve_scaled(b)
        ctx.solver = solver
        ctx.save_for_backward(y)
        return y.to(b.device)

    @staticmethod
    def backward(ctx, gy):
        (y,) = ctx.saved_tensors
        s = ctx.solver
        lam = s.solve_scaled_adjoint(gy)
        dvals = s.pattern_outer(lam, y, -1.0) if ctx.needs_input_grad[0] else None
        db = lam if ctx.needs_input_grad[1] else None
        return dvals, db, None


def solve(values, b, solver: SparseSolver):
    """x = A^{-1} b, differentiable w.r.t. A's CSC values and b.  `values`
    must equal the values the solver was built from (the factorization is
    the preconditioner of exactly that matrix)."""
    if not torch.equal(values.detach().to("cpu", torch.float64), solver.values):
        raise ValueError("solve: values differ from those the solver was factored with; build a new SparseSolver")
    dev = solver.device
    s = solver.s.to(dev)
    as_values = values.to(dev, torch.float64) * s[solver.col.to(dev)]
    y = _ScaledSolve.apply(as_values, b.to(dev, torch.float64), solver)
    return s * y
