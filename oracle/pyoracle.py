"""ORACLE — test infrastructure only.

ctypes binding of the CPU spaQR restatement (oracle/_build/liboracle.so).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker; the product never does.
"""
from __future__ import annotations

import ctypes as C
import glob
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    cmd = ["make", "-C", HERE, "-j8"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = C.CDLL(LIB)
        L.orc_tree_new.restype = C.c_void_p
        L.orc_tree_new.argtypes = [C.c_int, C.c_int]
        L.orc_tree_free.argtypes = [C.c_void_p]
        L.orc_tree_set.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_float, C.c_int, C.c_int]
        L.orc_tree_set_tensor_coords.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_tree_set_row_parts.argtypes = [C.c_void_p, _ip, C.c_int]
        L.orc_hypergraph_labels.argtypes = [C.c_int, C.c_int, _ip, _ip, C.c_int, _ip, _ip]
        L.orc_partition.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _dp]
        L.orc_assemble.argtypes = [C.c_void_p]
        L.orc_factorize.argtypes = [C.c_void_p]
        L.orc_solve.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_solve_nrml.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_gmres.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_int, C.c_double, _dp, C.c_int]
        L.orc_cgls.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, _dp, C.c_int]
        L.orc_perms.argtypes = [C.c_void_p, _ip, _ip]
        L.orc_level_stats.argtypes = [C.c_void_p, _dp]
        L.orc_counters.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong), _dp]
        L.orc_num_sparsified.argtypes = [C.c_void_p]
        L.orc_sparsified.argtypes = [C.c_void_p, _ip]
        L.orc_num_extra_sparsified.argtypes = [C.c_void_p]
        L.orc_extra_sparsified.argtypes = [C.c_void_p, _ip]
        L.orc_num_clusters.argtypes = [C.c_void_p]
        L.orc_cluster_table.argtypes = [C.c_void_p, _ip]
        L.orc_use_openblas.argtypes = [C.c_char_p]
        L.orc_backend.restype = C.c_char_p
        L.orc_last_error.restype = C.c_char_p
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_geqrf.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_larft.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_larfb.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_char]
        L.orc_geqp3.argtypes = [C.c_int, C.c_int, _dp, _ip, _dp]
        L.orc_ormqr_left.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_char]
        L.orc_trsm_right_upper.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_trsv_upper.argtypes = [C.c_int, _dp, _dp, C.c_char]
        L.orc_choose_rank.argtypes = [C.c_int, _dp, C.c_double]
        L.orc_colscale.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp]
        L.orc_random.argtypes = [C.c_int, C.c_int, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def openblas_path():
    """The OpenBLAS shared library bundled with scipy (exports LAPACK)."""
    import scipy
    cands = glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs", "libscipy_openblas*.so"))
    return os.path.realpath(cands[0]) if cands else None


def set_backend(name: str, threads: int = 1) -> str:
    L = lib()
    if name == "openblas":
        p = openblas_path()
        if p is None or L.orc_use_openblas(p.encode()) != 0:
            raise RuntimeError("OpenBLAS backend unavailable")
        L.orc_set_threads(threads)
    else:
        L.orc_use_own()
    return L.orc_backend().decode()


# ------------------------------------------------------------- dense kernels
def geqrf(A):
    A = np.array(A, dtype=np.float64, order="F")
    m, n = A.shape
    tau = np.zeros(min(m, n))
    lib().orc_geqrf(m, n, _d(A), _d(tau))
    return A, tau


def singular_values(A):
    """util.cpp:767-782 (gesvd_values) as restated: descending."""
    A = np.array(A, dtype=np.float64, order="F")
    m, n = A.shape
    S = np.zeros(min(m, n))
    lib().orc_singular_values(m, n, _d(A), _d(S))
    return S


def larft(V, tau):
    V = np.array(V, dtype=np.float64, order="F")
    m, k = V.shape
    T = np.zeros((k, k), order="F")
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    lib().orc_larft(m, k, _d(V), _d(tau), _d(T))
    return T


def larfb(V, T, Cm, trans="T"):
    V = np.array(V, dtype=np.float64, order="F")
    T = np.array(T, dtype=np.float64, order="F")
    Cm = np.array(Cm, dtype=np.float64, order="F")
    m, k = V.shape
    lib().orc_larfb(m, Cm.shape[1], k, _d(V), _d(T), _d(Cm), trans.encode())
    return Cm


def geqp3(A):
    A = np.array(A, dtype=np.float64, order="F")
    m, n = A.shape
    jpvt = np.zeros(n, dtype=np.int32)
    tau = np.zeros(min(m, n))
    lib().orc_geqp3(m, n, _d(A), _i(jpvt), _d(tau))
    return A, jpvt, tau


def ormqr_left(V, tau, A, trans="T"):
    V = np.array(V, dtype=np.float64, order="F")
    A = np.array(A, dtype=np.float64, order="F")
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    lib().orc_ormqr_left(A.shape[0], A.shape[1], V.shape[1], _d(V), _d(tau), _d(A), trans.encode())
    return A


def trsm_right_upper(U, B):
    U = np.array(U, dtype=np.float64, order="F")
    B = np.array(B, dtype=np.float64, order="F")
    lib().orc_trsm_right_upper(B.shape[0], B.shape[1], _d(U), _d(B))
    return B


def trsv_upper(U, x, trans="N"):
    U = np.array(U, dtype=np.float64, order="F")
    x = np.array(x, dtype=np.float64)
    lib().orc_trsv_upper(U.shape[0], _d(U), _d(x), trans.encode())
    return x


def choose_rank(s, tol):
    s = np.ascontiguousarray(s, dtype=np.float64)
    return lib().orc_choose_rank(len(s), _d(s), tol)


def colscale(A):
    """spaQR.cpp:131-147 applied to a scipy CSC matrix; returns (A_scaled, D)."""
    A = A.tocsc().copy()
    A.sort_indices()
    m, n = A.shape
    cp = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ri = np.ascontiguousarray(A.indices, dtype=np.int32)
    v = np.ascontiguousarray(A.data, dtype=np.float64)
    D = np.zeros(n)
    lib().orc_colscale(m, n, _i(cp), _i(ri), _d(v), _d(D))
    A.data = v
    return A, D


def random_vec(n, seed=2021):
    out = np.zeros(n)
    lib().orc_random(n, seed, _d(out))
    return out


# ------------------------------------------------------------- Tree
class Tree:
    """CPU restatement of the reference Tree (tree.h:36-152)."""

    def __init__(self, nlevels, skip=0):
        self.nlevels = nlevels
        self.h = lib().orc_tree_new(nlevels, skip)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    def configure(self, tol, scale=1, order=1.0, square=True, verbose=0):
        lib().orc_tree_set(self.h, tol, scale, order, 1 if square else 0, verbose)

    def set_tensor_coords(self, n, dim):
        lib().orc_tree_set_tensor_coords(self.h, n, dim)

    def set_row_parts(self, parts):
        self._parts = np.ascontiguousarray(parts, dtype=np.int32)
        lib().orc_tree_set_row_parts(self.h, _i(self._parts), len(self._parts))

    def partition(self, A):
        A = A.tocsc()
        A.sort_indices()
        self.m, self.n = A.shape
        self._cp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        self._ri = np.ascontiguousarray(A.indices, dtype=np.int32)
        self._v = np.ascontiguousarray(A.data, dtype=np.float64)
        if lib().orc_partition(self.h, self.m, self.n, _i(self._cp), _i(self._ri), _d(self._v)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def assemble(self):
        if lib().orc_assemble(self.h) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def factorize(self):
        if lib().orc_factorize(self.h) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        lib().orc_solve(self.h, _d(b), _d(x))
        return x

    def solve_nrml(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        lib().orc_solve_nrml(self.h, _d(b), _d(x))
        return x

    def gmres(self, b, iters=300, tol=1e-12, verbose=0, restart=None):
        """is.cpp:95-272; restart defaults to iters (spaQR.cpp:304)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        err = C.c_double(0)
        it = lib().orc_gmres(self.h, _d(b), _d(x), iters, restart or iters, tol, C.byref(err), verbose)
        return x, it, err.value

    def cgls(self, b, iters=300, tol=1e-12, verbose=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        err = C.c_double(0)
        it = lib().orc_cgls(self.h, _d(b), _d(x), iters, tol, C.byref(err), verbose)
        return x, it, err.value

    def perms(self):
        cp = np.zeros(self.n, dtype=np.int32)
        rp = np.zeros(self.m, dtype=np.int32)
        lib().orc_perms(self.h, _i(cp), _i(rp))
        return cp, rp

    def level_stats(self):
        out = np.zeros(8 * self.nlevels)
        lib().orc_level_stats(self.h, _d(out))
        return out.reshape(self.nlevels, 8)

    def counters(self):
        nnz = (C.c_ulonglong * 4)()
        dbl = np.zeros(5)
        lib().orc_counters(self.h, nnz, _d(dbl))
        return dict(nnzA=nnz[0], nnzR=nnz[1], nnzH=nnz[2], nnzQ=nnz[3], flops=dbl[0],
                    t_partition=dbl[1], t_assemble=dbl[2], t_factorize=dbl[3], top_size=int(dbl[4]))

    def sparsified(self):
        k = lib().orc_num_sparsified(self.h)
        out = np.zeros(5 * max(k, 1), dtype=np.int32)
        lib().orc_sparsified(self.h, _i(out))
        return out[: 5 * k].reshape(k, 5)

    def extra_sparsified(self):
        """Least squares: (ilvl, lvl, sep, extra rows before, after) of every
        sparsify_extra compression (tree.cpp:1260-1307)."""
        k = lib().orc_num_extra_sparsified(self.h)
        out = np.zeros(5 * max(k, 1), dtype=np.int32)
        lib().orc_extra_sparsified(self.h, _i(out))
        return out[: 5 * k].reshape(k, 5)

    def cluster_table(self):
        k = lib().orc_num_clusters(self.h)
        out = np.zeros(12 * k, dtype=np.int32)
        lib().orc_cluster_table(self.h, _i(out))
        return out.reshape(k, 12)


def default_levels(ncols: int) -> int:
    """spaQR.cpp:111: ceil(log2(ncols/64)) with integer division."""
    import math
    return int(math.ceil(math.log2(ncols // 64)))


def hypergraph_labels(A, nlevels, parts):
    """partition.cpp:450-499 + 582-623: column labels from a row partition,
    as an (n, 6) array of self, left, right (lvl, sep)."""
    import scipy.sparse as sp
    A = sp.csc_matrix(A)
    A.sort_indices()
    m, n = A.shape
    cp = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ri = np.ascontiguousarray(A.indices, dtype=np.int32)
    p = np.ascontiguousarray(parts, dtype=np.int32)
    out = np.zeros((n, 6), dtype=np.int32)
    if lib().orc_hypergraph_labels(m, n, _i(cp), _i(ri), nlevels, _i(p), _i(out)) != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return out


def run_reference_flow(A, n, dim, tol, nlevels=None, skip=0, scale=1, order=1.0, iters=300, res=1e-12,
                       rhs=None, verbose=0, parts=None):
    """The flow of spaQR.cpp:85-320 on a scipy matrix (column scaling, tree,
    one-shot solve, GMRES/CGLS) ; returns a result dict.  n = None: no
    coordinates, the non-geometric partition from the row `parts`."""
    import time
    import scipy.sparse as sp
    A = sp.csc_matrix(A)
    if A.shape[0] < A.shape[1]:
        A = A.T.tocsc()
    m, ncols = A.shape
    if nlevels is None:
        nlevels = default_levels(ncols)
    As, D = colscale(A)
    if tol == 0:
        skip = nlevels - 1
    t = Tree(nlevels, skip)
    t.configure(tol, scale=scale if m == ncols else 1, order=order, square=(m == ncols), verbose=verbose)
    if n is not None:
        t.set_tensor_coords(n, dim)
    if parts is not None:
        t.set_row_parts(parts)
    t.partition(As)
    t.assemble()
    t.factorize()
    out = dict(nlevels=nlevels, tree=t, A=As, D=D)
    b = random_vec(m, 2021) if rhs is None else np.asarray(rhs, dtype=np.float64)
    t0 = time.perf_counter()
    if m == ncols:
        x = t.solve(b)
        out["t_solve"] = time.perf_counter() - t0
        out["oneshot_res"] = np.linalg.norm(As @ x - b) / np.linalg.norm(b)
        t0 = time.perf_counter()
        x, it, err = t.gmres(b, iters, res, verbose)
        out["t_krylov"] = time.perf_counter() - t0
        out["final_res"] = np.linalg.norm(As @ x - b) / np.linalg.norm(b)
    else:
        x = t.solve_nrml(As.T @ b)
        out["t_solve"] = time.perf_counter() - t0
        out["oneshot_res"] = np.linalg.norm(As.T @ (As @ x - b)) / np.linalg.norm(As.T @ b)
        t0 = time.perf_counter()
        x, it, err = t.cgls(b, iters, res, verbose)
        out["t_krylov"] = time.perf_counter() - t0
        out["final_res"] = np.linalg.norm(As.T @ (As @ x - b)) / np.linalg.norm(As.T @ b)
    out.update(iters=it, est=err, x=x, b=b)
    out.update(t.counters())
    out["levels"] = t.level_stats()
    return out
