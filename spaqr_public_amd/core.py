"""ctypes binding of libspaqr.so (include/spaqr.h).

Mirrors the reference's host API: ``Tree(nlevels, skip)`` with the setters of
tree.cpp:14-20, ``partition`` / ``assemble`` / ``factorize`` / ``solve``
(tree.h:139-147), and ``gmres`` (is.cpp:95).  All numerics run in the HIP
library; there is no CPU fallback: a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libspaqr.so")

FP64 = 8
FP32 = 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


class SpaqrError(RuntimeError):
    pass


SYMBOLS = [
    "spaqr_last_error", "spaqr_tree_create", "spaqr_tree_destroy", "spaqr_set_tol", "spaqr_set_skip",
    "spaqr_set_scale", "spaqr_set_square", "spaqr_set_order", "spaqr_set_hsl", "spaqr_set_coordinates",
    "spaqr_set_tensor_coordinates", "spaqr_set_verbose", "spaqr_rows", "spaqr_cols", "spaqr_levels",
    "spaqr_partition", "spaqr_assemble", "spaqr_factorize", "spaqr_solve", "spaqr_solve_device",
    "spaqr_solve_nrml", "spaqr_matrix_create", "spaqr_matrix_destroy", "spaqr_gmres", "spaqr_gmres_device",
    "spaqr_colscale", "spaqr_random", "spaqr_perms", "spaqr_level_stats", "spaqr_counters",
    "spaqr_num_sparsified", "spaqr_sparsified", "spaqr_num_clusters", "spaqr_cluster_table",
    "spaqr_set_phase_sync", "spaqr_kernel_timing", "spaqr_kernel_stat_count", "spaqr_kernel_stat",
    "spaqr_kernel_stats_reset", "spaqr_solve_nrml_device", "spaqr_cgls", "spaqr_cgls_device",
    "spaqr_solve_transpose", "spaqr_solve_transpose_device", "spaqr_gmres_transpose_device",
    "spaqr_csc_outer_device", "spaqr_num_extra_sparsified", "spaqr_extra_sparsified", "spaqr_profile_json",
    "spaqr_gmres_device_noalloc", "spaqr_set_row_parts", "spaqr_row_parts",
]


def load_library(path: str = LIB_PATH):
    """Load libspaqr.so and declare its C signatures (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SpaqrError(f"libspaqr.so not built ({path}); run __graft_entry__.build()")
    L = C.CDLL(path)
    L.spaqr_last_error.restype = C.c_char_p
    L.spaqr_tree_create.restype = C.c_void_p
    L.spaqr_tree_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    L.spaqr_tree_destroy.argtypes = [C.c_void_p]
    for n, t in [("spaqr_set_tol", C.c_double), ("spaqr_set_skip", C.c_int), ("spaqr_set_scale", C.c_int),
                 ("spaqr_set_square", C.c_int), ("spaqr_set_order", C.c_float), ("spaqr_set_hsl", C.c_int),
                 ("spaqr_set_verbose", C.c_int)]:
        getattr(L, n).argtypes = [C.c_void_p, t]
    L.spaqr_set_coordinates.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int]
    L.spaqr_set_tensor_coordinates.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.spaqr_set_row_parts.argtypes = [C.c_void_p, _ip, C.c_int]
    L.spaqr_row_parts.argtypes = [C.c_int, C.c_int, _ip, _ip, C.c_int, _ip]
    for n in ("spaqr_rows", "spaqr_cols", "spaqr_levels", "spaqr_num_sparsified", "spaqr_num_clusters",
              "spaqr_num_extra_sparsified"):
        getattr(L, n).argtypes = [C.c_void_p]
    L.spaqr_partition.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _dp]
    L.spaqr_assemble.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _dp]
    L.spaqr_factorize.argtypes = [C.c_void_p]
    L.spaqr_solve.argtypes = [C.c_void_p, _dp, _dp]
    L.spaqr_solve_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.spaqr_solve_nrml.argtypes = [C.c_void_p, _dp, _dp]
    L.spaqr_solve_nrml_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.spaqr_cgls.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_int, _ip, _dp]
    L.spaqr_cgls_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_int,
                                    _ip, _dp]
    L.spaqr_matrix_create.restype = C.c_void_p
    L.spaqr_matrix_create.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _dp]
    L.spaqr_matrix_destroy.argtypes = [C.c_void_p]
    L.spaqr_gmres.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_int, _ip, _dp]
    L.spaqr_gmres_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                     C.c_double, C.c_int, _ip, _dp]
    L.spaqr_solve_transpose.argtypes = [C.c_void_p, _dp, _dp]
    L.spaqr_solve_transpose_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.spaqr_gmres_transpose_device.argtypes = L.spaqr_gmres_device.argtypes
    L.spaqr_gmres_device_noalloc.argtypes = L.spaqr_gmres_device.argtypes
    L.spaqr_csc_outer_device.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    L.spaqr_colscale.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp]
    L.spaqr_random.argtypes = [C.c_int, C.c_int, _dp]
    L.spaqr_perms.argtypes = [C.c_void_p, _ip, _ip]
    L.spaqr_level_stats.argtypes = [C.c_void_p, _dp]
    L.spaqr_counters.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong), _dp]
    L.spaqr_sparsified.argtypes = [C.c_void_p, _ip]
    L.spaqr_extra_sparsified.argtypes = [C.c_void_p, _ip]
    L.spaqr_profile_json.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
    L.spaqr_cluster_table.argtypes = [C.c_void_p, _ip]
    L.spaqr_set_phase_sync.argtypes = [C.c_void_p, C.c_int]
    L.spaqr_kernel_timing.argtypes = [C.c_void_p, C.c_int]
    L.spaqr_kernel_stat_count.argtypes = [C.c_void_p]
    L.spaqr_kernel_stat.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int, _dp]
    L.spaqr_kernel_stats_reset.argtypes = [C.c_void_p]
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _check(rc, what):
    if rc != 0:
        raise SpaqrError(f"{what} failed ({rc}): {load_library().spaqr_last_error().decode()}")


def _csc_arrays(A):
    A = sp.csc_matrix(A, copy=True)  # never alias (or reorder) the caller's arrays
    A.sum_duplicates()
    A.sort_indices()
    return (A.shape[0], A.shape[1], np.ascontiguousarray(A.indptr, dtype=np.int32),
            np.ascontiguousarray(A.indices, dtype=np.int32), np.ascontiguousarray(A.data, dtype=np.float64))


def colscale(A):
    """spaQR.cpp:131-147 on a scipy matrix: returns (A D * 10, D)."""
    m, n, cp, ri, v = _csc_arrays(A)
    D = np.zeros(n)
    _check(load_library().spaqr_colscale(m, n, _i(cp), _i(ri), _d(v), _d(D)), "colscale")
    return sp.csc_matrix((v, ri, cp), shape=(m, n)), D


def random_vec(n, seed=2021):
    """util.cpp:943-952: mt19937(seed), uniform(-1, 1)."""
    out = np.zeros(n)
    _check(load_library().spaqr_random(n, seed, _d(out)), "random")
    return out


def default_levels(ncols):
    """spaQR.cpp:111 (integer division inside the log)."""
    return int(math.ceil(math.log2(ncols // 64)))


class Tree:
    """The reference's Tree (tree.h:36-152) on one MI355X."""

    def __init__(self, nlevels, skip=0, precision=FP64, device=0):
        L = load_library()
        self.nlevels = nlevels
        self.h = L.spaqr_tree_create(nlevels, skip, precision, device)
        if not self.h:
            raise SpaqrError("tree creation failed: " + L.spaqr_last_error().decode())
        self._A = None

    def close(self):
        if getattr(self, "h", None):
            load_library().spaqr_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # setters (tree.cpp:14-20)
    def set_tol(self, t): _check(load_library().spaqr_set_tol(self.h, t), "set_tol")
    def set_skip(self, s): _check(load_library().spaqr_set_skip(self.h, s), "set_skip")
    def set_scale(self, s): _check(load_library().spaqr_set_scale(self.h, s), "set_scale")
    def set_square(self, s): _check(load_library().spaqr_set_square(self.h, s), "set_square")
    def set_order(self, o): _check(load_library().spaqr_set_order(self.h, o), "set_order")
    def set_hsl(self, h): _check(load_library().spaqr_set_hsl(self.h, h), "set_hsl")
    def set_verbose(self, v): _check(load_library().spaqr_set_verbose(self.h, v), "set_verbose")

    def set_tensor_coordinates(self, n, dim):
        _check(load_library().spaqr_set_tensor_coordinates(self.h, n, dim), "set_tensor_coordinates")

    def set_row_parts(self, parts):
        """Row partition for a tree without coordinates: parts[i] in
        [0, 2^(nlevels-1)), parts 2g and 2g+1 the halves of part g (the
        output of partition_patoh, partition.cpp:422-448).  Without it the
        library partitions the rows itself (row_parts)."""
        p = np.ascontiguousarray(parts, dtype=np.int32)
        _check(load_library().spaqr_set_row_parts(self.h, _i(p), len(p)), "set_row_parts")

    def set_coordinates(self, X):
        """X: dim x n array of coordinates (the reference's Xcoo)."""
        X = np.asarray(X, dtype=np.float64)
        dim, n = X.shape
        flat = np.ascontiguousarray(X.T)
        _check(load_library().spaqr_set_coordinates(self.h, _d(flat), dim, n), "set_coordinates")

    def rows(self): return load_library().spaqr_rows(self.h)
    def cols(self): return load_library().spaqr_cols(self.h)
    def levels(self): return load_library().spaqr_levels(self.h)

    def partition(self, A):
        self._A = _csc_arrays(A)
        m, n, cp, ri, v = self._A
        self.m, self.n = m, n
        _check(load_library().spaqr_partition(self.h, m, n, _i(cp), _i(ri), _d(v)), "partition")

    def assemble(self, A=None):
        if A is not None:
            self._A = _csc_arrays(A)
        m, n, cp, ri, v = self._A
        _check(load_library().spaqr_assemble(self.h, m, n, _i(cp), _i(ri), _d(v)), "assemble")

    def factorize(self):
        _check(load_library().spaqr_factorize(self.h), "factorize")

    def solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        _check(load_library().spaqr_solve(self.h, _d(b), _d(x)), "solve")
        return x

    def solve_device(self, b_ptr, x_ptr):
        """Device-pointer apply (e.g. torch tensors' data_ptr())."""
        _check(load_library().spaqr_solve_device(self.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr)), "solve_device")

    def solve_transpose(self, b):
        """x = M^{-T} b (square trees): the adjoint of solve()."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.m)
        _check(load_library().spaqr_solve_transpose(self.h, _d(b), _d(x)), "solve_transpose")
        return x

    def solve_transpose_device(self, b_ptr, x_ptr):
        _check(load_library().spaqr_solve_transpose_device(self.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr)),
               "solve_transpose_device")

    def solve_nrml(self, b):
        """tree.cpp:1471-1517: x = (W^T W)^{-1} b (least-squares tree, b of length cols)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        _check(load_library().spaqr_solve_nrml(self.h, _d(b), _d(x)), "solve_nrml")
        return x

    def solve_nrml_device(self, b_ptr, x_ptr):
        _check(load_library().spaqr_solve_nrml_device(self.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr)),
               "solve_nrml_device")

    # inspection
    def perms(self):
        cp = np.zeros(self.n, dtype=np.int32)
        rp = np.zeros(self.m, dtype=np.int32)
        _check(load_library().spaqr_perms(self.h, _i(cp), _i(rp)), "perms")
        return cp, rp

    def level_stats(self):
        out = np.zeros(8 * self.nlevels)
        _check(load_library().spaqr_level_stats(self.h, _d(out)), "level_stats")
        return out.reshape(self.nlevels, 8)

    def counters(self):
        nnz = (C.c_ulonglong * 4)()
        dbl = np.zeros(8)
        _check(load_library().spaqr_counters(self.h, nnz, _d(dbl)), "counters")
        return dict(nnzA=nnz[0], nnzR=nnz[1], nnzH=nnz[2], nnzQ=nnz[3], flops=dbl[0], t_partition=dbl[1],
                    t_assemble=dbl[2], t_factorize=dbl[3], top_size=int(dbl[4]), device_bytes=dbl[5],
                    factor_bytes=dbl[6], apply_bytes=dbl[7])

    def sparsified(self):
        L = load_library()
        k = L.spaqr_num_sparsified(self.h)
        out = np.zeros(5 * max(k, 1), dtype=np.int32)
        _check(L.spaqr_sparsified(self.h, _i(out)), "sparsified")
        return out[: 5 * k].reshape(k, 5)

    def profile(self):
        """The factorization's profile (profile.h:13-106) as a dict: levels,
        interfaces (cols before / after, neighbours, wave), merged clusters,
        least-squares extra rows and, with kernel timing, kernel families."""
        import json
        L = load_library()
        n = L.spaqr_profile_json(self.h, None, 0)
        _check(min(n, 0), "profile_json")
        buf = C.create_string_buffer(n + 1)
        _check(min(L.spaqr_profile_json(self.h, buf, n + 1), 0), "profile_json")
        return json.loads(buf.value.decode())

    def extra_sparsified(self):
        """Least squares: (ilvl, lvl, sep, extra rows before, kept) of every
        sparsify_extra compression (tree.cpp:1260-1307), in call order."""
        L = load_library()
        k = L.spaqr_num_extra_sparsified(self.h)
        out = np.zeros(5 * max(k, 1), dtype=np.int32)
        _check(L.spaqr_extra_sparsified(self.h, _i(out)), "extra_sparsified")
        return out[: 5 * k].reshape(k, 5)

    def set_phase_sync(self, on):
        _check(load_library().spaqr_set_phase_sync(self.h, 1 if on else 0), "set_phase_sync")

    def kernel_timing(self, on=True):
        _check(load_library().spaqr_kernel_timing(self.h, 1 if on else 0), "kernel_timing")

    def kernel_stats(self, reset=False):
        """{family: dict(launches, ms, flops, bytes)} from HIP events."""
        L = load_library()
        n = L.spaqr_kernel_stat_count(self.h)
        if n < 0:
            _check(n, "kernel_stat_count")
        out = {}
        buf = C.create_string_buffer(64)
        v = np.zeros(4)
        for i in range(n):
            _check(L.spaqr_kernel_stat(self.h, i, buf, 64, _d(v)), "kernel_stat")
            out[buf.value.decode()] = dict(launches=int(v[0]), ms=v[1], flops=v[2], bytes=v[3])
        if reset:
            _check(L.spaqr_kernel_stats_reset(self.h), "kernel_stats_reset")
        return out

    def cluster_table(self):
        L = load_library()
        k = L.spaqr_num_clusters(self.h)
        out = np.zeros(12 * max(k, 1), dtype=np.int32)
        _check(L.spaqr_cluster_table(self.h, _i(out)), "cluster_table")
        return out[: 12 * k].reshape(k, 12)


class Matrix:
    """Device copy of the operator for the Krylov drivers."""

    def __init__(self, tree: Tree, A):
        m, n, cp, ri, v = _csc_arrays(A)
        self.m, self.n, self.nnz = m, n, int(cp[-1])
        self.h = load_library().spaqr_matrix_create(tree.h, m, n, _i(cp), _i(ri), _d(v))
        if not self.h:
            raise SpaqrError("matrix upload failed: " + load_library().spaqr_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            load_library().spaqr_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def gmres(tree: Tree, M: Matrix, b, x0=None, iters=300, restart=None, tol=1e-12, verbose=0):
    """is.cpp:95-272 with the tree as left preconditioner; returns (x, iters, est)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(M.n) if x0 is None else np.array(x0, dtype=np.float64)
    it = C.c_int(0)
    est = C.c_double(0)
    _check(load_library().spaqr_gmres(tree.h, M.h, _d(b), _d(x), iters, restart or iters, tol, verbose,
                                      C.byref(it), C.byref(est)), "gmres")
    return x, it.value, est.value


def gmres_device(tree: Tree, M: Matrix, b_ptr, x_ptr, iters=300, restart=None, tol=1e-12, verbose=0):
    it = C.c_int(0)
    est = C.c_double(0)
    _check(load_library().spaqr_gmres_device(tree.h, M.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), iters,
                                             restart or iters, tol, verbose, C.byref(it), C.byref(est)),
           "gmres_device")
    return it.value, est.value


def gmres_transpose_device(tree: Tree, M: Matrix, b_ptr, x_ptr, iters=300, restart=None, tol=1e-12, verbose=0):
    """GMRES on A^T x = b preconditioned with M^{-T} (the adjoint solve)."""
    it = C.c_int(0)
    est = C.c_double(0)
    _check(load_library().spaqr_gmres_transpose_device(tree.h, M.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), iters,
                                                       restart or iters, tol, verbose, C.byref(it), C.byref(est)),
           "gmres_transpose_device")
    return it.value, est.value


def csc_outer_device(M: Matrix, alpha, u_ptr, v_ptr, out_ptr):
    """out[k] = alpha * u[i_k] * v[j_k] over M's nonzeros in CSC order."""
    _check(load_library().spaqr_csc_outer_device(M.h, alpha, C.c_void_p(u_ptr), C.c_void_p(v_ptr),
                                                 C.c_void_p(out_ptr)), "csc_outer_device")


def cgls(tree: Tree, M: Matrix, b, x0=None, iters=300, tol=1e-10, verbose=0):
    """is.cpp:275-346 preconditioned by the tree's normal-equation apply;
    returns (x, iters, est) with est = ||A^T r|| / ||A^T b||."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(M.n) if x0 is None else np.array(x0, dtype=np.float64)
    it = C.c_int(0)
    est = C.c_double(0)
    _check(load_library().spaqr_cgls(tree.h, M.h, _d(b), _d(x), iters, tol, verbose, C.byref(it), C.byref(est)),
           "cgls")
    return x, it.value, est.value


def cgls_device(tree: Tree, M: Matrix, b_ptr, x_ptr, iters=300, tol=1e-10, verbose=0):
    it = C.c_int(0)
    est = C.c_double(0)
    _check(load_library().spaqr_cgls_device(tree.h, M.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), iters, tol, verbose,
                                            C.byref(it), C.byref(est)), "cgls_device")
    return it.value, est.value


def row_parts(A, nlevels):
    """The library's row partition of A into 2^(nlevels-1) parts (recursive
    bisection of the column-net hypergraph with Fiduccia-Mattheyses
    refinement), in place of partition.cpp:422-448 (PaToH).  Host only."""
    A = sp.csc_matrix(A)
    A.sort_indices()
    m, n = A.shape
    ip = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ix = np.ascontiguousarray(A.indices, dtype=np.int32)
    out = np.zeros(m, dtype=np.int32)
    _check(load_library().spaqr_row_parts(m, n, _i(ip), _i(ix), nlevels, _i(out)), "row_parts")
    return out


def run_flow(A, n, dim, tol, nlevels=None, skip=0, scale=1, order=1.0, iters=300, res=1e-12, precision=FP64,
             rhs=None, verbose=0, device=0, parts=None):
    """The driver flow of spaQR.cpp:85-320 on the GPU: GMRES for square
    systems, CGLS on the normal equations for tall ones.  n = None: no
    coordinates (the non-geometric partition, from `parts` when given)."""
    import time
    A = sp.csc_matrix(A)
    if A.shape[0] < A.shape[1]:
        A = A.T.tocsc()
    m, ncols = A.shape
    if nlevels is None:
        nlevels = default_levels(ncols)
    As, D = colscale(A)
    if tol == 0:
        skip = nlevels - 1
    t = Tree(nlevels, skip, precision, device)
    t.set_scale(scale if m == ncols else 1)
    t.set_tol(tol)
    t.set_order(order)
    t.set_square(1 if m == ncols else 0)
    t.set_verbose(verbose)
    if n is not None:
        t.set_tensor_coordinates(n, dim)
    if parts is not None:
        t.set_row_parts(parts)
    t.partition(As)
    t.assemble(As)
    t.factorize()
    out = dict(nlevels=nlevels, tree=t, A=As, D=D)
    b = random_vec(m, 2021) if rhs is None else np.asarray(rhs, dtype=np.float64)
    M = Matrix(t, As)
    t0 = time.perf_counter()
    if m == ncols:
        x = t.solve(b)
        out["t_solve"] = time.perf_counter() - t0
        out["oneshot_res"] = float(np.linalg.norm(As @ x - b) / np.linalg.norm(b))
        t0 = time.perf_counter()
        x, it, est = gmres(t, M, b, iters=iters, tol=res, verbose=verbose)
        out["t_krylov"] = time.perf_counter() - t0
        out["final_res"] = float(np.linalg.norm(As @ x - b) / np.linalg.norm(b))
    else:
        x = t.solve_nrml(As.T @ b)
        out["t_solve"] = time.perf_counter() - t0
        out["oneshot_res"] = float(np.linalg.norm(As.T @ (As @ x - b)) / np.linalg.norm(As.T @ b))
        t0 = time.perf_counter()
        x, it, est = cgls(t, M, b, iters=iters, tol=res, verbose=verbose)
        out["t_krylov"] = time.perf_counter() - t0
        out["final_res"] = float(np.linalg.norm(As.T @ (As @ x - b)) / np.linalg.norm(As.T @ b))
    out.update(iters=it, est=est, x=x, b=b, M=M)
    out.update(t.counters())
    out["levels"] = t.level_stats()
    return out
