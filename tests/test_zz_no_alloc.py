"""spaqr_gmres_device_noalloc: GMRES whose work vectors live in the
library's fixed device array instead of allocated memory.  Checked against
spaqr_gmres_device on the same trees and right-hand sides, to the GMRES bar
(same iteration count within +-1, both at the 1e-12 stop, solutions equal to
1e-10 relative), on consecutive calls whose second input is smaller (the
array keeps the previous call's contents), for an FP64 and an FP32 tree, and
the size limits refused with SPAQR_E_ARG before any GPU work.
"""
import ctypes as C

import numpy as np
import pytest

from spaqr_public_amd import gen

pytestmark = pytest.mark.gpu

SPAQR_E_ARG = -2


def _lib():
    import spaqr_public_amd as sq
    L = sq.load_library()
    f = L.spaqr_gmres_device_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int,
                  C.POINTER(C.c_int), C.POINTER(C.c_double)]
    f.restype = C.c_int
    return L


def _setup(name, n, precision):
    import spaqr_public_amd as sq
    A = gen.config_matrix(name)
    As, _ = sq.colscale(A)
    t = sq.Tree(sq.default_levels(As.shape[1]), 0, precision)
    t.set_scale(1)
    t.set_tol(1e-2)
    t.set_square(1)
    t.set_tensor_coordinates(n, 2)
    t.partition(As)
    t.assemble(As)
    t.factorize()
    return t, sq.Matrix(t, As), As


def _both(L, t, M, As, restart):
    import spaqr_public_amd as sq
    from spaqr_public_amd import hip
    b = np.random.default_rng(As.shape[0]).standard_normal(As.shape[0])
    db = hip.DeviceArray.from_host(b)
    x1 = hip.DeviceArray.from_host(np.zeros_like(b))
    x2 = hip.DeviceArray.from_host(np.zeros_like(b))
    it1, est1 = sq.gmres_device(t, M, db.ptr, x1.ptr, 300, restart, 1e-12)
    it2, est2 = C.c_int(0), C.c_double(0)
    rc = L.spaqr_gmres_device_noalloc(t.h, M.h, db.ptr, x2.ptr, 300, restart, 1e-12, 0, C.byref(it2),
                                      C.byref(est2))
    assert rc == 0, L.spaqr_last_error().decode()
    a, c = x1.to_host(), x2.to_host()
    for arr in (db, x1, x2):
        arr.free()
    r1 = np.linalg.norm(As @ a - b) / np.linalg.norm(b)
    r2 = np.linalg.norm(As @ c - b) / np.linalg.norm(b)
    print(f"n={As.shape[0]} restart {restart}: allocating {it1} its ({r1:.2e}), no-alloc {it2.value} its ({r2:.2e})")
    assert abs(it1 - it2.value) <= 1
    assert r1 <= 1e-12 and r2 <= 1e-12
    assert np.linalg.norm(a - c) / np.linalg.norm(a) <= 1e-10


def test_noalloc_matches_allocating_gmres(gpu_lib):
    import spaqr_public_amd as sq
    L = _lib()
    # larger input first, then smaller ones on the same GPU: the second and
    # third calls find the first call's values in the array
    for name, n, prec in (("advdiff2d_64", 64, sq.FP64), ("advdiff2d_32", 32, sq.FP32),
                          ("advdiff2d_32", 32, sq.FP64)):
        t, M, As = _setup(name, n, prec)
        _both(L, t, M, As, 300)
        M.close()
        t.close()


def test_noalloc_limits(gpu_lib):
    """n (restart + 1) beyond the array, or restart beyond 1024, is refused
    before any GPU work with a message naming the limit."""
    import spaqr_public_amd as sq
    L = _lib()
    t, M, As = _setup("advdiff2d_128", 128, sq.FP64)  # n = 16384
    it, est = C.c_int(0), C.c_double(0)
    rc = L.spaqr_gmres_device_noalloc(t.h, M.h, None, None, 300, 2000, 1e-12, 0, C.byref(it), C.byref(est))
    assert rc == SPAQR_E_ARG and b"exceeds the maximum" in L.spaqr_last_error()
    rc = L.spaqr_gmres_device_noalloc(t.h, M.h, None, None, 1000, 1000, 1e-12, 0, C.byref(it), C.byref(est))
    assert rc == SPAQR_E_ARG and b"doubles of scratch" in L.spaqr_last_error()
    # within the limit, the same problem runs and matches the allocating form
    _both(L, t, M, As, 100)
    M.close()
    t.close()
