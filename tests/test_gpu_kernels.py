"""Kernel-level parity on the GPU: each batched HIP kernel family against the
oracle's LAPACK restatement (or numpy for plain BLAS shapes), FP64.

Bar: R/V/tau/T to 1e-12 relative (same Householder conventions as LAPACK's
dgeqrf/dlarft), pivots and ranks exact, solves to 1e-12 relative.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
NB = 32


def d(a):
    return a.ctypes.data_as(dp)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


@pytest.fixture(scope="module")
def K(ktest):
    ktest.ktest_gemm.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, C.c_int,
                                 C.c_int, C.c_double, C.c_double]
    ktest.ktest_gemm_split.argtypes = ktest.ktest_gemm.argtypes + [C.c_int]
    fp = C.POINTER(C.c_float)
    ktest.ktest_gemm_f32.argtypes = [C.c_int, C.c_int, C.c_int, fp, C.c_int, C.c_int, fp, C.c_int, C.c_int, fp,
                                     C.c_int, C.c_int, C.c_float, C.c_float]
    ktest.ktest_smin.argtypes = [C.c_int, dp, C.c_int, dp]
    ktest.ktest_qr.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_int, dp]
    ktest.ktest_larft.argtypes = [C.c_int, C.c_int, dp, dp, dp]
    ktest.ktest_qp3.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_double, C.c_int, C.c_int]
    ktest.ktest_refl.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int]
    ktest.ktest_bwd.argtypes = [C.c_int, dp, C.c_int, dp, dp, dp, C.c_int, C.c_int]
    ktest.ktest_trsm.argtypes = [C.c_int, C.c_int, dp, dp]
    ktest.ktest_copy.argtypes = [C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, C.c_int]
    return ktest


def block_T(oracle, V, tau):
    """Oracle larft per NB block -> nb x roundup(k) layout of the device."""
    m, k = V.shape
    nblk = (k + NB - 1) // NB
    out = np.zeros((NB, NB * nblk), order="F")
    for b in range(nblk):
        jb = b * NB
        w = min(NB, k - jb)
        Vb = np.asfortranarray(V[jb:, jb:jb + w])
        out[:w, jb:jb + w] = oracle.larft(Vb, tau[jb:jb + w])
    return out


@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("shape", [(1, 1), (33, 70), (100, 7), (64, 64)])
def test_copy(K, shape, trans):
    rng = np.random.default_rng(1)
    r, c = shape
    lds = r + 3
    src = np.asfortranarray(rng.standard_normal((lds, c)))
    dr, dc = (c, r) if trans else (r, c)
    ldd = dr + 5
    dst = np.asfortranarray(np.full((ldd, dc), 9.0))
    assert K.ktest_copy(r, c, d(src), lds, d(dst), ldd, trans) == 0
    ref = src[:r, :c].T if trans else src[:r, :c]
    assert np.array_equal(dst[:dr, :dc], ref)
    assert np.all(dst[dr:, :] == 9.0)


def struct_mask(M, s):
    M = M.copy()
    if s == 1:
        M = np.tril(M, -1) + np.eye(*M.shape)
    elif s == 2:
        M = np.triu(M)
    elif s == 3:
        M = np.tril(M)
    return M


@pytest.mark.parametrize("ta,tb,sa,sb", [(0, 0, 0, 0), (1, 0, 1, 0), (1, 0, 2, 0), (0, 0, 1, 0), (1, 0, 1, 1),
                                         (0, 1, 0, 0), (1, 1, 3, 2), (0, 0, 0, 2)])
@pytest.mark.parametrize("mnk", [(1, 1, 1), (65, 17, 130), (200, 129, 33), (7, 300, 64)])
def test_gemm(K, mnk, ta, tb, sa, sb):
    m, n, k = mnk
    rng = np.random.default_rng(m * 7 + n)
    Ar, Ac = (k, m) if ta else (m, k)
    Br, Bc = (n, k) if tb else (k, n)
    lda, ldb, ldc = Ar + 2, Br + 1, m + 3
    A = np.asfortranarray(rng.standard_normal((lda, Ac)))
    B = np.asfortranarray(rng.standard_normal((ldb, Bc)))
    Cm = np.asfortranarray(rng.standard_normal((ldc, n)))
    alpha, beta = -1.5, 0.5
    Am = struct_mask(A[:Ar, :Ac], sa)
    Bm = struct_mask(B[:Br, :Bc], sb)
    opA = Am.T if ta else Am
    opB = Bm.T if tb else Bm
    ref = alpha * opA @ opB + beta * Cm[:m, :n]
    flags = ta | (tb << 1) | (sa << 4) | (sb << 6)
    assert K.ktest_gemm(m, n, k, d(A), lda, Ac, d(B), ldb, Bc, d(Cm), ldc, flags, alpha, beta) == 0
    assert rel(Cm[:m, :n], ref) < 1e-13


# FP32 matrix-core path (the mixed form's factorization), every tile shape:
# 32 x 128 (m <= 32), 128 x 32 (n <= 32) and 64 x 64, against an FP64 product
# of the same FP32 inputs
@pytest.mark.parametrize("ta,tb,sa,sb", [(0, 0, 0, 0), (1, 0, 1, 0), (0, 1, 0, 0), (1, 1, 3, 2)])
@pytest.mark.parametrize("mnk", [(1, 1, 1), (32, 300, 700), (65, 17, 130), (200, 129, 33), (7, 300, 64)])
def test_gemm_f32(K, mnk, ta, tb, sa, sb):
    m, n, k = mnk
    rng = np.random.default_rng(m * 5 + n)
    Ar, Ac = (k, m) if ta else (m, k)
    Br, Bc = (n, k) if tb else (k, n)
    lda, ldb, ldc = Ar + 1, Br + 3, m + 2
    A = np.asfortranarray(rng.standard_normal((lda, Ac)).astype(np.float32))
    B = np.asfortranarray(rng.standard_normal((ldb, Bc)).astype(np.float32))
    Cm = np.asfortranarray(rng.standard_normal((ldc, n)).astype(np.float32))
    alpha, beta = -1.0, 1.0
    Am = struct_mask(A[:Ar, :Ac].astype(np.float64), sa)
    Bm = struct_mask(B[:Br, :Bc].astype(np.float64), sb)
    ref = alpha * (Am.T if ta else Am) @ (Bm.T if tb else Bm) + beta * Cm[:m, :n].astype(np.float64)
    flags = ta | (tb << 1) | (sa << 4) | (sb << 6)
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    assert K.ktest_gemm_f32(m, n, k, f(A), lda, Ac, f(B), ldb, Bc, f(Cm), ldc, flags, alpha, beta) == 0
    assert rel(Cm[:m, :n].astype(np.float64), ref) < 2e-6 * np.sqrt(k)


# split-K: K cut into ksplit chunks (some empty when k < 16 * ksplit), partial
# tiles summed in chunk order by the reduce pass, which applies beta
@pytest.mark.parametrize("mnk,ks,ta,sa", [((32, 300, 5000), 9, 1, 1), ((32, 70, 1000), 4, 1, 0),
                                          ((100, 33, 37), 5, 0, 0), ((64, 64, 20), 3, 1, 2), ((7, 5, 3), 2, 0, 3)])
def test_gemm_split(K, mnk, ks, ta, sa):
    m, n, k = mnk
    rng = np.random.default_rng(m * 11 + n + ks)
    Ar, Ac = (k, m) if ta else (m, k)
    lda, ldb, ldc = Ar + 1, k + 2, m + 1
    A = np.asfortranarray(rng.standard_normal((lda, Ac)))
    B = np.asfortranarray(rng.standard_normal((ldb, n)))
    Cm = np.asfortranarray(rng.standard_normal((ldc, n)))
    alpha, beta = 0.75, -1.0
    Am = struct_mask(A[:Ar, :Ac], sa)
    opA = Am.T if ta else Am
    ref = alpha * opA @ B[:k, :n] + beta * Cm[:m, :n]
    flags = ta | (sa << 4)
    before = Cm.copy()
    assert K.ktest_gemm_split(m, n, k, d(A), lda, Ac, d(B), ldb, n, d(Cm), ldc, flags, alpha, beta, ks) == 0
    assert rel(Cm[:m, :n], ref) < 1e-13
    assert np.array_equal(Cm[m:, :], before[m:, :])  # padding rows untouched


# sigma_min of R from a QR of random matrices with prescribed singular values
# (graded, clustered at the bottom, and a 1e8 condition number) against
# numpy's SVD
@pytest.mark.parametrize("k,kind", [(1, "graded"), (5, "graded"), (33, "graded"), (64, "cluster"), (100, "graded"),
                                    (300, "cluster"), (650, "graded"), (120, "ill")])
def test_smin(K, k, kind):
    rng = np.random.default_rng(k)
    m = k + 17
    if kind == "graded":
        s = np.logspace(0, -3, k)
    elif kind == "cluster":
        s = np.concatenate([np.logspace(1, 0, k - 3), [1e-2 * (1 + 1e-9), 1e-2 * (1 + 2e-9), 1e-2]]) if k > 3 else np.ones(k)
    else:
        s = np.logspace(0, -8, k)
    U = np.linalg.qr(rng.standard_normal((m, k)))[0]
    Vt = np.linalg.qr(rng.standard_normal((k, k)))[0]
    A = (U * s) @ Vt
    R = np.linalg.qr(A, mode="r")
    ld = k + 3
    Rp = np.zeros((ld, k), order="F")
    Rp[:k, :] = R
    out = np.zeros(1)
    assert K.ktest_smin(k, d(Rp), ld, d(out)) == 0
    ref = np.linalg.svd(R, compute_uv=False).min()
    # a cluster of width 2e-9 at the bottom is resolved to within its width
    tol = {"graded": 1e-12, "cluster": 1e-8, "ill": 1e-7}[kind]
    assert abs(out[0] - ref) <= tol * ref, (out[0], ref)


# panel steps run on ceil((m - jb) / 256) workgroups: (513, 65), (2000, 70)
# use the flat inter-workgroup barrier, (5000, 40) the two-level one
@pytest.mark.parametrize("mk", [(1, 1), (5, 3), (40, 40), (97, 33), (150, 64), (300, 100), (64, 80), (513, 65),
                                (2000, 70), (5000, 40)])
def test_blocked_qr(K, oracle, mk):
    m, k = mk
    rng = np.random.default_rng(m + 1000 * k)
    A = np.asfortranarray(rng.standard_normal((m, k)))
    if k > 1:
        A[:, k // 2] = 0.0  # a zero column gives tau = 0 (dlarfg special case)
    nc = 37
    Cm = np.asfortranarray(rng.standard_normal((m, nc)))
    oracle.set_backend("own")
    QR_ref, tau_ref = oracle.geqrf(A)
    kk = min(m, k)
    nblk = (k + NB - 1) // NB
    P = A.copy(order="F")
    tau = np.zeros(k)
    Tb = np.zeros((NB, NB * nblk), order="F")
    C2 = Cm.copy(order="F")
    assert K.ktest_qr(m, k, d(P), d(tau), d(Tb), nc, d(C2)) == 0
    scale = np.linalg.norm(A)
    assert np.linalg.norm(np.triu(P[:kk]) - np.triu(QR_ref[:kk])) / scale < 1e-12
    assert np.linalg.norm(np.tril(P, -1) - np.tril(QR_ref, -1)) / max(np.linalg.norm(np.tril(QR_ref, -1)), 1) < 1e-12
    assert np.allclose(tau[:kk], tau_ref, rtol=1e-12, atol=1e-14)
    Tref = block_T(oracle, np.asfortranarray(QR_ref[:, :kk]), tau_ref)
    for b in range((kk + NB - 1) // NB):
        w = min(NB, kk - b * NB)
        blk = slice(b * NB, b * NB + w)
        assert rel(np.triu(Tb[:w, blk]), Tref[:w, blk]) < 1e-12
        # the trailing update takes T_b as a dense factor (W = V_b T_b^T):
        # its strict lower part must be stored as exact zeros
        assert not np.any(np.tril(Tb[:w, blk], -1))
    Cref = oracle.larfb(np.asfortranarray(QR_ref[:, :kk]), oracle.larft(np.asfortranarray(QR_ref[:, :kk]), tau_ref),
                        Cm, "T")
    assert rel(C2, Cref) < 1e-12


@pytest.mark.parametrize("mr", [(10, 3), (64, 64), (200, 70)])
def test_larft(K, oracle, mr):
    m, r = mr
    rng = np.random.default_rng(m)
    oracle.set_backend("own")
    QR, tau = oracle.geqrf(rng.standard_normal((m, r)))
    V = np.asfortranarray(QR)
    nblk = (r + NB - 1) // NB
    Tb = np.zeros((NB, NB * nblk), order="F")
    assert K.ktest_larft(m, r, d(V), d(np.ascontiguousarray(tau)), d(Tb)) == 0
    Tref = block_T(oracle, V, tau)
    for b in range(nblk):
        w = min(NB, r - b * NB)
        blk = slice(b * NB, b * NB + w)
        assert rel(np.triu(Tb[:w, blk]), Tref[:w, blk]) < 1e-12


# form 0 = streaming kernel over nwg workgroups; form 1 / 2 = register kernel
# with 256 / 1024-thread workgroups and 32 rows per lane (lane group size and
# workgroup count follow from k and N: with 256 threads (120, 2000) runs on 32
# workgroups with 4 lanes per column and (300, 1500) on 94 with 16, both
# through the two-level barrier; with 1024 threads on 8 and 24); form
# 10 * rc + 1 / 2 = the same with rc rows per lane (16 to 28: other lane
# groups and padding, e.g. (300, 1500) at rc = 20 on 16 lanes, 320 rows; 40
# and 48, 256 threads only: fewer lanes, e.g. (300, 1500) on 8)
@pytest.mark.parametrize("form,nwg", [(0, 1), (0, 3), (1, 0), (2, 0), (161, 0), (201, 0), (241, 0), (282, 0), (401, 0),
                                      (481, 0)])
@pytest.mark.parametrize("kN,true_rank,tol", [((20, 100), 6, 1e-3), ((64, 300), 20, 1e-6), ((50, 40), 50, 1e-10),
                                              ((30, 30), 10, 1e-2), ((1, 5), 1, 1e-3), ((120, 2000), 90, 1e-8),
                                              ((300, 1500), 150, 1e-7)])
def test_qp3(K, oracle, kN, true_rank, tol, form, nwg):
    k, N = kN
    rng = np.random.default_rng(k * N)
    r = min(true_rank, k, N)
    U = np.linalg.qr(rng.standard_normal((k, k)))[0][:, :r]
    s = np.logspace(0, -12, r)
    W = rng.standard_normal((r, N))
    Cmat = np.asfortranarray((U * s) @ W + 1e-15 * rng.standard_normal((k, N)))
    oracle.set_backend("own")
    Rref, jpvt, tau_ref = oracle.geqp3(Cmat)
    kk = min(k, N)
    rank_ref = oracle.choose_rank(np.diag(Rref)[:kk], tol)
    Ct = np.asfortranarray(Cmat.T.copy())
    V = np.zeros((k, k), order="F")
    tau = np.zeros(k)
    rank = K.ktest_qp3(k, N, d(Ct), d(V), d(tau), tol, min(nwg, N), form)
    assert rank == rank_ref
    Crank = Ct[:, :rank].T  # rank x N, original column order
    # reference Crank = triu(R[:rank]) P^T
    pos = np.empty(N, dtype=int)
    pos[jpvt] = np.arange(N)
    ref = np.zeros((rank, N))
    for j in range(N):
        p = pos[j]
        ref[:, j] = np.where(np.arange(rank) <= p, Rref[:rank, p], 0.0)
    assert rel(Crank, ref) < 1e-10
    # a reflector is determined to ~eps * |r_00| / |r_ii|: compare the well-conditioned ones
    dd = np.abs(np.diag(Rref))[:rank]
    well = dd / dd[0] > 1e-6
    assert np.allclose(tau[:rank][well], tau_ref[:rank][well], rtol=1e-9, atol=1e-14)
    Vref = np.tril(Rref[:, :rank], -1) + np.eye(k, rank)
    assert rel((np.tril(V[:, :rank], -1) + np.eye(k, rank))[:, well], Vref[:, well]) < 1e-9


@pytest.mark.parametrize("form", [0, 1, 2])
@pytest.mark.parametrize("trans", [1, 0])
# form 0: m >= 2048 runs the row-sliced multi-workgroup form (4 workgroups);
# form 1 (register form, 7 row waves): ceil(m / 448) workgroups, so 600 rows
# use 2 and 5100 rows 12; form 2: 1 row wave for 40 rows, 2 for 64 < m <= 128,
# 7 beyond; 97 segments exceed one wave of the segment-table scan
@pytest.mark.parametrize("mk,nseg", [((40, 40), 1), ((100, 60), 4), ((150, 70), 3), ((448, 64), 2), ((600, 100), 4),
                                     ((20000, 40), 2), ((5100, 300), 3), ((97 * 7, 90), 97)])
def test_refl(K, oracle, mk, nseg, trans, form):
    m, k = mk
    assert m % nseg == 0  # the test splits x into nseg equal segments
    rng = np.random.default_rng(m + k)
    oracle.set_backend("own")
    QR, tau = oracle.geqrf(rng.standard_normal((m, k)))
    V = np.asfortranarray(QR)
    Tb = block_T(oracle, V, tau)
    x = rng.standard_normal(m)
    Tfull = oracle.larft(V, tau)
    ref = oracle.larfb(V, Tfull, x.reshape(m, 1), "T" if trans else "N")[:, 0]
    xx = x.copy()
    assert K.ktest_refl(m, k, d(V), d(Tb), d(xx), trans, nseg, 5, form) == 0
    assert rel(xx, ref) < 1e-12


@pytest.mark.parametrize("trans", [0, 1])
# form 0: k >= 512 runs the multi-workgroup form (4 workgroups over whole
# 32-row blocks); form 1 (latency form) has no R12 products: s = 0 only
@pytest.mark.parametrize("ks,form", [((1, 0), 0), ((50, 30), 0), ((300, 10), 0), ((700, 40), 0), ((1000, 0), 0),
                                     ((1, 0), 1), ((31, 0), 1), ((50, 0), 1), ((300, 0), 1), ((733, 0), 1),
                                     ((1000, 0), 1), ((2100, 0), 1)])
def test_bwd(K, ks, trans, form):
    k, s = ks
    rng = np.random.default_rng(k)
    R = np.asfortranarray(np.triu(rng.standard_normal((k, k))) / np.sqrt(k) + 4 * np.eye(k))
    A1 = np.asfortranarray(rng.standard_normal((k, max(s, 1))))
    xs = rng.standard_normal(k)
    x1 = rng.standard_normal(max(s, 1))
    xs2, x12 = xs.copy(), x1.copy()
    assert K.ktest_bwd(k, d(R), s, d(A1), d(xs2), d(x12), trans, form) == 0
    if trans == 0:
        ref = np.linalg.solve(R, xs - (A1[:, :s] @ x1[:s] if s else 0))
        assert rel(xs2, ref) < 1e-12
    else:
        ref = np.linalg.solve(R.T, xs)
        assert rel(xs2, ref) < 1e-12
        if s:
            assert rel(x12[:s], x1[:s] - A1[:, :s].T @ ref) < 1e-12


@pytest.mark.parametrize("rk", [(1, 1), (300, 45), (70, 200)])
def test_trsm(K, rk):
    rows, k = rk
    rng = np.random.default_rng(rows)
    # well conditioned upper factor (random triangular matrices are not)
    R = np.asfortranarray(np.triu(rng.standard_normal((k, k))) / np.sqrt(k) + 4 * np.eye(k))
    B = np.asfortranarray(rng.standard_normal((rows, k)))
    ref = np.linalg.solve(R.T, B.T).T
    BB = B.copy(order="F")
    assert K.ktest_trsm(rows, k, d(R), d(BB)) == 0
    assert rel(BB, ref) < 1e-12
