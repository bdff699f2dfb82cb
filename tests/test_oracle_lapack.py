"""The oracle's LAPACK restatement against scipy-LAPACK golden vectors
(tests/golden/lapack_golden.npz) and against the mathematics (orthogonality,
compact-WY identity, pivoted-QR invariants).  CPU only."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "lapack_golden.npz")


@pytest.fixture(scope="module")
def G():
    return np.load(GOLD, allow_pickle=False)


@pytest.fixture(params=["own", "openblas"])
def be(request, oracle):
    oracle.set_backend(request.param)
    yield request.param
    oracle.set_backend("own")


def close(a, b, tol=1e-12):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) <= tol * max(np.linalg.norm(np.asarray(b)), 1.0)


@pytest.mark.parametrize("i", range(4))
def test_geqrf_golden(oracle, G, be, i):
    qr, tau = oracle.geqrf(G[f"geqrf{i}_A"])
    assert close(qr, G[f"geqrf{i}_qr"])
    assert close(tau, G[f"geqrf{i}_tau"])


def crank(qr, jpvt, rows):
    """triu(R[:rows]) P^T: R rows in original column order (what sparsifyD
    writes back, tree.cpp:1082-1084)."""
    n = qr.shape[1]
    pos = np.empty(n, dtype=int)
    pos[jpvt] = np.arange(n)
    R = np.triu(qr[:rows])
    return R[:, pos]


@pytest.mark.parametrize("i", range(3))
def test_geqp3_golden(oracle, G, be, i):
    A = G[f"geqp3{i}_A"]
    qr, jpvt, tau = oracle.geqp3(A)
    ref_qr, ref_p, ref_tau = G[f"geqp3{i}_qr"], G[f"geqp3{i}_jpvt"], G[f"geqp3{i}_tau"]
    d = np.abs(np.diag(ref_qr))
    well = int(np.sum(d / d[0] > 1e-6))  # rows whose reflector is well determined
    assert np.array_equal(jpvt[:well], ref_p[:well])
    assert close(crank(qr, jpvt, well), crank(ref_qr, ref_p, well), 1e-9)
    assert close(tau[:well], ref_tau[:well], 1e-9)
    # the rank decision the sparsifier takes is identical
    for tol in (1e-1, 1e-3, 1e-6):
        assert oracle.choose_rank(np.diag(qr), tol) == oracle.choose_rank(np.diag(ref_qr), tol)


def test_ormqr_golden(oracle, G, be):
    qr, tau = G["geqrf0_qr"], G["geqrf0_tau"]
    assert close(oracle.ormqr_left(qr, tau, G["ormqr_C"], "T"), G["ormqr_QtC"])
    assert close(oracle.ormqr_left(qr, tau, G["ormqr_C"], "N"), G["ormqr_QC"])


def test_trsv_trsm_golden(oracle, G, be):
    U, b = G["trtrs_U"], G["trtrs_b"]
    assert close(oracle.trsv_upper(U, b[:, 0], "N"), G["trtrs_x"][:, 0])
    assert close(oracle.trsv_upper(U, b[:, 0], "T"), G["trtrs_xt"][:, 0])
    # B U^{-1} = (U^{-T} B^T)^T
    assert close(oracle.trsm_right_upper(U, b.T), G["trtrs_xt"].T)


@pytest.mark.parametrize("mk", [(30, 30), (70, 33), (12, 5)])
def test_larft_compact_wy_identity(oracle, be, mk):
    """H_1 ... H_k == I - V T V^T (the defining identity of dlarft 'F','C')."""
    rng = np.random.default_rng(mk[0])
    m, k = mk
    qr, tau = oracle.geqrf(rng.standard_normal((m, k)))
    V = np.tril(qr, -1) + np.eye(m, k)
    T = oracle.larft(qr, tau)
    H = np.eye(m)
    for i in range(k):
        v = V[:, i:i + 1]
        H = H @ (np.eye(m) - tau[i] * v @ v.T)
    assert close(np.eye(m) - V @ T @ V.T, H, 1e-13)
    assert np.allclose(np.tril(T, -1), 0)
    # larfb 'T' applies H^T, 'N' applies H
    Cm = rng.standard_normal((m, 4))
    assert close(oracle.larfb(qr, T, Cm, "T"), H.T @ Cm)
    assert close(oracle.larfb(qr, T, Cm, "N"), H @ Cm)


def test_geqp3_invariants(oracle, be):
    """Pivoted QR: A P = Q R, |r_ii| non-increasing, |r_ii| = norm of the
    pivot column's remaining part (the largest remaining norm)."""
    rng = np.random.default_rng(7)
    A = rng.standard_normal((25, 60)) * np.logspace(0, -6, 60)
    qr, jpvt, tau = oracle.geqp3(A)
    R = np.triu(qr)
    Q = oracle.ormqr_left(qr[:, :25], tau, np.eye(25), "N")
    assert close(Q @ R, A[:, jpvt], 1e-12)
    d = np.abs(np.diag(R))
    assert np.all(d[1:] <= d[:-1] * (1 + 1e-12))
    assert sorted(jpvt.tolist()) == list(range(60))


def test_choose_rank(oracle):
    """util.cpp:795-816: first index with |s_i|/|s_0| < tol, at least 1."""
    s = np.array([4.0, 2.0, 0.5, 0.01, 0.001])
    assert oracle.choose_rank(s, 0.1) == 3
    assert oracle.choose_rank(s, 0.0) == 5
    assert oracle.choose_rank(s, 1.0) == 0
    assert oracle.choose_rank(s, 1e-9) == 5
    assert oracle.choose_rank(s[:1], 0.5) == 1
    assert oracle.choose_rank(np.zeros(3), 0.1) == 1  # 0/0 is not >= tol


def test_planted_sign_error_is_caught(oracle, G):
    """A plausible mistake (flipping the dlarfg sign convention) must fail the
    golden check: the check has discriminating power."""
    oracle.set_backend("own")
    qr, tau = oracle.geqrf(G["geqrf1_A"])
    assert not close(-qr, G["geqrf1_qr"])


@pytest.mark.parametrize("m,n,cond", [(7, 7, 1e2), (40, 25, 1e6), (90, 60, 1e3), (5, 12, 1e1)])
def test_singular_values(oracle, m, n, cond):
    """The oracle's singular values (util.cpp:767-782, one-sided Jacobi) match
    LAPACK's SVD (numpy) on matrices with prescribed spectra."""
    rng = np.random.default_rng(m * n)
    k = min(m, n)
    s = np.logspace(0, -np.log10(cond), k)
    U = np.linalg.qr(rng.standard_normal((m, k)))[0]
    V = np.linalg.qr(rng.standard_normal((n, k)))[0]
    A = (U * s) @ V.T
    S = oracle.singular_values(A)
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.all(np.diff(S) <= 0)
    assert np.max(np.abs(S - ref)) <= 1e-13 * ref[0] * max(m, n)
