"""Writes tests/golden/lapack_golden.npz: input/output vectors of the LAPACK
routines the reference calls through its wrappers (util.cpp:306-477, 785-793),
produced by scipy.linalg.lapack (scipy's bundled OpenBLAS/LAPACK).

The oracle's own restatement of these routines is checked against the file
(tests/test_oracle_lapack.py), so the fixtures pin the Householder sign
convention (dlarfg), the tau values, the pivot order of dgeqp3 and the
orthogonal applies.  Run from the repository root: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy
import scipy.linalg.lapack as la

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lapack_golden.npz")


def main():
    rng = np.random.default_rng(20211)
    data = {"scipy_version": np.array(scipy.__version__)}
    # dgeqrf on tall, square, wide and with an exactly zero column (tau = 0)
    shapes = [(6, 4), (9, 9), (5, 8), (40, 17)]
    for i, (m, n) in enumerate(shapes):
        A = rng.standard_normal((m, n))
        if i == 3:
            A[:, 5] = 0.0
        qr, tau, work, info = la.dgeqrf(A)
        assert info == 0
        data[f"geqrf{i}_A"] = A
        data[f"geqrf{i}_qr"] = qr
        data[f"geqrf{i}_tau"] = tau
    # dgeqp3 on wide matrices with decaying spectrum (the sparsify C blocks)
    for i, (m, n, r) in enumerate([(8, 20, 5), (12, 30, 12), (30, 45, 9)]):
        U = np.linalg.qr(rng.standard_normal((m, m)))[0][:, :r]
        W = rng.standard_normal((r, n))
        A = (U * np.logspace(0, -9, r)) @ W
        qr, jpvt, tau, work, info = la.dgeqp3(A)
        assert info == 0
        data[f"geqp3{i}_A"] = A
        data[f"geqp3{i}_qr"] = qr
        data[f"geqp3{i}_jpvt"] = jpvt - 1
        data[f"geqp3{i}_tau"] = tau
    # dormqr: Q^T C and Q C with the reflectors of geqrf0
    qr, tau = data["geqrf0_qr"], data["geqrf0_tau"]
    C = rng.standard_normal((6, 3))
    lwork = 64
    cq, work, info = la.dormqr("L", "T", qr, tau, C, lwork)
    assert info == 0
    data["ormqr_C"] = C
    data["ormqr_QtC"] = cq
    cq, work, info = la.dormqr("L", "N", qr, tau, C, lwork)
    data["ormqr_QC"] = cq
    # dtrtrs: upper solves (trsv / trsm semantics)
    U = np.triu(rng.standard_normal((7, 7))) + 3 * np.eye(7)
    b = rng.standard_normal((7, 2))
    x, info = la.dtrtrs(U, b, lower=0, trans=0)
    data["trtrs_U"] = U
    data["trtrs_b"] = b
    data["trtrs_x"] = x
    xt, info = la.dtrtrs(U, b, lower=0, trans=1)
    data["trtrs_xt"] = xt
    np.savez_compressed(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
