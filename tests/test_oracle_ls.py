"""Least-squares parts of the oracle pinned on the CPU (configs[3] family,
A = [L_h; mu D_h]).

* sparsify_extra (tree.cpp:1260-1307): the extra rows E of an interface are
  replaced by the R of their RRQR, truncated at tol, so E^T E ~ R^T R.  At
  tol 1e-6 and 1e-10 no interface of ls_lapgrad_32 is sparsified by sparsifyD
  while every sparsify_extra call drops rows, so the normal-equation apply
  (W^T W)^{-1} A^T A - I isolates the extra-row compression: it must be of the
  order of tol.  Keeping the untransformed rows, dropping the wrong ones or
  forgetting the pivot permutation leaves an O(1) error; a no-op compression
  fails the "rows dropped" check.
* form_rmap (tree.cpp:103-163): a row joins the cluster holding the largest
  sum of squares of its entries, ties to the first cluster in column order.
  Planted rows whose winner is known by construction (a 3 in one cluster
  against three 2s in another: sum of squares 9 < 12, largest entry 3 > 2;
  and a 2 against a 2) pin the rule against the plausible mistakes (largest
  entry, smallest sum, last of equal sums).  The planted rows use columns of
  one existing row of L_h, so the A^T A pattern and hence the partition are
  unchanged.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from spaqr_public_amd import gen


def ls_consistency(tree, As, solve_nrml):
    N = As.shape[1]
    AtA = (As.T @ As).toarray()
    Minv = np.stack([solve_nrml(AtA[:, j]) for j in range(N)], axis=1)
    return np.linalg.norm(Minv - np.eye(N), 2)


@pytest.mark.parametrize("tol", [1e-6, 1e-10])
def test_sparsify_extra_consistent(oracle, tol):
    oracle.set_backend("own")
    r = oracle.run_reference_flow(gen.config_matrix("ls_lapgrad_32"), 32, 2, tol, res=1e-10)
    t = r["tree"]
    ex, sp_ = t.extra_sparsified(), t.sparsified()
    assert len(ex) >= 10
    assert np.all(ex[:, 4] < ex[:, 3])           # every call drops rows
    assert np.all(sp_[:, 4] == sp_[:, 3])        # sparsifyD keeps every column at this tol
    err = ls_consistency(t, r["A"], t.solve_nrml)
    assert err <= 100 * tol, err


def cluster_maps(cperm, rperm, table):
    """Original column / row -> index of its level-0 cluster."""
    lvl0 = table[table[:, 0] == 0]
    pos_c = np.empty(len(cperm), dtype=np.int64)
    pos_c[cperm] = np.arange(len(cperm))
    pos_r = np.empty(len(rperm), dtype=np.int64)
    pos_r[rperm] = np.arange(len(rperm))
    col_of = np.full(len(cperm), -1)
    row_of = np.full(len(rperm), -1)
    for q, rec in enumerate(lvl0):
        cs, cn, rs, rn = rec[7], rec[8], rec[9], rec[10]
        col_of[(pos_c >= cs) & (pos_c < cs + cn)] = q
        row_of[(pos_r >= rs) & (pos_r < rs + rn)] = q
    return col_of, row_of, lvl0


def planted_rows(A, col_of, lvl0):
    """Rows appended to A: (row, expected level-0 cluster), built from one
    row of L_h (its 5 columns) whose columns lie in two clusters, at least
    three of them in one."""
    N = A.shape[1]
    L = A[:N].tocsr()
    for i in range(N):
        cols = L.indices[L.indptr[i]:L.indptr[i + 1]]
        cl = col_of[cols]
        u, cnt = np.unique(cl, return_counts=True)
        if len(u) != 2 or cnt.max() < 3:
            continue
        big, small = (u[0], u[1]) if cnt[0] >= cnt[1] else (u[1], u[0])
        cb, cs = cols[cl == big], cols[cl == small]
        rows, expect = [], []
        # sum of squares decides: 3^2 = 9 in `small` against 3 * 2^2 = 12 in `big`
        r = np.zeros(N)
        r[cs[0]] = 3.0
        r[cb[:3]] = 2.0
        rows.append(r)
        expect.append(big)
        # equal sums: the cluster whose columns come first in the ordering wins
        r = np.zeros(N)
        r[cs[0]] = 2.0
        r[cb[0]] = -2.0
        rows.append(r)
        first = big if lvl0[big][7] < lvl0[small][7] else small
        expect.append(first)
        return sp.csc_matrix(np.array(rows)), expect
    raise AssertionError("no row of L_h straddles two clusters")


def partition_only(make_tree, A, n):
    t = make_tree(A)
    t.partition(A)
    cperm, rperm = t.perms()
    return t, cperm, rperm, t.cluster_table()


def oracle_tree(oracle, nlevels):
    def make(A):
        t = oracle.Tree(nlevels, 0)
        t.configure(1e-2, scale=1, order=1.0, square=False)
        t.set_tensor_coords(16, 2)
        return t
    return make


def run_form_rmap_planted(make_tree):
    base = gen.config_matrix("ls_lapgrad_16")
    N = base.shape[1]
    _, cperm, rperm, table = partition_only(make_tree, base, 16)
    col_of, _, lvl0 = cluster_maps(cperm, rperm, table)
    extra, expect = planted_rows(base, col_of, lvl0)
    A = sp.vstack([base, extra]).tocsc()
    A.sort_indices()
    _, cperm2, rperm2, table2 = partition_only(make_tree, A, 16)
    # the column partition is unchanged by the planted rows
    assert np.array_equal(cperm2, cperm)
    colpart = [0, 1, 2, 3, 4, 5, 6, 7, 8, 11]
    assert np.array_equal(table2[:, colpart], table[:, colpart])
    _, row_of, _ = cluster_maps(cperm2, rperm2, table2)
    m0 = base.shape[0]
    got = [row_of[m0 + q] for q in range(len(expect))]
    assert got == expect, (got, expect)


def test_form_rmap_planted(oracle):
    from pyoracle import default_levels
    oracle.set_backend("own")
    run_form_rmap_planted(oracle_tree(oracle, default_levels(256)))
