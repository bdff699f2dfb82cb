"""Non-geometric partition on the CPU.

* The oracle's labelling from a row partition (HypergraphPartition,
  partition.cpp:450-499, and getInterfacesUsingA_HUND, partition.cpp:582-623)
  against vectors derived by hand from those lines on small patterns: 1-D
  chains at two and three levels, and a pattern built so that one separator
  meets leaves of two parts on one side (highest common ancestor), a
  lower-level separator through a row, a side with no lower-level column
  (falls back to self) and a first row in the odd part (index pi / 2).
  Sides follow the part of the row a neighbour is met through, not the
  neighbour's own part; the chain at three levels has a level-2 separator
  whose left side is the leaf next to it, not the level-1 separator.
* Invariants of the labelling on random patterns and partitions: a leaf's
  rows share one part; a level-l separator's rows share their part above
  level l and split at level l; interfaces lie below their separator.
* The library's row partitioner (spaqr_row_parts, in place of PaToH; a
  binary-only library, so its partitions themselves are parity unpinned):
  parts in range, balanced halves at every level, deterministic, and on
  grids no more cut columns than recursive coordinate bisection (far fewer
  than a random balanced partition on the least-squares matrix).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from spaqr_public_amd import gen


def chain(n):
    return sp.diags([np.ones(n - 1), 2 * np.ones(n), np.ones(n - 1)], [-1, 0, 1]).tocsc()


def labels_of(rows):
    return [tuple(r) for r in rows]


def test_hund_chain_two_levels(oracle):
    parts = [0, 0, 0, 0, 1, 1, 1]
    got = labels_of(oracle.hypergraph_labels(chain(7), 2, parts))
    leaf0, leaf1 = (0, 0, 0, 0, 0, 0), (0, 1, 0, 1, 0, 1)
    assert got == [leaf0, leaf0, leaf0, (1, 0, 0, 0, 0, 1), (1, 0, 0, 0, 0, 1), leaf1, leaf1]


def test_hund_chain_three_levels(oracle):
    parts = [0] * 4 + [1] * 4 + [2] * 4 + [3] * 4
    got = labels_of(oracle.hypergraph_labels(chain(16), 3, parts))
    leaf = [(0, p, 0, p, 0, p) for p in range(4)]
    want = [leaf[0]] * 3 + [(1, 0, 0, 0, 0, 1)] * 2 + [leaf[1]] * 2 + [(2, 0, 0, 1, 0, 2)] * 2 \
        + [leaf[2]] * 2 + [(1, 1, 0, 2, 0, 3)] * 2 + [leaf[3]] * 3
    assert got == want


def test_hund_planted_pattern(oracle):
    cols = [[0], [1], [2], [3], [0, 1, 2], [2, 3], [0, 4], [3, 5]]
    r = [i for c in cols for i in c]
    c = [j for j, cc in enumerate(cols) for _ in cc]
    A = sp.csc_matrix((np.ones(len(r)), (r, c)), shape=(6, 8))
    parts = [0, 1, 2, 3, 1, 2]
    got = labels_of(oracle.hypergraph_labels(A, 3, parts))
    assert got == [(0, 0, 0, 0, 0, 0), (0, 1, 0, 1, 0, 1), (0, 2, 0, 2, 0, 2), (0, 3, 0, 3, 0, 3),
                   (2, 0, 1, 0, 1, 1),   # left: highest common of (0,0), (1,0), (0,1); right: of (0,2), (1,1)
                   (1, 1, 0, 2, 0, 3),
                   (1, 0, 0, 0, 1, 0),   # no lower-level column on the right: self
                   (1, 1, 1, 1, 0, 3)]   # first row in part 3: sep 3 / 2; empty left: self


def check_invariants(A, nlevels, parts, lab):
    A = sp.csc_matrix(A)
    parts = np.asarray(parts)
    for j in range(A.shape[1]):
        rows = A.indices[A.indptr[j]:A.indptr[j + 1]]
        lvl, sep, ll, ls, rl, rs = lab[j]
        if lvl == 0:
            assert np.all(parts[rows] == sep) and (ll, ls, rl, rs) == (0, sep, 0, sep)
            continue
        assert len(set(parts[rows] >> lvl)) == 1 and (parts[rows[0]] >> lvl) == sep
        assert len(set(parts[rows] >> (lvl - 1))) == 2
        for il, isp in ((ll, ls), (rl, rs)):
            assert (il, isp) == (lvl, sep) or (il < lvl and (isp >> (lvl - il)) == sep)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_hund_invariants_random(oracle, seed):
    rng = np.random.default_rng(seed)
    m, n, nlevels = 60, 50, 4
    A = sp.random(m, n, density=0.06, random_state=seed, format="csc")
    A = (A + sp.csc_matrix((np.ones(n), (rng.integers(0, m, n), np.arange(n))), shape=(m, n))).tocsc()
    parts = rng.integers(0, 2 ** (nlevels - 1), m)
    check_invariants(A, nlevels, parts, oracle.hypergraph_labels(A, nlevels, parts))


def cut_columns(A, parts):
    A = sp.csc_matrix(A)
    p = np.asarray(parts)
    return sum(len(set(p[A.indices[A.indptr[j]:A.indptr[j + 1]]])) > 1 for j in range(A.shape[1]))


def coordinate_parts(n, dim, nlevels):
    """Recursive coordinate bisection of an n^dim grid's rows (the longest
    extent halved, lower half first): a yardstick for the cut."""
    coords = np.stack(np.unravel_index(np.arange(n ** dim), (n,) * dim), 1)
    part = np.zeros(n ** dim, dtype=np.int64)
    for d in range(nlevels - 1):
        new = part.copy()
        for g in range(2 ** d):
            s = np.where(part == g)[0]
            c = coords[s]
            order = s[np.argsort(c[:, np.argmax(c.max(0) - c.min(0))], kind="stable")]
            new[order[:(len(s) + 1) // 2]] = 2 * g
            new[order[(len(s) + 1) // 2:]] = 2 * g + 1
        part = new
    return part


@pytest.mark.parametrize("name,nlevels,grid", [("advdiff2d_32", 5, (32, 2)), ("advdiff3d_16", 6, (16, 3)),
                                               ("advdiff3d_32", 9, (32, 3)), ("ls_lapgrad_32", 5, None)])
def test_row_parts_properties(gpu_lib, oracle, name, nlevels, grid):
    import spaqr_public_amd as sq
    A = gen.config_matrix(name)
    if A.shape[0] < A.shape[1]:
        A = A.T.tocsc()
    p = sq.row_parts(A, nlevels)
    assert np.array_equal(p, sq.row_parts(A, nlevels))
    assert p.min() >= 0 and p.max() < 2 ** (nlevels - 1)
    m = A.shape[0]
    for lvl in range(nlevels - 1):  # each side within 3 % of its part over one half, per bisection
        sizes = np.bincount(p >> lvl, minlength=2 ** (nlevels - 1 - lvl))
        assert sizes.min() > 0
        assert sizes.max() <= m / len(sizes) * 1.06 ** (nlevels - 1 - lvl) + nlevels
    cut = cut_columns(A, p)
    if grid:  # no worse than bisecting the grid by coordinates
        ref = cut_columns(A, coordinate_parts(*grid, nlevels))
        print(f"{name}: {cut} cut columns (coordinate bisection: {ref})")
        assert cut <= 1.1 * ref
    else:
        rng = np.random.default_rng(0)
        rand = rng.permutation(np.repeat(np.arange(2 ** (nlevels - 1)), -(-m // 2 ** (nlevels - 1)))[:m])
        print(f"{name}: {cut} cut columns (random balanced partition: {cut_columns(A, rand)})")
        assert cut < 0.5 * cut_columns(A, rand)
    check_invariants(A, nlevels, p, oracle.hypergraph_labels(A, nlevels, p))


def test_row_parts_top_separator_on_grid(gpu_lib):
    """A 64 x 64 grid: the top bisection cuts about one grid line, not a
    diagonal band (breadth-first layers from a corner are diagonals; the
    refinement straightens them)."""
    import spaqr_public_amd as sq
    A = gen.config_matrix("advdiff2d_64")
    p = sq.row_parts(A, 2)
    cut = cut_columns(A, p)
    print(f"64x64 grid, one bisection: {cut} cut columns")
    assert cut <= 4 * 64


def test_row_parts_rejects_bad_levels(gpu_lib):
    import spaqr_public_amd as sq
    with pytest.raises(sq.SpaqrError):
        sq.row_parts(chain(8), 0)


@pytest.mark.parametrize("name", ["advdiff2d_32", "advdiff3d_16", "ls_lapgrad_16"])
def test_nongeometric_exact_at_tol0(gpu_lib, oracle, name):
    """At tol 0 (no sparsification) the non-geometric tree is an exact sparse
    QR: the one-shot solve equals spsolve / lstsq.  A row sparsity that drops
    a cluster coupling numerically to c through a neighbour (the
    column-sparsity rule of the geometric path, or no numeric test at all,
    tree.cpp:846-866) leaves fill out and fails this; the numeric test
    |A(n,c)^T A(n,nin)| > 1e-14 only drops products that are exactly zero."""
    import scipy.sparse.linalg as spl
    import spaqr_public_amd as sq
    A = gen.config_matrix(name)
    At = A.T.tocsc() if A.shape[0] < A.shape[1] else A
    nl = sq.default_levels(At.shape[1])
    parts = sq.row_parts(At, nl)
    oracle.set_backend("own")
    r = oracle.run_reference_flow(A, None, None, 0.0, nlevels=nl, parts=parts, iters=1)
    As, b = r["A"], r["b"]
    if As.shape[0] == As.shape[1]:
        x = r["tree"].solve(b)
        xe = spl.spsolve(As.tocsc(), b)
    else:
        x = r["tree"].solve_nrml(As.T @ b)
        xe = np.linalg.lstsq(As.toarray(), b, rcond=None)[0]
    err = np.linalg.norm(x - xe) / np.linalg.norm(xe)
    print(f"{name}: non-geometric tree at tol 0, |x - x_exact| / |x_exact| = {err:.2e}")
    assert err < 1e-10
