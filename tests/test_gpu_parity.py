"""End-to-end parity of the GPU spaQR against the oracle (the CPU restatement
of the reference, itself pinned to the reference's published log).

Bar (north star): identical cluster tree and ordering; one-shot residual
within 2x of the oracle's; GMRES iterations within +-1; final relative
residual <= 1e-12 (FP64 form).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run_both(oracle, A, n, dim, tol, **kw):
    import spaqr_public_amd as sq
    oracle.set_backend("own")
    ref = oracle.run_reference_flow(A, n, dim, tol, **kw)
    gpu = sq.run_flow(A, n, dim, tol, **kw)
    return gpu, ref


def _check_tree(gpu, ref):
    cg, rg = gpu["tree"].perms()
    co, ro = ref["tree"].perms()
    assert np.array_equal(cg, co)
    assert np.array_equal(rg, ro)
    assert np.array_equal(gpu["tree"].cluster_table(), ref["tree"].cluster_table())


def test_readme_known_answer(oracle, gpu_lib):
    """configs[0]-shape run of README.md:48-84 (central convection): every
    printed integer of the reference log must come out of the GPU path."""
    from spaqr_public_amd import gen
    A = gen.config_matrix("advdiff2d_128_central")
    gpu, ref = _run_both(oracle, A, 128, 2, 1e-3)
    _check_tree(gpu, ref)
    tops = gpu["levels"][:, 7].astype(int).tolist()
    assert tops == [254, 216, 163, 132, 99, 63, 63, 63]          # README.md:63-70
    assert gpu["top_size"] == 63                                  # README.md:73
    assert (gpu["nnzA"], gpu["nnzR"], gpu["nnzH"], gpu["nnzQ"]) == (81408, 1594464, 2697960, 201985)
    assert gpu["oneshot_res"] <= 2 * 1.338e-3                      # README.md:78
    assert abs(gpu["iters"] - 5) <= 1                              # README.md:81
    assert gpu["final_res"] <= 1e-12
    assert np.array_equal(gpu["tree"].sparsified(), ref["tree"].sparsified())


@pytest.mark.parametrize("name,n,dim", [("advdiff2d_32", 32, 2), ("advdiff3d_12", 12, 3)])
def test_exact_factorization(oracle, gpu_lib, name, n, dim):
    """tol = 0 makes spaQR an exact sparse QR (spaQR.cpp:150): one solve is a
    direct solve."""
    import scipy.sparse.linalg as spl
    from spaqr_public_amd import gen
    A = gen.config_matrix(name)
    gpu, ref = _run_both(oracle, A, n, dim, 0.0)
    _check_tree(gpu, ref)
    assert gpu["oneshot_res"] < 1e-13
    x = gpu["tree"].solve(gpu["b"])
    xs = spl.spsolve(gpu["A"].tocsc(), gpu["b"])
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-12


@pytest.mark.parametrize("name,n,dim,tol", [
    ("advdiff2d_128", 128, 2, 1e-3),
    ("advdiff2d_64", 64, 2, 1e-1),
    ("advdiff3d_16", 16, 3, 1e-2),
    ("advdiff3d_32", 32, 3, 1e-2),
])
def test_parity_configs(oracle, gpu_lib, name, n, dim, tol):
    from spaqr_public_amd import gen
    A = gen.config_matrix(name)
    gpu, ref = _run_both(oracle, A, n, dim, tol)
    _check_tree(gpu, ref)
    assert gpu["oneshot_res"] <= 2.0 * ref["oneshot_res"]
    assert abs(gpu["iters"] - ref["iters"]) <= 1
    assert gpu["final_res"] <= 1e-12
    sg, so = gpu["tree"].sparsified(), ref["tree"].sparsified()
    assert sg.shape == so.shape
    assert np.mean(sg[:, 4] == so[:, 4]) >= 0.95  # ranks: near-threshold ties may flip a few


def test_solve_device_pointer_alias(oracle, gpu_lib):
    """b may alias x in the device apply (the reference takes b by value)."""
    import spaqr_public_amd as sq
    from spaqr_public_amd import gen, hip
    A = gen.config_matrix("advdiff2d_32")
    r = sq.run_flow(A, 32, 2, 1e-2)
    t = r["tree"]
    b = r["b"]
    x_host = t.solve(b)
    buf = hip.DeviceArray.from_host(b)
    t.solve_device(buf.ptr, buf.ptr)
    out = buf.to_host()
    assert np.allclose(out, x_host, rtol=0, atol=1e-13 * np.abs(x_host).max())


@pytest.mark.parametrize("name,n,dim,tol", [
    ("advdiff2d_128_central", 128, 2, 1e-3),
    ("advdiff3d_16", 16, 3, 1e-2),
    ("advdiff3d_32", 32, 3, 1e-2),
])
def test_order_1_5(oracle, gpu_lib, name, n, dim, tol):
    """--order 1.5 (sparsifyD_imp, tree.cpp:1117-1257): fine clusters keep the
    coupling E2 = (Q^T Apm_top)[rank:, :] and a tau = 0 QR op."""
    from spaqr_public_amd import gen
    A = gen.config_matrix(name)
    gpu, ref = _run_both(oracle, A, n, dim, tol, order=1.5)
    _check_tree(gpu, ref)
    assert gpu["oneshot_res"] <= 2.0 * ref["oneshot_res"]
    assert abs(gpu["iters"] - ref["iters"]) <= 1
    assert gpu["final_res"] <= 1e-12
    sg, so = gpu["tree"].sparsified(), ref["tree"].sparsified()
    assert sg.shape == so.shape
    assert np.mean(sg[:, 4] == so[:, 4]) >= 0.95
    if name == "advdiff2d_128_central":
        assert np.array_equal(sg, so)
        assert gpu["oneshot_res"] < 0.8 * 1.338e-3  # below the order-1 README value


@pytest.mark.parametrize("n,tol", [(16, 0.0), (32, 0.0), (32, 1e-2), (64, 1e-2)])
def test_least_squares(oracle, gpu_lib, n, tol):
    """Least squares [L_h; mu D_h] (configs[3] family): R-only elimination,
    reassign_rows, sparsify_extra, the normal-equation apply and CGLS on the
    GPU against the oracle (tree.cpp:319-411, 1260-1307, 1471-1517;
    is.cpp:275-346)."""
    from spaqr_public_amd import gen
    A = gen.config_matrix(f"ls_lapgrad_{n}")
    gpu, ref = _run_both(oracle, A, n, 2, tol, res=1e-10)
    _check_tree(gpu, ref)
    if tol == 0:
        As, b = gpu["A"], gpu["b"]
        x = gpu["tree"].solve_nrml(As.T @ b)
        xe = np.linalg.lstsq(As.toarray(), b, rcond=None)[0]
        assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-11
    assert gpu["oneshot_res"] <= 2.0 * ref["oneshot_res"] + 1e-14
    assert abs(gpu["iters"] - ref["iters"]) <= 1
    assert gpu["final_res"] <= 1e-10
    sg, so = gpu["tree"].sparsified(), ref["tree"].sparsified()
    assert sg.shape == so.shape
    if len(so):
        assert np.mean(sg[:, 4] == so[:, 4]) >= 0.95


@pytest.mark.parametrize("name,n,dim,tol", [
    ("advdiff2d_128", 128, 2, 1e-3),
    ("advdiff3d_16", 16, 3, 1e-2),
    ("hicontrast3d_16", 16, 3, 1e-2),
    ("hicontrast3d_32", 32, 3, 1e-2),
])
def test_mixed_precision(oracle, gpu_lib, name, n, dim, tol):
    """Mixed form: FP32 factorization, FP64 GMRES as iterative refinement.
    Same tree as the FP64 oracle; the one-shot solve stays within 2x of the
    oracle's residual and close to its solution; refinement reaches 1e-12."""
    import spaqr_public_amd as sq
    from spaqr_public_amd import gen
    A = gen.config_matrix(name)
    oracle.set_backend("own")
    ref = oracle.run_reference_flow(A, n, dim, tol)
    gpu = sq.run_flow(A, n, dim, tol, precision=sq.FP32)
    _check_tree(gpu, ref)
    x32 = gpu["tree"].solve(gpu["b"])
    x64 = ref["tree"].solve(ref["b"])
    err = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    print(f"{name}: one-shot {gpu['oneshot_res']:.3e} (oracle {ref['oneshot_res']:.3e}), "
          f"|x32 - x64|/|x64| = {err:.2e}, refinement its {gpu['iters']} (oracle GMRES {ref['iters']}), "
          f"final {gpu['final_res']:.2e}")
    assert gpu["oneshot_res"] <= 2.0 * ref["oneshot_res"]
    assert err < 1e-3
    assert gpu["final_res"] <= 1e-12
    assert gpu["iters"] <= 3 * ref["iters"] + 6


@pytest.mark.parametrize("n,tol", [(32, 1e-2), (64, 1e-2)])
def test_least_squares_mixed(oracle, gpu_lib, n, tol):
    """Least squares with FP32 factors and FP64 CGLS (configs[3] form)."""
    import spaqr_public_amd as sq
    from spaqr_public_amd import gen
    A = gen.config_matrix(f"ls_lapgrad_{n}")
    oracle.set_backend("own")
    ref = oracle.run_reference_flow(A, n, 2, tol, res=1e-10)
    gpu = sq.run_flow(A, n, 2, tol, res=1e-10, precision=sq.FP32)
    _check_tree(gpu, ref)
    print(f"ls {n}: one-shot {gpu['oneshot_res']:.3e} (oracle {ref['oneshot_res']:.3e}), "
          f"CGLS its {gpu['iters']} (oracle {ref['iters']}), final {gpu['final_res']:.2e}")
    assert gpu["oneshot_res"] <= 2.0 * ref["oneshot_res"] + 1e-6
    assert gpu["final_res"] <= 1e-10
    assert gpu["iters"] <= 2 * ref["iters"] + 4


@pytest.mark.parametrize("name,n,dim,tol", [
    ("advdiff2d_64", 64, 2, 1e-2),
    ("advdiff2d_128", 128, 2, 1e-3),
    ("advdiff3d_16", 16, 3, 1e-2),
    ("advdiff3d_32", 32, 3, 1e-2),
])
def test_scale_0(oracle, gpu_lib, name, n, dim, tol):
    """--scale 0 (sparsify, tree.cpp:887-989): sigma_min normalisation,
    App Q, split_edges and the inline fine-cluster elimination on the GPU.

    This branch is rounding-sensitive in the reference itself: the oracle run
    with its own BLAS and with OpenBLAS already disagrees on 7 of 463
    (advdiff2d_128) and 4 of 117 (advdiff3d_16) ranks, where scale 1 agrees
    exactly.  So the tree is compared structurally (permutations, cluster
    ids, offsets, creation order; sizes follow the ranks) and the ranks to
    90 %, next to the north-star bar on residual and iterations."""
    from spaqr_public_amd import gen
    A = gen.config_matrix(name)
    gpu, ref = _run_both(oracle, A, n, dim, tol, scale=0)
    cg, rg = gpu["tree"].perms()
    co, ro = ref["tree"].perms()
    assert np.array_equal(cg, co) and np.array_equal(rg, ro)
    tg, to = gpu["tree"].cluster_table(), ref["tree"].cluster_table()
    cols = [0, 1, 2, 3, 4, 5, 6, 7, 9, 11]
    assert np.array_equal(tg[:, cols], to[:, cols])
    print(f"{name}: one-shot {gpu['oneshot_res']:.3e} (oracle {ref['oneshot_res']:.3e}), its {gpu['iters']} "
          f"(oracle {ref['iters']}), final {gpu['final_res']:.2e}")
    assert gpu["oneshot_res"] <= 2.0 * ref["oneshot_res"]
    assert abs(gpu["iters"] - ref["iters"]) <= 1
    assert gpu["final_res"] <= 1e-12
    sg, so = gpu["tree"].sparsified(), ref["tree"].sparsified()
    assert sg.shape == so.shape
    assert np.mean(sg[:, 4] == so[:, 4]) >= 0.90
    if name == "advdiff2d_64":
        assert np.array_equal(sg, so)


def test_scale_0_consistent():
    """GPU --scale 0 at tol 1e-6: M^{-1} A - I is of the order of tol while
    interfaces are compressed (the oracle test of the same name pins the
    restatement)."""
    import spaqr_public_amd as sq
    from spaqr_public_amd import gen
    A = gen.config_matrix("advdiff2d_32")
    r = sq.run_flow(A, 32, 2, 1e-6, scale=0)
    sp = r["tree"].sparsified()
    assert (sp[:, 4] < sp[:, 3]).sum() >= 3
    N = A.shape[0]
    E = np.eye(N)
    Minv = np.stack([r["tree"].solve(E[:, j]) for j in range(N)], axis=1)
    assert np.linalg.norm(Minv @ r["A"].toarray() - E) <= 1e-4
