"""The C ABI library (libspaqr.so) loads without a GPU and exports every
function include/spaqr.h declares; host-only helpers match the oracle."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "spaqr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(spaqr_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_tree_surface():
    names = declared_functions()
    for n in ("spaqr_tree_create", "spaqr_partition", "spaqr_assemble", "spaqr_factorize", "spaqr_solve",
              "spaqr_solve_nrml", "spaqr_gmres", "spaqr_tree_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol(gpu_lib):
    missing = [n for n in declared_functions() if not hasattr(gpu_lib, n)]
    assert not missing, missing


def test_python_binding_lists_every_symbol():
    from spaqr_public_amd import core
    assert sorted(core.SYMBOLS) == declared_functions()


def test_host_helpers_match_oracle(gpu_lib, oracle):
    import spaqr_public_amd as sq
    from spaqr_public_amd import gen
    assert np.array_equal(sq.random_vec(1000, 2021), oracle.random_vec(1000, 2021))
    A = gen.config_matrix("advdiff2d_32")
    a1, d1 = sq.colscale(A)
    a2, d2 = oracle.colscale(A)
    assert np.array_equal(a1.data, a2.data) and np.array_equal(d1, d2)


def test_missing_library_fails_loudly(tmp_path):
    from spaqr_public_amd import core
    try:
        core.load_library.__wrapped__ if hasattr(core.load_library, "__wrapped__") else None
        saved = core._lib
        core._lib = None
        raised = False
        try:
            core.load_library(str(tmp_path / "nope.so"))
        except core.SpaqrError:
            raised = True
        assert raised
    finally:
        core._lib = saved
