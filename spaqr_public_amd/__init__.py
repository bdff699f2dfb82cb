"""spaqr_public_amd — MI355X-native spaQR (sparsified hierarchical QR).

The hot path of the reference (Tree::factorize per-level phases and
Tree::solve inside GMRES) runs as batched gfx950 HIP kernels behind the C ABI
of include/spaqr.h; this package is the thin Python binding of that library.
The differentiable solve (torch.autograd) lives in spaqr_public_amd.autograd.
"""
from .core import (FP32, FP64, Matrix, SpaqrError, Tree, cgls, cgls_device, colscale, csc_outer_device,
                   default_levels, gmres, gmres_device, gmres_transpose_device, load_library, random_vec, row_parts,
                   run_flow)

__all__ = ["FP32", "FP64", "Matrix", "SpaqrError", "Tree", "cgls", "cgls_device", "colscale", "csc_outer_device",
           "default_levels", "gmres", "gmres_device", "gmres_transpose_device", "load_library", "random_vec",
           "row_parts", "run_flow"]
