"""Writes tests/golden/oracle64_<workload>.npz: the oracle (the CPU
restatement of the reference, OpenBLAS LAPACK) on the full-size 3-D
workloads at t = 1e-2, the flow of spaQR.cpp:85-320 (column scaling, tree,
one-shot solve with random(n, 2021), GMRES to 1e-12):

  advdiff3d_64    : -Lap u + v.grad u, 64^3, upwind v = (100, 100, 100)
  hicontrast3d_64 : -div(a grad u), 64^3, a piecewise constant on 8^3 blocks

Stored per workload: SHA-256 of cperm, rperm and the cluster table (the tree
must be identical), the sparsified-interface table (ilvl, lvl, sep, cols
before, after), the top-separator size, the printed nnz counters, the
one-shot relative residual, the GMRES iteration count and final residual;
for hicontrast3d_64 also the one-shot solution x = M^{-1} b in float32, so
the mixed-precision GPU form can report its error against the FP64 oracle.

The GPU tests (tests/test_gpu_fullsize.py) read these files; the oracle does
not run on the GPU box.  About 7 minutes on 8 cores.  Run from the
repository root: python tests/golden/make_golden_64.py
"""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

WORKLOADS = [("advdiff3d_64", 64, 3, 1e-2, False), ("hicontrast3d_64", 64, 3, 1e-2, True)]
# cluster-table columns (l, self, left, right, cstart, order) that do not depend on ranks
STRUCT_COLS = [0, 1, 2, 3, 4, 5, 6, 7, 11]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int32).tobytes()).hexdigest()


def main():
    import pyoracle as po
    from spaqr_public_amd import gen
    threads = min(os.cpu_count() or 1, 16)
    backend = po.set_backend("openblas", threads=threads)
    for name, n, dim, tol, keep_x in WORKLOADS:
        t0 = time.time()
        r = po.run_reference_flow(gen.config_matrix(name), n, dim, tol)
        t = r["tree"]
        cperm, rperm = t.perms()
        data = dict(
            workload=np.array(name), tol=np.array(tol), backend=np.array(backend),
            cperm_sha=np.array(sha(cperm)), rperm_sha=np.array(sha(rperm)),
            cluster_table_sha=np.array(sha(t.cluster_table())),
            # the tree's structure: level, ids, column start and creation
            # order of every cluster (the size columns of merged clusters
            # follow the ranks of the levels below)
            cluster_struct_sha=np.array(sha(t.cluster_table()[:, STRUCT_COLS])),
            sparsified=t.sparsified().astype(np.int32),
            top_size=np.array(r["top_size"]), nlevels=np.array(r["nlevels"]),
            nnz=np.array([r["nnzA"], r["nnzR"], r["nnzH"], r["nnzQ"]], dtype=np.uint64),
            oneshot_res=np.array(r["oneshot_res"]), iters=np.array(r["iters"]), final_res=np.array(r["final_res"]),
        )
        if keep_x:
            data["x_oneshot_f32"] = t.solve(r["b"]).astype(np.float32)
        out = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"oracle64_{name}.npz")
        np.savez_compressed(out, **data)
        print(f"{name}: one-shot {r['oneshot_res']:.3e}, {r['iters']} its, final {r['final_res']:.2e}, "
              f"{len(data['sparsified'])} sparsified, top {r['top_size']} ({time.time() - t0:.0f} s) -> {out}",
              flush=True)


if __name__ == "__main__":
    main()
