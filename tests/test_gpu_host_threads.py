"""The host planner's worker pool changes nothing that the factorization
computes.  A tree planned with SPAQR_HOST_THREADS=1 (every planning loop on the
calling thread, in a subprocess) matches one planned on the pool (this process,
8 threads) in:
- the cluster table;
- the sparsified interfaces and their ranks;
- the fill counters and the top separator;
- the Krylov iteration count;
- the one-shot residual, to rounding (the solve's R12 products accumulate with
  atomics, so the apply is not bitwise reproducible).
The pool's chunking depends only on loop lengths, and every table is written at
fixed offsets, so the descriptor tables, and with them the kernels' inputs, are
the same for any thread count."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import spaqr_public_amd as sq
from spaqr_public_amd import gen
A = gen.config_matrix(sys.argv[2])
r = sq.run_flow(A, int(sys.argv[3]), int(sys.argv[4]), float(sys.argv[5]))
t = r["tree"]
out = dict(table=t.cluster_table().tolist(), sparsified=t.sparsified().tolist(), nnzR=int(r["nnzR"]),
           nnzH=int(r["nnzH"]), top=int(r["top_size"]), iters=int(r["iters"]), oneshot=float(r["oneshot_res"]))
r["M"].close()
t.close()
print("RESULT " + json.dumps(out), flush=True)
"""


def test_serial_planning_matches_pool(gpu_lib):
    import spaqr_public_amd as sq
    from spaqr_public_amd import gen
    name, n, dim, tol = "advdiff3d_32", 32, 3, 1e-2
    print(f"{name}: planning on the pool (this process)", flush=True)
    r = sq.run_flow(gen.config_matrix(name), n, dim, tol)
    t = r["tree"]
    pool = dict(table=t.cluster_table(), sparsified=t.sparsified(), nnzR=int(r["nnzR"]), nnzH=int(r["nnzH"]),
                top=int(r["top_size"]), iters=int(r["iters"]), oneshot=float(r["oneshot_res"]))
    r["M"].close()
    t.close()
    print(f"{name}: planning on one thread (subprocess, SPAQR_HOST_THREADS=1)", flush=True)
    env = dict(os.environ, SPAQR_HOST_THREADS="1")
    p = subprocess.run([sys.executable, "-c", CHILD, REPO, name, str(n), str(dim), str(tol)], env=env,
                       capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr[-2000:]
    line = [x for x in p.stdout.splitlines() if x.startswith("RESULT ")][-1]
    one = json.loads(line[len("RESULT "):])
    assert np.array_equal(np.asarray(one["table"]), pool["table"])
    assert np.array_equal(np.asarray(one["sparsified"]), pool["sparsified"])
    assert (one["nnzR"], one["nnzH"], one["top"], one["iters"]) == (pool["nnzR"], pool["nnzH"], pool["top"],
                                                                     pool["iters"])
    assert abs(one["oneshot"] - pool["oneshot"]) <= 1e-9 * pool["oneshot"]
    print(f"{name}: identical tree, {len(pool['sparsified'])} sparsified interfaces, {pool['iters']} iterations, "
          f"one-shot {pool['oneshot']:.6e} / {one['oneshot']:.6e}")
