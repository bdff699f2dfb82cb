"""Profile JSON (the reference's Profile, profile.h:13-106) and the launch-form
autotuner, on the GPU path."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from spaqr_public_amd import gen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_profile_json(gpu_lib):
    import spaqr_public_amd as sq
    r = sq.run_flow(gen.config_matrix("advdiff2d_64"), 64, 2, 1e-2)
    t = r["tree"]
    p = t.profile()
    assert p["nlevels"] == r["nlevels"] and p["rows"] == p["cols"] == 4096
    assert len(p["levels"]) == r["nlevels"]
    sp = t.sparsified()
    got = np.array([[i["ilvl"], i["lvl"], i["sep"], i["cols_before"], i["cols_after"]] for i in p["interfaces"]])
    assert np.array_equal(got, sp)
    # an interface's RRQR matrix has its neighbours' columns (at least one)
    assert all(i["neighbors"] > 0 for i in p["interfaces"] if i["cols_after"] < i["cols_before"])
    assert sum(lv["sparsified"] for lv in p["levels"]) == len(sp)
    for m in p["merged"]:
        assert m["rows"] >= m["cols"] >= 0
        if m["cols"]:
            assert abs(m["aspect_ratio"] - m["rows"] / m["cols"]) < 1e-12
    tops = [lv["top_cols"] for lv in p["levels"]]
    assert tops == r["levels"][:, 7].astype(int).tolist()


TUNE_SCRIPT = r"""
import json, sys
sys.path.insert(0, {root!r})
import numpy as np
import spaqr_public_amd as sq
from spaqr_public_amd import gen
r = sq.run_flow(gen.config_matrix("advdiff3d_16"), 16, 3, 1e-2)
print(json.dumps({{"iters": r["iters"], "final": r["final_res"], "ranks": r["tree"].sparsified()[:, 4].tolist()}}))
"""


def run_tuned(path, tune):
    env = dict(os.environ, SPAQR_TUNE=str(tune), SPAQR_TUNE_FILE=path)
    out = subprocess.run([sys.executable, "-c", TUNE_SCRIPT.format(root=ROOT)], env=env, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_autotuner_records_and_reuses(gpu_lib, tmp_path):
    path = str(tmp_path / "tune.json")
    first = run_tuned(path, 1)
    rec = json.load(open(path))
    keys = [e["key"] for e in rec["entries"]]
    assert keys and all(k.startswith("qp3r/f64/") for k in keys)
    for e in rec["entries"]:
        assert e["choice"] in (16, 20, 24, 28, 32, 40, 48, 64, 80, 100)
        assert str(e["choice"]) in e["ms"] and len(e["ms"]) >= 5
        assert e["ms"][str(e["choice"])] == min(e["ms"].values())
    again = run_tuned(path, 0)  # reads the file, tunes nothing
    assert json.load(open(path)) == rec
    assert first["final"] <= 1e-12 and again["final"] <= 1e-12
    assert abs(first["iters"] - again["iters"]) <= 1
    agree = np.mean(np.array(first["ranks"]) == np.array(again["ranks"]))
    assert agree >= 0.95
