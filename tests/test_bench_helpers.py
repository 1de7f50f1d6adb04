"""bench.py's helpers on the CPU: the committed traffic figure it puts in
`roofline.traffic`, the solver rate, and the red-black SOR baseline (the
reference's own build when it is there and the state cannot converge before
the cap, else the oracle's restatement)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
import nusif_api as A  # noqa: E402


def test_measured_traffic_reads_the_committed_profile():
    t = bench.measured_traffic("gmg_smooth32_l0", 256)
    assert t is not None
    # between the algorithmic bytes (12 B x 256^3) and twice that
    assert 12 * 256 ** 3 <= t <= 2 * 12 * 256 ** 3
    assert bench.measured_traffic("gmg_smooth32_l0", 128) is None
    assert bench.measured_traffic("no_such_kernel", 256) is None


def test_solver_rate():
    res = {"solve": {"ms": 20.0, "solves": 10, "cycles": 30.0}}
    r = bench.solver_rate(res, 1e6)
    assert r["solver_mlups"] == pytest.approx(30.0 * 1e6 / 0.02 * 1e-6)
    assert r["solve_ms_per_step"] == pytest.approx(2.0)
    assert bench.solver_rate({"solve": {"ms": 0.0, "solves": 0, "cycles": 0.0}}, 1e6) == {}


CPU_RB = """
import json, sys
sys.path.insert(0, %r)
import bench, nusif_api as A
src = A.Case(A.load(A.oracle_lib_path("rb_fast")), **bench.cavity_params(32)).run(2)
out = [bench.cpu_rb(src, 32, 10, 1000, iters=12)]
if %r:
    out.append(bench.cpu_rb(src, 32, 1000, 1000, iters=12))
src.close()
print(json.dumps(out))
"""


def test_cpu_rb_baseline_kinds():
    """In a child process: the -ffast-math builds this loads set the
    process's flush-to-zero mode, which the oracle checks must not inherit."""
    has_ref = os.path.exists(A.ref_lib_path("rb_fast"))
    r = subprocess.run([sys.executable, "-c", CPU_RB % (ROOT, has_ref)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    port = out[0]
    assert port["kind"] == "port" and port["iterations"] <= 12 and port["solver_mlups"] > 0
    if has_ref:
        ref = out[1]
        assert ref["kind"] == "reference" and ref["iterations"] == 12 and ref["solver_mlups"] > 0


CPU_GMG = """
import json, sys
sys.path.insert(0, %r)
import bench, nusif_api as A
src = A.Case(A.load(A.oracle_lib_path("rb_fast")), **bench.cavity_params(32)).run(2)
out = bench.cpu_baseline(src, 32, budget_s=0.05)
src.close()
print(json.dumps(out))
"""


def test_cpu_gmg_baseline():
    """The headline CPU baseline (in a child process, as above): the oracle's
    FP64 multigrid with the GPU modes' smoother omega, the host CPU named"""
    r = subprocess.run([sys.executable, "-c", CPU_GMG % ROOT], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["kind"] == "port" and out["cores"] == 1 and out["value"] > 0
    assert "omega 1.3" in out["sample"] and out["host_cpu"] in out["sample"]
    assert out["cycles_per_step"] >= 1
