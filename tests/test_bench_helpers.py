"""bench.py's helpers on the CPU: the committed traffic figure it puts in
`roofline.traffic`, the solver rate, and the red-black SOR baseline (the
reference's own build when it is there and the state cannot converge before
the cap, else the oracle's restatement)."""
import os
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


def test_cpu_rb_baseline_kinds():
    src = A.Case(A.load(A.oracle_lib_path("rb_fast")), **bench.cavity_params(32)).run(2)
    port = bench.cpu_rb(src, 32, 10, 1000, iters=12)
    assert port["kind"] == "port" and port["iterations"] <= 12 and port["solver_mlups"] > 0
    if os.path.exists(A.ref_lib_path("rb_fast")):
        ref = bench.cpu_rb(src, 32, 1000, 1000, iters=12)
        assert ref["kind"] == "reference" and ref["iterations"] == 12 and ref["solver_mlups"] > 0
    src.close()
