"""The GPU geometric multigrid (FP64 and FP32-in-FP64 defect correction)
against the oracle's restatement, through the C ABI."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import make_golden as G
import nusif_api as A
from conftest import lib

pytestmark = pytest.mark.gpu
DP = A._DP


def oracle_solve(p, rhs, prm, opts, first=None):
    """The oracle's gmg solve; `first` = cycles before the first stop test
    (default: the mg_inner option, 1 when absent)."""
    prm = dict(A.DEFAULTS, **prm)
    rel = C.c_double()
    if first is None:
        first = max(1, int(opts.get("mg_inner", 1)))
    cyc = lib("orc_gmg").orcGmgSolveGrouped(
        prm["imax"], prm["jmax"], prm["kmax"], prm["xlength"] / prm["imax"],
        prm["ylength"] / prm["jmax"], prm["zlength"] / prm["kmax"], p.ctypes.data_as(DP),
        rhs.ctypes.data_as(DP), prm["levels"], prm["presmooth"], prm["postsmooth"],
        float(opts.get("mg_omega", 1.0)),
        float(opts["mg_tol"]), int(opts.get("mg_maxcycles", 100)),
        int(opts.get("mg_project", 0)), 1 if opts.get("precision") == "mixed" else 0,
        first, C.byref(rel))
    return cyc, rel.value


CASES = [
    ((32, 32, 32), 4, "fp64"), ((32, 32, 32), 4, "mixed"),
    ((48, 24, 16), 3, "fp64"), ((48, 24, 16), 3, "mixed"),
    ((64, 64, 64), 6, "mixed"),
    # finest extent >= 128 and nk >= 32: the fused one-launch sweep (partial k chunk)
    ((128, 12, 72), 3, "fp64"), ((128, 12, 72), 3, "mixed"),
]


@pytest.mark.parametrize("shape,levels,precision", CASES)
def test_gpu_gmg_solve_matches_oracle(shape, levels, precision, gpu_ok):
    print("gmg", shape, levels, precision, flush=True)
    ni, nj, nk = shape
    prm = dict(G.KAT_PARAMS, imax=ni, jmax=nj, kmax=nk, xlength=ni / 32.0, ylength=nj / 32.0,
               zlength=nk / 32.0, levels=levels, presmooth=2, postsmooth=2)
    opts = {"solver": "gmg", "precision": precision, "mg_tol": 1e-10, "mg_project": 1}
    c = A.Case(lib("gpu_rb"), options=opts, **prm)
    f = G.random_fields(c.shape, 31)
    p0 = np.zeros(c.shape)
    c.set_fields(p=p0, rhs=f["rhs"])
    res = c.lib.solve(C.byref(c.s), c.d.p, c.d.rhs)
    out = c.fields()
    cyc_g = c.lib.nusifSolverIterations(C.byref(c.s))
    p = p0.copy()
    rhs = np.ascontiguousarray(f["rhs"]).copy()
    cyc_o, rel_o = oracle_solve(p, rhs, prm, opts)
    assert cyc_g == cyc_o
    assert res == pytest.approx(rel_o, rel=1e-6)
    assert res <= 1e-10
    a = out["p"][1:-1, 1:-1, 1:-1]
    b = p[1:-1, 1:-1, 1:-1]
    tol = 1e-12 if precision == "fp64" else 1e-9
    assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(b))
    c.close()


@pytest.mark.parametrize("shape,levels,precision,inner", [((64, 64, 64), 6, "fp64", 3),
                                                         ((64, 64, 64), 6, "mixed", 3),
                                                         ((128, 12, 72), 3, "mixed", 2)])
def test_gpu_gmg_grouped_cycles_match_oracle(shape, levels, precision, inner, gpu_ok):
    """mg_inner k: the first stop test after k cycles (mixed: k FP32 cycles on
    one defect, continuing e): same cycle count and p as the oracle's grouped
    loop."""
    print("gmg grouped", shape, levels, precision, inner, flush=True)
    ni, nj, nk = shape
    prm = dict(G.KAT_PARAMS, imax=ni, jmax=nj, kmax=nk, xlength=ni / 32.0, ylength=nj / 32.0,
               zlength=nk / 32.0, levels=levels, presmooth=2, postsmooth=2)
    opts = {"solver": "gmg", "precision": precision, "mg_tol": 1e-10, "mg_project": 1,
            "mg_inner": inner}
    c = A.Case(lib("gpu_rb"), options=opts, **prm)
    f = G.random_fields(c.shape, 33)
    p0 = np.zeros(c.shape)
    c.set_fields(p=p0, rhs=f["rhs"])
    res = c.lib.solve(C.byref(c.s), c.d.p, c.d.rhs)
    out = c.fields()
    cyc_g = c.lib.nusifSolverIterations(C.byref(c.s))
    p = p0.copy()
    rhs = np.ascontiguousarray(f["rhs"]).copy()
    cyc_o, rel_o = oracle_solve(p, rhs, prm, opts)
    assert cyc_g == cyc_o
    assert res == pytest.approx(rel_o, rel=1e-6)
    a = out["p"][1:-1, 1:-1, 1:-1]
    b = p[1:-1, 1:-1, 1:-1]
    tol = 1e-12 if precision == "fp64" else 1e-9
    assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(b))
    c.close()


@pytest.mark.parametrize("precision,inner", [("fp64", 1), ("mixed", 1), ("mixed", 0)])
def test_gpu_gmg_cavity_steps(precision, inner, gpu_ok):
    """Whole steps of the 32^3 cavity with the gmg solver: GPU vs the oracle's
    step functions + oracle gmg; per-field error stated against FP64. With
    mg_inner 0 the library sizes each solve's first group from the previous
    solve: the oracle is given the group the library reports, the cycle
    counts agree step by step, and each group follows from the previous
    solve (its count when the group did not converge, else the same group or
    one less; in mixed precision a group that stopped reducing at the FP32
    bound of one correction caps the later groups below itself)."""
    print("gmg cavity", precision, inner, flush=True)
    prm = dict(G.TRAJ["c1_rb"][1], re=1000.0, levels=4, presmooth=2, postsmooth=2)
    opts = {"solver": "gmg", "precision": precision, "mg_tol": 1e-8, "mg_inner": inner}
    g = A.Case(lib("gpu_rb"), options=opts, **prm)
    o = A.Case(lib("orc_rb"), **prm)
    prev = None  # (group, cycles) of the previous solve
    cap = None   # mixed precision: the group cap the library set
    for step in range(6):
        g.step()
        d = o.d
        if d.tau > 0:
            o.lib.computeTimestep(C.byref(d))
        for fn in ("setBoundaryConditions", "setSpecialBoundaryCondition", "computeFG",
                   "computeRHS"):
            o.call(fn)
        if step == 0:
            o.call("normalizePressure")
        n = int(np.prod(o.shape))
        pv = np.ctypeslib.as_array(d.p, shape=(n,))
        rv = np.ctypeslib.as_array(d.rhs, shape=(n,))
        first = g.lib.nusifSolverFirstGroup(C.byref(g.s)) if inner == 0 else inner
        cyc, _ = oracle_solve(pv, rv, prm, opts, first)
        assert g.lib.nusifSolverIterations(C.byref(g.s)) == cyc, step
        if inner == 0:
            if prev is None:
                assert first == 1
            elif prev[1] > prev[0]:
                if first != prev[1]:
                    # set below the stalled group, or kept from an earlier stall
                    assert precision == "mixed" and (first < prev[0] or first == cap), (step, prev, first, cap)
                    cap = first
                    print("step %d: group capped at %d after %s" % (step, first, prev))
            else:
                assert first in (prev[0], max(1, prev[0] - 1)), (step, prev, first)
        prev = (first, cyc)
        o.call("adaptUV")
    fg, fo = g.fields(), o.fields()
    for k in ("u", "v", "w", "p"):
        err = A.rel_maxnorm(fg[k], fo[k])
        print(precision, k, "rel err vs oracle", err)
        assert err <= (1e-10 if precision == "fp64" else 1e-6), k
    g.close()
    o.close()


@pytest.mark.parametrize("shape,levels,precision", [((64, 64, 64), 6, "fp64"),
                                                   ((64, 64, 64), 6, "mixed"),
                                                   ((128, 16, 32), 4, "mixed")])
def test_gpu_gmg_tail_cycle_equals_per_level_kernels(shape, levels, precision, gpu_ok):
    """The one-workgroup coarse tail (mg_tail 1) runs the same passes in the
    same operand order as the per-level kernels (mg_tail 0): same p, bit for
    bit, and the same cycle count."""
    print("gmg tail", shape, levels, precision, flush=True)
    ni, nj, nk = shape
    prm = dict(G.KAT_PARAMS, imax=ni, jmax=nj, kmax=nk, xlength=ni / 32.0, ylength=nj / 32.0,
               zlength=nk / 32.0, levels=levels, presmooth=2, postsmooth=2)
    out = []
    for tail in (0, 1):
        opts = {"solver": "gmg", "precision": precision, "mg_tol": 1e-9, "mg_project": 1,
                "mg_tail": tail}
        c = A.Case(lib("gpu_rb"), options=opts, **prm)
        f = G.random_fields(c.shape, 7)
        c.set_fields(p=np.zeros(c.shape), rhs=f["rhs"])
        c.lib.solve(C.byref(c.s), c.d.p, c.d.rhs)
        out.append((c.fields()["p"], c.lib.nusifSolverIterations(C.byref(c.s))))
        c.close()
    for o in out[1:]:
        assert o[1] == out[0][1]
        np.testing.assert_array_equal(o[0], out[0][0])


@pytest.mark.parametrize("shape,levels,precision,post", [((128, 128, 64), 5, "mixed", 2),
                                                        ((128, 128, 64), 5, "fp64", 2),
                                                        ((96, 64, 32), 3, "mixed", 1),
                                                        ((96, 64, 32), 3, "fp64", 3),
                                                        ((64, 32, 64), 4, "mixed", 2)])
def test_gpu_gmg_fused_prolongation_equals_separate_pass(shape, levels, precision, post, gpu_ok):
    """The prolongation inside the first post-smoothing sweep (mg_fuse_prolong
    1) against kProlongMarch followed by a plain sweep (mg_fuse_prolong 0):
    the same p, bit for bit, and the same cycle count. The shapes give tiles
    of 64 x 16, 64 x 8 and 64 x 4 cells, k-chunks of one to sixteen planes
    and ragged tiles (96 columns); 1 to 3 post-smoothing sweeps."""
    print("gmg fused prolongation", shape, levels, precision, post, flush=True)
    ni, nj, nk = shape
    prm = dict(G.KAT_PARAMS, imax=ni, jmax=nj, kmax=nk, xlength=ni / 32.0, ylength=nj / 32.0,
               zlength=nk / 32.0, levels=levels, presmooth=2, postsmooth=post)
    out = []
    for fuse in (0, 1):
        opts = {"solver": "gmg", "precision": precision, "mg_tol": 1e-9, "mg_project": 1,
                "mg_fuse_prolong": fuse}
        c = A.Case(lib("gpu_rb"), options=opts, **prm)
        f = G.random_fields(c.shape, 11)
        c.set_fields(p=np.zeros(c.shape), rhs=f["rhs"])
        c.lib.solve(C.byref(c.s), c.d.p, c.d.rhs)
        out.append((c.fields()["p"], c.lib.nusifSolverIterations(C.byref(c.s))))
        c.close()
    assert out[1][1] == out[0][1]
    assert out[0][1] >= 2
    np.testing.assert_array_equal(out[1][0], out[0][0])


@pytest.mark.parametrize("precision,speculate,cache", [("fp64", 1, 64), ("mixed", 1, 64), ("fp64", 0, 2),
                                                     ("mixed", 0, 2), ("mixed", 1, 2)])
def test_gpu_gmg_graph_replay_equals_eager_launches(precision, speculate, cache, gpu_ok):
    """Cycles replayed from captured HIP graphs (mg_graph 1) against the same
    cycles launched kernel by kernel (mg_graph 0): whole cavity steps, each
    with several cycles (replays of the cached graphs) and the step's own
    eager kernels and host synchronisations between them; every field equal
    bit for bit after every step. Speculative cycles: the replayed fold
    kernel must write the cycle-control slot of the cycle that runs, not of
    the captured one. A cache of two graphs with groups sized from the
    previous solve (mg_inner 0) evicts and recaptures throughout."""
    print("gmg graph", precision, "speculate", speculate, "cache", cache, flush=True)
    prm = dict(G.TRAJ["c1_rb"][1], imax=64, jmax=64, kmax=64, re=1000.0, levels=5,
               presmooth=2, postsmooth=2)
    cases = []
    for graph in (0, 1):
        opts = {"solver": "gmg", "precision": precision, "mg_tol": 1e-8, "mg_graph": graph,
                "mg_speculate": speculate, "mg_inner": 0, "mg_graph_cache": cache}
        cases.append(A.Case(lib("gpu_rb"), options=opts, **prm))
    total_cycles = 0
    for step in range(6):
        for c in cases:
            c.step()
        it = [c.lib.nusifSolverIterations(C.byref(c.s)) for c in cases]
        assert it[0] == it[1]
        total_cycles += it[0]
        fa, fb = cases[0].fields(), cases[1].fields()
        for k in ("u", "v", "w", "p"):
            np.testing.assert_array_equal(fa[k], fb[k], err_msg="%s step %d" % (k, step))
    assert total_cycles >= 6
    for c in cases:
        c.close()


WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cfd-bench_amd", "weights",
                       "unet_cavity.bin")


def test_gpu_gmg_graph_replay_run_steps_with_eviction(gpu_ok):
    """nusifRunSteps (adaptUV queued behind the first cycle group, no host
    synchronisation between steps but the norm read-backs) with graphs
    evicted and recaptured from a cache of two: trace and fields equal to the
    eager launches bit for bit."""
    print("gmg graph run steps, cache 2", flush=True)
    prm = dict(G.TRAJ["c1_rb"][1], imax=64, jmax=64, kmax=64, re=1000.0, levels=5,
               presmooth=2, postsmooth=2)
    n = 10
    out = []
    for graph in (0, 1):
        opts = {"solver": "gmg", "precision": "mixed", "mg_tol": 1e-8, "mg_graph": graph, "mg_inner": 0,
                "mg_graph_cache": 2}
        c = A.Case(lib("gpu_rb"), options=opts, **prm)
        tr = (C.c_double * (4 * n))()
        assert c.lib.nusifRunSteps(C.byref(c.d), C.byref(c.s), n, tr) == n
        out.append((np.array(tr), c.fields()))
        c.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for k in A.FIELDS:
        np.testing.assert_array_equal(out[0][1][k], out[1][1][k], err_msg=k)


def test_gpu_graph_replay_refused_with_warmstart(gpu_ok, tmp_path):
    """mg_graph 1 together with warmstart unet stops in initSolver with a
    message (the combination is not supported), in a child process"""
    script = tmp_path / "refuse.py"
    script.write_text("""
import sys
sys.path.insert(0, %r)
import nusif_api as A
lib = A.load(A.gpu_lib_path("rb"))
A.Case(lib, options={"solver": "gmg", "precision": "mixed", "mg_graph": 1, "warmstart": "unet",
                     "unet_weights": %r}, name="dcavity", imax=32, jmax=32, kmax=32, levels=4)
print("initialised")
""" % (os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cfd-bench_amd"), WEIGHTS))
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "initialised" not in r.stdout
    assert "does not combine with mg_graph 1" in r.stderr


@pytest.mark.parametrize("opts", [{"precision": "fp64", "mg_inner": 0}, {"precision": "mixed", "mg_inner": 0},
                                  {"precision": "mixed", "mg_inner": 1}, {"precision": "mixed", "mg_speculate": 1},
                                  {"precision": "mixed", "mg_inner": 0, "warmstart": "unet", "unet_weights": WEIGHTS}],
                         ids=["fp64", "mixed", "mixed-inner1", "mixed-speculate", "learned"])
def test_gpu_run_steps_early_adapt_equals_step_calls(opts, gpu_ok):
    """nusifRunSteps queues adaptUV right after the first cycle group, before
    the host reads its norms, and runs it again when later groups change p
    (mg_inner 1: every solve needs more than its first group). Against the
    same steps called one by one (adaptUV after the solve): every field and
    the trace (t, dt, residual, cycles) equal bit for bit."""
    print("run steps early adaptUV", opts, flush=True)
    prm = dict(G.TRAJ["c1_rb"][1], imax=64, jmax=64, kmax=64, re=1000.0, levels=5, presmooth=2,
               postsmooth=2)
    o = dict({"solver": "gmg", "mg_tol": 1e-8}, **opts)
    n = 8
    a = A.Case(lib("gpu_rb"), options=o, **prm).run(n)
    b = A.Case(lib("gpu_rb"), options=o, **prm)
    tr = (C.c_double * (4 * n))()
    assert b.lib.nusifRunSteps(C.byref(b.d), C.byref(b.s), n, tr) == n
    t = np.array(tr).reshape(n, 4)
    np.testing.assert_array_equal(t[:, :3], np.array(a.trace))
    assert t[-1, 3] == a.iterations()
    fa, fb = a.fields(), b.fields()
    for k in A.FIELDS:
        np.testing.assert_array_equal(fa[k], fb[k], err_msg=k)
    a.close()
    b.close()


@pytest.mark.parametrize("shape,levels", [((128, 144, 240), 5), ((64, 12, 14), 2)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_gmg_tall_tile_sweep_equals_default(shape, levels, precision, gpu_ok):
    """The fused sweep with 64 x 8 and 64 x 16 tiles against 64 x 4 tiles:
    same p bit for bit. At 128 x 144 x 240 the three heights run with
    different k-chunks on the finest level (FP64: 8, 4 and 16 planes; FP32:
    8, 4 and 4) and the next levels' 72 and 36 rows are no multiple of 16; at
    64 x 12 x 14 the 64 x 16 tile is partial in j and its last 4-plane chunk
    holds 2 planes."""
    print("gmg tile rows", shape, precision, flush=True)
    ni, nj, nk = shape
    prm = dict(G.KAT_PARAMS, imax=ni, jmax=nj, kmax=nk, xlength=ni / 64.0, ylength=nj / 64.0,
               zlength=nk / 64.0, levels=levels, presmooth=2, postsmooth=2)
    out = []
    for rows in (4, 8, 16):
        opts = {"solver": "gmg", "precision": precision, "mg_tol": 1e-9, "mg_project": 1,
                "mg_tile_rows": rows}
        c = A.Case(lib("gpu_rb"), options=opts, **prm)
        f = G.random_fields(c.shape, 9)
        c.set_fields(p=np.zeros(c.shape), rhs=f["rhs"])
        c.lib.solve(C.byref(c.s), c.d.p, c.d.rhs)
        out.append((c.fields()["p"], c.lib.nusifSolverIterations(C.byref(c.s))))
        c.close()
    for o in out[1:]:
        assert o[1] == out[0][1]
        np.testing.assert_array_equal(o[0], out[0][0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_channel_steps_per_field_error(precision, gpu_ok):
    """The channel form (canal.par: outflow left/right, no-slip walls,
    uniform inflow u = 1 at x = 0, domain 4 x 1 x 1, Re 100) at 64 x 16 x 16
    with the gmg solver to 1e-8 (rhs mean removed: the inflow/outflow rhs is
    not compatible with the all-Neumann pressure problem): GPU steps against
    the oracle's step functions + the oracle's FP64 gmg. The error per field
    is printed and bounded: FP64 <= 1e-10, mixed (FP32 V-cycles inside FP64
    defect correction) <= 1e-6."""
    print("channel", precision, flush=True)
    prm = dict(name="canal", re=100.0, imax=64, jmax=16, kmax=16, xlength=4.0, ylength=1.0,
               zlength=1.0, bcLeft=3, bcRight=3, bcBottom=1, bcTop=1, bcFront=1, bcBack=1,
               u_init=1.0, tau=0.5, te=1e9, gamma=0.9, levels=3, presmooth=2, postsmooth=2)
    opts = {"solver": "gmg", "precision": precision, "mg_tol": 1e-8, "mg_project": 1,
            "u_inflow": 1.0}
    g = A.Case(lib("gpu_rb"), options=opts, **prm)
    o_lib = lib("orc_rb")
    o_lib.orcSetProblemExtras(1.0, 1.0)
    try:
        o = A.Case(o_lib, **prm)
        for step in range(8):
            g.step()
            d = o.d
            o.lib.computeTimestep(C.byref(d))
            for fn in ("setBoundaryConditions", "setSpecialBoundaryCondition", "computeFG",
                       "computeRHS"):
                o.call(fn)
            if step == 0:
                o.call("normalizePressure")
            n = int(np.prod(o.shape))
            pv = np.ctypeslib.as_array(d.p, shape=(n,))
            rv = np.ctypeslib.as_array(d.rhs, shape=(n,))
            oracle_solve(pv, rv, dict(prm), dict(opts, precision="fp64"))
            o.call("adaptUV")
    finally:
        o_lib.orcSetProblemExtras(1.0, 2.0)
    fg, fo = g.fields(), o.fields()
    assert np.all(fg["u"][1:-1, 1:-1, 0] == 1.0)
    for k in ("u", "v", "w", "p"):
        err = A.rel_maxnorm(fg[k], fo[k])
        print(precision, "channel", k, "rel err vs FP64 oracle %.3e" % err)
        assert err <= (1e-10 if precision == "fp64" else 1e-6), k
    g.close()
    o.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_gmg_speculative_cycles_equal_one_at_a_time(precision, gpu_ok):
    """Cycles enqueued ahead of the convergence read-back (mg_speculate 1:
    the device skips those past convergence and the host restores the
    buffer arrangement) against one cycle per read-back (mg_speculate 0):
    the same cycle count every step and every field equal bit for bit,
    over steps whose cycle counts go up and down (a batch sized by the
    previous step then overshoots or falls short)."""
    print("gmg speculate", precision, flush=True)
    prm = dict(G.TRAJ["c1_rb"][1], imax=64, jmax=64, kmax=64, re=1000.0, levels=5,
               presmooth=2, postsmooth=2)
    cases = []
    for spec in (0, 1):
        opts = {"solver": "gmg", "precision": precision, "mg_tol": 1e-7, "mg_omega": 1.3,
                "mg_speculate": spec}
        cases.append(A.Case(lib("gpu_rb"), options=opts, **prm))
    counts = []
    for step in range(10):
        for c in cases:
            c.step()
        it = [c.lib.nusifSolverIterations(C.byref(c.s)) for c in cases]
        assert it[0] == it[1], (step, it)
        counts.append(it[0])
        fa, fb = cases[0].fields(), cases[1].fields()
        for k in ("u", "v", "w", "p"):
            np.testing.assert_array_equal(fa[k], fb[k], err_msg="%s step %d" % (k, step))
    print("cycles per step", counts)
    assert min(counts) >= 1
    for c in cases:
        c.close()


def _np_lap7(x, ni, nj, nk, idx2, idy2, idz2):
    """The solvers' 7-point Neumann Laplacian on the interior, in their
    operand order ((E - 2C + W) * idx2 + (N - 2C + S) * idy2 + (T - 2C + B) * idz2)."""
    c = x[1:-1, 1:-1, 1:-1]
    e, w = x[1:-1, 1:-1, 2:], x[1:-1, 1:-1, :-2]
    n, s = x[1:-1, 2:, 1:-1], x[1:-1, :-2, 1:-1]
    t, b = x[2:, 1:-1, 1:-1], x[:-2, 1:-1, 1:-1]
    return (e - 2.0 * c + w) * idx2 + (n - 2.0 * c + s) * idy2 + (t - 2.0 * c + b) * idz2


@pytest.mark.parametrize("shape", [(64, 8, 32), (96, 22, 18)])
def test_gpu_input_restriction_matches_numpy(shape, gpu_ok):
    """The warm start's input pass (r0 = rhs - A p in FP64 with the 8-child
    averages of r0 and rhs written by the residual kernel itself) against a
    numpy restatement: the children summed in the order i, j, k as double,
    times 1/8, as FP32: bit for bit. The r0^2 sum within 1e-12 relative
    (workgroup partials, another summation order)."""
    ni, nj, nk = shape
    dx, dy, dz = 1.0 / ni, 0.7 / nj, 1.3 / nk
    rng = np.random.default_rng(4)
    p = rng.uniform(-1, 1, (nk + 2, nj + 2, ni + 2))
    rhs = rng.uniform(-1, 1, (nk + 2, nj + 2, ni + 2))
    gc = (nk // 2 + 2, nj // 2 + 2, ni // 2 + 2)
    rc = np.zeros(gc, dtype=np.float32)
    bc = np.zeros(gc, dtype=np.float32)
    ss = C.c_double()
    fp = C.POINTER(C.c_float)
    lib("gpu_rb").nusifRestrictResidualCheck(ni, nj, nk, dx, dy, dz, p.ctypes.data_as(DP), rhs.ctypes.data_as(DP),
                                             rc.ctypes.data_as(fp), bc.ctypes.data_as(fp), C.byref(ss))
    r = rhs[1:-1, 1:-1, 1:-1] - _np_lap7(p, ni, nj, nk, 1.0 / (dx * dx), 1.0 / (dy * dy), 1.0 / (dz * dz))
    b = rhs[1:-1, 1:-1, 1:-1]
    for src, out in ((r, rc), (b, bc)):
        s = np.zeros((nk // 2, nj // 2, ni // 2))
        for dk in (0, 1):
            for dj in (0, 1):
                for di in (0, 1):
                    s = s + src[dk::2, dj::2, di::2]
        np.testing.assert_array_equal(out[1:-1, 1:-1, 1:-1], (0.125 * s).astype(np.float32))
    assert ss.value == pytest.approx(float(np.sum(r * r)), rel=1e-12)


@pytest.mark.parametrize("shape", [(32, 4, 16), (48, 11, 7)])
def test_gpu_prolong_set_matches_numpy(shape, gpu_ok):
    """The warm start's FP32 trilinear prolongation (the multigrid's
    column-marching kernel with the old value not added) against numpy:
    weights 3/4, 1/4 per direction, coarse neighbour clamped at the walls,
    in-plane terms first; interior and face ghosts bit for bit."""
    nic, njc, nkc = shape
    rng = np.random.default_rng(6)
    xc = rng.uniform(-1, 1, (nkc + 2, njc + 2, nic + 2)).astype(np.float32)
    ni, nj, nk = 2 * nic, 2 * njc, 2 * nkc
    xf = np.zeros((nk + 2, nj + 2, ni + 2), dtype=np.float32)
    fp = C.POINTER(C.c_float)
    lib("gpu_rb").nusifProlongSetCheck(nic, njc, nkc, xc.ctypes.data_as(fp), xf.ctypes.data_as(fp))
    a, q = np.float32(0.75), np.float32(0.25)

    def idx(n, nc):
        f = np.arange(1, n + 1)
        I = (f + 1) // 2
        I2 = np.clip(I + np.where(f & 1, -1, 1), 1, nc)
        return I, I2

    I, I2 = idx(ni, nic)
    J, J2 = idx(nj, njc)
    K, K2 = idx(nk, nkc)

    def S(Kp):
        pl = xc[Kp]  # (n fine k, nj+2, ni+2)
        c1 = a * pl[:, J][:, :, I] + q * pl[:, J][:, :, I2]
        c2 = a * pl[:, J2][:, :, I] + q * pl[:, J2][:, :, I2]
        return a * c1 + q * c2

    ref = a * S(K) + q * S(K2)
    np.testing.assert_array_equal(xf[1:-1, 1:-1, 1:-1], ref)
    np.testing.assert_array_equal(xf[1:-1, 1:-1, 0], ref[:, :, 0])
    np.testing.assert_array_equal(xf[1:-1, 1:-1, -1], ref[:, :, -1])
    np.testing.assert_array_equal(xf[1:-1, 0, 1:-1], ref[:, 0, :])
    np.testing.assert_array_equal(xf[1:-1, -1, 1:-1], ref[:, -1, :])
    np.testing.assert_array_equal(xf[0, 1:-1, 1:-1], ref[0])
    np.testing.assert_array_equal(xf[-1, 1:-1, 1:-1], ref[-1])
