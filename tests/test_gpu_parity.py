"""GPU library (HIP, gfx950) against the reference's golden vectors and the
CPU oracle, through the C ABI (cfd-bench_amd/lib/libnusif_*.so).

Bar: pointwise kernels (BCs, computeFG, computeRHS, adaptUV, the CFL maxima,
every SOR sweep) evaluate the reference's expressions in its operand order
with contraction off, so they are bit-identical. Sums (normalizePressure,
residuals) run in a different order: relative 1e-13 on those values and
1e-10 relative max-norm on fields after whole runs (the target of the
plain FP64 path)."""
import ctypes as C

import numpy as np
import pytest

import make_golden as G
import nusif_api as A
from conftest import lib

pytestmark = pytest.mark.gpu
SUM_RTOL = 1e-13
FIELD_RTOL = 1e-10


def run_kat(libname, case):
    name, kind, prm, seed, funcs = case
    c = A.Case(lib("%s_%s" % (libname, kind)), **prm)
    inputs = G.random_fields(c.shape, seed)
    c.set_fields(**inputs)
    out = {}
    for fn in funcs:
        if fn == "solve":
            out["res"] = c.lib.solve(C.byref(c.s), c.d.p, c.d.rhs)
        else:
            c.call(fn)
    out.update(c.fields())
    out["dt"] = c.d.dt
    out["iterations"] = c.iterations()
    c.close()
    return inputs, out


@pytest.mark.parametrize("case", G.kat_cases(), ids=lambda c: c[0])
def test_gpu_kat(case, kat, gpu_ok):
    name = case[0]
    print("case", name, flush=True)
    inputs, out = run_kat("gpu", case)
    for k in A.FIELDS:
        key = "%s/%s" % (name, k)
        expect = kat[key] if key in kat.files else inputs[k]
        if name == "normalize" and k == "p":
            np.testing.assert_allclose(out[k], expect, rtol=0, atol=1e-15, err_msg=key)
        else:
            np.testing.assert_array_equal(out[k], expect, err_msg=key)
    assert out["dt"] == float(kat[name + "/dt"])
    if name + "/res" in kat.files:
        assert out["res"] == pytest.approx(float(kat[name + "/res"]), rel=SUM_RTOL)


@pytest.mark.parametrize("shape", [(40, 36, 28), (33, 17, 9), (5, 3, 2)])
def test_gpu_pointwise_vs_oracle(shape, gpu_ok):
    """Odd and small shapes (partial workgroups, degenerate faces)."""
    print("shape", shape, flush=True)
    ni, nj, nk = shape
    prm = dict(G.KAT_PARAMS, imax=ni, jmax=nj, kmax=nk, bcLeft=3, bcRight=2, bcBottom=1,
               bcTop=1, bcFront=2, bcBack=3)
    for fn in ("setBoundaryConditions", "setSpecialBoundaryCondition", "computeFG",
               "computeRHS", "adaptUV", "computeTimestep"):
        outs = []
        for src in ("gpu", "orc"):
            c = A.Case(lib(src + "_rb"), **prm)
            c.set_fields(**G.random_fields(c.shape, 21))
            c.call(fn)
            outs.append((c.fields(), c.d.dt))
            c.close()
        (fg, dg), (fo, do) = outs
        assert dg == do, fn
        for k in A.FIELDS:
            np.testing.assert_array_equal(fg[k], fo[k], err_msg="%s %s" % (fn, k))


@pytest.mark.parametrize("name", list(G.TRAJ))
def test_gpu_trajectory_vs_reference(name, traj, gpu_ok):
    print("trajectory", name, flush=True)
    kind, prm, n = G.TRAJ[name]
    g = A.Case(lib("gpu_" + kind), **prm).run(n)
    tg = np.array(g.trace)
    te = traj[name + "/trace"]
    np.testing.assert_array_equal(tg[:, 1], te[:, 1])           # dt sequence exact
    np.testing.assert_allclose(tg[:, 2], te[:, 2], rtol=1e-10)  # residual per step
    f = g.fields()
    for k in ("p", "u", "v", "w"):
        err = A.rel_maxnorm(f[k], traj[name + "/" + k])
        assert err <= FIELD_RTOL, (k, err)
    g.close()


def test_gpu_rb_iteration_counts_match_oracle(gpu_ok):
    prm = G.TRAJ["c1_rb"][1]
    g = A.Case(lib("gpu_rb"), **prm)
    o = A.Case(lib("orc_rb"), **prm)
    for step in range(30):
        g.step()
        o.step()
        assert g.iterations() == o.iterations(), step
    g.close()
    o.close()


def test_gpu_run_steps_matches_step_calls(gpu_ok):
    prm = dict(G.TRAJ["c1_rb"][1], imax=24, jmax=20, kmax=16)
    a = A.Case(lib("gpu_rb"), **prm).run(12)
    b = A.Case(lib("gpu_rb"), **prm)
    tr = (C.c_double * (4 * 12))()
    assert b.lib.nusifRunSteps(C.byref(b.d), C.byref(b.s), 12, tr) == 12
    t = np.array(tr).reshape(12, 4)
    np.testing.assert_array_equal(t[:, :3], np.array(a.trace))
    fa, fb = a.fields(), b.fields()
    for k in A.FIELDS:
        np.testing.assert_array_equal(fa[k], fb[k], err_msg=k)
    a.close()
    b.close()


@pytest.mark.parametrize("bcs,name,shape,extra", [
    ("noslip", "dcavity", (24, 20, 16), {}),
    ("mixed", "dcavity", (23, 9, 7), {}),
    ("mixed2", "canal", (16, 5, 12), {"inflow_noise": "0.01"}),
    ("mixed", "canal", (3, 3, 3), {}),
    ("noslip", "canal", (24, 8, 8), {}),
])
def test_gpu_fused_boundary_step_equals_reference_calls(bcs, name, shape, extra, gpu_ok):
    """nusifRunSteps with each step's CFL maxima over the ghost shell, the six
    velocity faces and the lid or inflow in one launch (bc_fused 1) against
    computeTimestep, setBoundaryConditions and setSpecialBoundaryCondition
    one by one (bc_fused 0): eight steps, the trace (t, dt, res) and every
    field bit for bit. The BC sets mix NOSLIP, SLIP, OUTFLOW and PERIODIC
    faces (the fused kernel moves the ghost updates whose source a later
    face overwrites into that face's task, and applies an earlier face's wall
    rule where a later face reads the wall velocity); 3 x 3 x 3 is the
    smallest grid it takes."""
    print("fused boundary step", bcs, name, shape, flush=True)
    prm = dict(G.KAT_PARAMS, **G.BC_SETS[bcs], name=name, imax=shape[0], jmax=shape[1], kmax=shape[2])
    out = []
    for fused in (0, 1):
        c = A.Case(lib("gpu_rb"), options=dict(extra, bc_fused=str(fused)), **prm)
        c.set_fields(**G.random_fields(c.shape, 5))
        tr = (C.c_double * (4 * 8))()
        assert c.lib.nusifRunSteps(C.byref(c.d), C.byref(c.s), 8, tr) == 8
        out.append((np.array(tr), c.fields()))
        c.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for k in A.FIELDS:
        np.testing.assert_array_equal(out[0][1][k], out[1][1][k], err_msg=k)


@pytest.mark.parametrize("shape", [(32, 32, 32), (70, 37, 9), (48, 20, 4)])
def test_gpu_rb_fused_iteration_matches_colour_passes(gpu_ok, shape):
    """One launch per SOR iteration (rb_fused 1, the default) against one
    launch per colour pass plus the ghost pass (rb_fused 0): the same p, u,
    v, w bits and the same iteration counts over a run (ragged tiles and
    k-chunks on the odd shapes; the result ends in either buffer)."""
    prm = dict(G.TRAJ["c1_rb"][1], imax=shape[0], jmax=shape[1], kmax=shape[2])
    a = A.Case(lib("gpu_rb"), options={"rb_fused": 1}, **prm)
    b = A.Case(lib("gpu_rb"), options={"rb_fused": 0}, **prm)
    its_a, its_b = [], []
    for _ in range(8):
        a.step()
        b.step()
        its_a.append(a.iterations())
        its_b.append(b.iterations())
    assert its_a == its_b
    fa, fb = a.fields(), b.fields()
    for k in A.FIELDS:
        np.testing.assert_array_equal(fa[k], fb[k], err_msg=k)
    a.close()
    b.close()


@pytest.mark.parametrize("shape", [(32, 32, 32), (70, 37, 9)])
def test_gpu_mg_fused_sweeps_match_colour_passes(gpu_ok, shape):
    """The reference MG's smoothing sweeps (sweep + ghost copy) as one fused
    launch each against colour passes + ghost pass: the same bits, residual
    included (V(2,2): the result ends on p after four sweeps, V(1,2) on the
    partner buffer after three)."""
    for pre, post in ((2, 2), (1, 2)):
        prm = dict(G.TRAJ["c1_rb"][1], imax=shape[0], jmax=shape[1], kmax=shape[2], levels=3,
                   presmooth=pre, postsmooth=post)
        a = A.Case(lib("gpu_mg"), options={"rb_fused": 1}, **prm)
        b = A.Case(lib("gpu_mg"), options={"rb_fused": 0}, **prm)
        for _ in range(6):
            assert a.step() == b.step()
        fa, fb = a.fields(), b.fields()
        for k in A.FIELDS:
            np.testing.assert_array_equal(fa[k], fb[k], err_msg="%s V(%d,%d)" % (k, pre, post))
        a.close()
        b.close()


def test_gpu_solve_timing(gpu_ok):
    """Events around each solve of nusifRunSteps (70 steps: more than one
    pool of event pairs) count every solve, sum to less than the run's wall
    time, and change no result."""
    import time
    prm = dict(G.TRAJ["c1_rb"][1], imax=24, jmax=20, kmax=16)
    a = A.Case(lib("gpu_rb"), **prm)
    b = A.Case(lib("gpu_rb"), **prm)
    n = 70
    ta, tb = (C.c_double * (4 * n))(), (C.c_double * (4 * n))()
    assert a.lib.nusifRunSteps(C.byref(a.d), C.byref(a.s), n, ta) == n
    b.lib.nusifSolveTiming(C.byref(b.d), 1)
    t0 = time.perf_counter()
    assert b.lib.nusifRunSteps(C.byref(b.d), C.byref(b.s), n, tb) == n
    wall_ms = 1e3 * (time.perf_counter() - t0)
    ms, solves = C.c_double(), C.c_longlong()
    b.lib.nusifSolveTimes(C.byref(b.d), C.byref(ms), C.byref(solves))
    assert solves.value == n
    assert 0.0 < ms.value < wall_ms
    b.lib.nusifSolveTimes(C.byref(b.d), C.byref(ms), C.byref(solves))
    assert solves.value == 0 and ms.value == 0.0  # reset by the read
    np.testing.assert_array_equal(np.array(ta), np.array(tb))
    fa, fb = a.fields(), b.fields()
    for k in A.FIELDS:
        np.testing.assert_array_equal(fa[k], fb[k], err_msg=k)
    a.close()
    b.close()


def test_gpu_solve_on_foreign_host_arrays(gpu_ok):
    """solve() on host arrays that are not the Discretization's own is staged
    through the device and gives the same result, and leaves the
    Discretization's own p and rhs as they were (the reference's solve
    touches only the arrays it is given)."""
    prm = dict(G.KAT_PARAMS)
    c = A.Case(lib("gpu_rb"), **prm)
    own = G.random_fields(c.shape, 3)
    c.set_fields(p=own["p"], rhs=own["rhs"])
    inputs = G.random_fields(c.shape, 10)
    p = np.ascontiguousarray(inputs["p"]).copy()
    rhs = np.ascontiguousarray(inputs["rhs"]).copy()
    dp = C.POINTER(C.c_double)
    res = c.lib.solve(C.byref(c.s), p.ctypes.data_as(dp), rhs.ctypes.data_as(dp))
    _, ref = run_kat("orc", ("solve_rb", "rb", prm, 10, ["solve"]))
    np.testing.assert_array_equal(p, ref["p"])
    assert res == pytest.approx(ref["res"], rel=SUM_RTOL)
    after = c.fields()
    np.testing.assert_array_equal(after["p"], own["p"])
    np.testing.assert_array_equal(after["rhs"], own["rhs"])
    c.close()


def test_gpu_canal_inflow_option(gpu_ok):
    prm = dict(G.TRAJ["canal_mg"][1])
    g = A.Case(lib("gpu_mg"), options={"u_inflow": 1.0}, **prm).run(5)
    o_lib = lib("orc_mg")
    o_lib.orcSetProblemExtras(1.0, 1.0)
    try:
        o = A.Case(o_lib, **prm).run(5)
    finally:
        o_lib.orcSetProblemExtras(1.0, 2.0)
    fg, fo = g.fields(), o.fields()
    assert np.all(fg["u"][1:-1, 1:-1, 0] == 1.0)
    for k in ("p", "u", "v", "w"):
        assert A.rel_maxnorm(fg[k], fo[k]) <= FIELD_RTOL, k
    g.close()
    o.close()


@pytest.mark.parametrize("shape", [(33, 17, 9), (64, 40, 36)])
def test_gpu_fused_rhs_equals_separate_pass(shape, gpu_ok):
    """computeFG forms rhs in its own pass (option fg_rhs, default on) and
    computeRHS publishes it: every field after boundary conditions, computeFG
    and computeRHS has the bits of the separate computeRHS pass (fg_rhs 0),
    on shapes with partial tiles in i, j and k and random fields. computeFG
    alone leaves rhs as it was, and a dt changed between the two calls makes
    computeRHS recompute with it."""
    print("fused rhs", shape, flush=True)
    ni, nj, nk = shape
    prm = dict(G.KAT_PARAMS, imax=ni, jmax=nj, kmax=nk, bcLeft=3, bcRight=2, bcBottom=1,
               bcTop=1, bcFront=2, bcBack=3)
    runs = {}
    for fused in (1, 0):
        c = A.Case(lib("gpu_rb"), options={"fg_rhs": fused}, **prm)
        c.set_fields(**G.random_fields(c.shape, 5))
        before = c.fields()["rhs"]
        for fn in ("setBoundaryConditions", "setSpecialBoundaryCondition", "computeFG"):
            c.call(fn)
        assert np.array_equal(c.fields()["rhs"], before)
        c.call("computeRHS")
        first = c.fields()
        # computeFG, then a new dt before computeRHS: the rhs must use it
        c.call("computeFG")
        c.d.dt = c.d.dt * 0.5
        c.call("computeRHS")
        runs[fused] = (first, c.fields())
        c.close()
    for q in range(2):
        for k in A.FIELDS:
            np.testing.assert_array_equal(runs[1][q][k], runs[0][q][k], err_msg="%d %s" % (q, k))


# Grids whose fused SOR iteration runs the 64x16 tile (jmax >= 128) and the
# 64x8 tile (jmax >= 64), the shapes of the 256^3 cavity's red-black solve;
# 70 x 130 leaves partial tiles in i and j. omega 1.9 and eps 3e-2 let the
# reference's stop test fire after ~350-1500 iterations (oracle counts).
WIDE = [(64, 128, 24), (48, 64, 16), (70, 130, 10)]


def _wide_prm(shape, itermax=3000):
    return dict(G.TRAJ["c1_rb"][1], imax=shape[0], jmax=shape[1], kmax=shape[2], omg=1.9, eps=3e-2,
                itermax=itermax)


@pytest.mark.parametrize("shape", WIDE)
def test_gpu_rb_fused_wide_tiles_match_colour_passes(gpu_ok, shape):
    """The fused iteration with the in-sweep residual on 64x16 and 64x8
    tiles against one launch per colour pass: p, u, v, w bit for bit and the
    same iteration counts, with the stop test firing before itermax."""
    print("rb fused wide", shape, flush=True)
    a = A.Case(lib("gpu_rb"), options={"rb_fused": 1}, **_wide_prm(shape))
    b = A.Case(lib("gpu_rb"), options={"rb_fused": 0}, **_wide_prm(shape))
    its_a, its_b = [], []
    for _ in range(4):
        ra, rb = a.step(), b.step()
        its_a.append(a.iterations())
        its_b.append(b.iterations())
        assert ra == pytest.approx(rb, rel=1e-12)
    print("iterations", its_a, flush=True)
    assert its_a == its_b
    assert max(its_a) < 3000  # every solve stopped on the residual
    fa, fb = a.fields(), b.fields()
    for k in A.FIELDS:
        np.testing.assert_array_equal(fa[k], fb[k], err_msg=k)
    a.close()
    b.close()


@pytest.mark.parametrize("shape", WIDE)
def test_gpu_rb_fused_wide_tiles_match_oracle(gpu_ok, shape):
    """The same tiles against the oracle's red-black SOR (solver-rb.c:27-172)
    with itermax capped at 20, so both run the same iterations: fields to
    1e-10 relative, the carried mean-square residual to 1e-13 (only its
    summation order differs)."""
    print("rb fused wide vs oracle", shape, flush=True)
    g = A.Case(lib("gpu_rb"), **_wide_prm(shape, itermax=20))
    o = A.Case(lib("orc_rb"), **_wide_prm(shape, itermax=20))
    for _ in range(3):
        rg, ro = g.step(), o.step()
        assert g.iterations() == o.iterations() == 20
        assert rg == pytest.approx(ro, rel=1e-13)
    fg, fo = g.fields(), o.fields()
    for k in ("u", "v", "w", "p"):
        assert A.rel_maxnorm(fg[k], fo[k]) <= 1e-10, k
    g.close()
    o.close()
