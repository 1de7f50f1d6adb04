"""Writes the golden fixtures of tests/golden/ from the reference itself.

The reference (HPC-Dwarfs/CFD-Bench, C99) is compiled from its own sources by
oracle/Makefile into oracle/_ref/ (strict IEEE flags) — in the container that
has the reference tree only. This script drives that build through the
reference's public API (ctypes, cfd-bench_amd/nusif_api.py) and stores inputs
and outputs as .npz data:

  kat.npz     per-function known-answer vectors on seeded random fields
              (12 x 10 x 8 cells, every boundary-condition type)
  traj.npz    whole-step trajectories: the 32^3 lid-driven cavity (Re=10,
              RB omega=1.7, eps=1e-3, itermax=500, 100 steps), a 16^3 cavity on
              the MG solver, and a 24 x 8 x 8 canal on the MG solver (levels 1)

Run: make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "cfd-bench_amd"))
import nusif_api as A  # noqa: E402

KAT_GRID = dict(imax=12, jmax=10, kmax=8, xlength=1.2, ylength=0.9, zlength=0.7)
KAT_PARAMS = dict(KAT_GRID, name="dcavity", re=100.0, gamma=0.9, tau=0.5, dt=1e-3, omg=1.7,
                  eps=1e-3, itermax=30, gx=0.1, gy=-0.2, gz=0.3, levels=2, presmooth=3,
                  postsmooth=2)
BC_SETS = {
    "noslip": dict(bcLeft=1, bcRight=1, bcBottom=1, bcTop=1, bcFront=1, bcBack=1),
    "mixed": dict(bcLeft=3, bcRight=3, bcBottom=2, bcTop=1, bcFront=4, bcBack=2),
    "mixed2": dict(bcLeft=2, bcRight=1, bcBottom=3, bcTop=3, bcFront=3, bcBack=3),
}
C1 = dict(name="dcavity", re=10.0, imax=32, jmax=32, kmax=32, tau=0.5, omg=1.7, eps=1e-3,
          itermax=500, gamma=0.9, te=100.0, dt=0.02)
MG16 = dict(name="dcavity", re=1000.0, imax=16, jmax=16, kmax=16, tau=0.5, omg=1.8, eps=1e-3,
            itermax=1000, gamma=0.9, te=100.0, levels=2, presmooth=20, postsmooth=5)
CANAL = dict(name="canal", re=100.0, imax=24, jmax=8, kmax=8, xlength=6.0, ylength=2.0,
             zlength=2.0, tau=0.5, omg=1.8, eps=1e-4, itermax=500, gamma=0.9, te=100.0,
             bcLeft=3, bcRight=3, u_init=1.0, levels=1, presmooth=10, postsmooth=5)
TRAJ = {"c1_rb": ("rb", C1, 100), "mg16": ("mg", MG16, 20), "canal_mg": ("mg", CANAL, 20)}


def splitmix_uniform(n, seed):
    """n values uniform in [-1, 1) from the splitmix64 sequence of `seed`
    (pure uint64 arithmetic, identical on every platform and numpy version)."""
    with np.errstate(over="ignore"):
        x = (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
             + np.uint64(seed) * np.uint64(0xD1B54A32D192ED03))
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def random_fields(shape, seed):
    n = int(np.prod(shape))
    return {k: splitmix_uniform(n, 100 * seed + q).reshape(shape)
            for q, k in enumerate(A.FIELDS)}


def kat_cases():
    """(case name, solver kind, params, seed, list of function names to call)."""
    cases = []
    for bcname, bcs in BC_SETS.items():
        prm = dict(KAT_PARAMS, **bcs)
        cases.append(("bc_" + bcname, "rb", prm, 1, ["setBoundaryConditions"]))
    cases.append(("lid", "rb", dict(KAT_PARAMS), 2, ["setSpecialBoundaryCondition"]))
    cases.append(("inflow", "rb", dict(KAT_PARAMS, name="canal"), 3,
                  ["setSpecialBoundaryCondition"]))
    cases.append(("timestep", "rb", dict(KAT_PARAMS), 4, ["computeTimestep"]))
    cases.append(("fg", "rb", dict(KAT_PARAMS), 5, ["computeFG"]))
    cases.append(("fg_upwind", "rb", dict(KAT_PARAMS, gamma=1.0, re=7.0), 6, ["computeFG"]))
    cases.append(("rhs", "rb", dict(KAT_PARAMS), 7, ["computeRHS"]))
    cases.append(("normalize", "rb", dict(KAT_PARAMS), 8, ["normalizePressure"]))
    cases.append(("adapt", "rb", dict(KAT_PARAMS), 9, ["adaptUV"]))
    cases.append(("solve_rb", "rb", dict(KAT_PARAMS), 10, ["solve"]))
    cases.append(("solve_rb_itermax", "rb", dict(KAT_PARAMS, eps=1e-9, itermax=7), 11, ["solve"]))
    cases.append(("solve_mg", "mg", dict(KAT_PARAMS), 12, ["solve"]))
    cases.append(("solve_mg_l1", "mg", dict(KAT_PARAMS, levels=1, presmooth=4), 13, ["solve"]))
    return cases


def run_kat(lib, kind, prm, seed, funcs):
    case = A.Case(lib, **prm)
    inputs = random_fields(case.shape, seed)
    case.set_fields(**inputs)
    out = {}
    for fn in funcs:
        if fn == "solve":
            out["res"] = np.array(case.lib.solve(A.C.byref(case.s), case.d.p, case.d.rhs))
        else:
            case.call(fn)
    # only the fields the call changed; the inputs are regenerated from the seed
    for k, v in case.fields().items():
        if not np.array_equal(v, inputs[k]):
            out[k] = v
    out["dt"] = np.array(case.d.dt)
    case.close()
    return out


def run_traj(lib, prm, nsteps):
    case = A.Case(lib, **prm).run(nsteps)
    f = case.fields()
    out = {k: f[k] for k in ("p", "u", "v", "w")}
    out["trace"] = np.array(case.trace)
    case.close()
    return out


def main():
    libs = {k: A.load(A.ref_lib_path(k)) for k in ("rb", "mg")}
    kat = {}
    for name, kind, prm, seed, funcs in kat_cases():
        out = run_kat(libs[kind], kind, prm, seed, funcs)
        for k, v in out.items():
            kat["%s/%s" % (name, k)] = v
    np.savez_compressed(os.path.join(HERE, "kat.npz"), **kat)
    traj = {}
    for name, (kind, prm, n) in TRAJ.items():
        for k, v in run_traj(libs[kind], prm, n).items():
            traj["%s/%s" % (name, k)] = v
    np.savez_compressed(os.path.join(HERE, "traj.npz"), **traj)
    print("wrote", len(kat), "kat arrays and", len(traj), "trajectory arrays")


if __name__ == "__main__":
    main()
