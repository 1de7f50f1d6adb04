"""Benchmark: time steps/s and MLUP/s of the 256^3 lid-driven cavity on one
MI355X, pressure solved each step to a relative residual of 1e-6.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode M] [--size N] [--opt KEY=VALUE]

A step is the reference time step (computeTimestep, boundary conditions,
computeFG, computeRHS, [normalizePressure], solve, adaptUV; reference
src/main.c:60-88) on the GPU library; the K timed steps run inside one
nusifRunSteps call with all fields resident in HBM. Modes:
  learned  U-Net warm start + FP32 V-cycles in FP64 defect correction (default)
  mixed    FP32 V-cycles in FP64 defect correction (previous-step guess)
  fp64     plain FP64 V(2,2) cycles (the comparison path)
The headline line also times the fp64 path on the same workload and reports
the speedup (and, for learned, the same FP32 solver without the network),
the roofline of the dominant kernel (HIP events on the library's stream) and
a CPU baseline (the oracle's FP64 multigrid on one host core, a bounded
sample of steps of the same workload). `rb_sor` times the red-black SOR solver
on the same state with the reference's own stop test, next to the oracle's
SOR on one host core (the reference's solver rate, iterations x cells / time).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cfd-bench_amd"))
import nusif_api as A  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "time steps/s + MLUP/s, 256^3 cavity: learned+FP32-MG vs plain FP64 MG"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cavity_params(n):
    return dict(name="dcavity", re=1000.0, imax=n, jmax=n, kmax=n, tau=0.5, omg=1.9, eps=1e-3,
                itermax=1000, gamma=0.9, te=1e9, levels=6, presmooth=2, postsmooth=2)


EXTRA_OPTS = {}


def mode_options(mode):
    # smoother over-relaxed (omega 1.3) in every mode, the FP64 comparison included:
    # 3.79 -> 3.08 V(2,2) cycles per step at 256^3 (measured over omega 1.0 .. 1.6);
    # first stop test after the smaller cycle count of the previous two solves
    # (mg_inner 0), also in every mode
    opts = {"solver": "gmg", "mg_tol": 1e-6, "mg_omega": 1.3, "mg_inner": 0}
    if mode in ("mixed", "learned"):
        opts["precision"] = "mixed"
    if mode == "learned":
        opts["warmstart"] = "unet"
        opts["unet_weights"] = os.path.join(ROOT, "cfd-bench_amd", "weights", "unet_cavity.bin")
    opts.update(EXTRA_OPTS)
    return opts


DOMINANT = {"fp64": "gmg_smooth64_l0", "mixed": "gmg_smooth32_l0", "learned": "gmg_smooth32_l0"}
TIMERS = ["gmg_smooth64_l0", "gmg_smooth32_l0", "gmg_smooth32z_l0", "gmg_restrict32_l0", "gmg_prolong32_l0",
          "gmg_restrict64_l0", "gmg_prolong64_l0", "gmg_residual64", "gmg_correct_residual",
          "ws_guess_residual",
          "computeFG", "computeFG_rhs",
          "computeRHS", "adaptUV", "absmax", "unet_forward", "unet_conv"]


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at n^3 from the committed rocprofv3
    counter summary (profiles/), or None when it was not measured."""
    path = os.path.join(ROOT, "profiles", "sweep32_cavity256_hbm_traffic.md")
    if n != 256 or not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return float(parts[2])
    return None


def run_mode(lib, mode, n, warmup, steps, profile_steps=20):
    """Warmup, then `steps` timed steps with no per-kernel timing, then
    `profile_steps` more steps with HIP-event timers around the kernels the
    roofline and the kernel table report (the events cost launch time, so
    they stay out of the headline timing), then `profile_steps` more with
    one event pair around each solve (the solver rate)."""
    case = A.Case(lib, options=mode_options(mode), **cavity_params(n))
    trace = (C.c_double * (4 * max(warmup, steps, profile_steps, 1)))()
    if warmup:
        lib.nusifRunSteps(C.byref(case.d), C.byref(case.s), warmup, trace)
    log("[bench] %s: %d warmup steps done" % (mode, warmup))
    lib.nusifDeviceSynchronize()
    tried0, taken0 = C.c_longlong(), C.c_longlong()
    lib.nusifWarmstartCounts(C.byref(case.s), C.byref(tried0), C.byref(taken0))
    t0 = time.perf_counter()
    lib.nusifRunSteps(C.byref(case.d), C.byref(case.s), steps, trace)
    lib.nusifDeviceSynchronize()
    t1 = time.perf_counter()
    tried, taken = C.c_longlong(), C.c_longlong()
    lib.nusifWarmstartCounts(C.byref(case.s), C.byref(tried), C.byref(taken))
    ws = None
    if tried.value > tried0.value:
        ws = (taken.value - taken0.value) / float(tried.value - tried0.value)
    cycles = [trace[4 * q + 3] for q in range(steps)]
    rels = [trace[4 * q + 2] for q in range(steps)]
    for t in TIMERS:
        lib.nusifTimerEnable(t.encode())
    lib.nusifTimerReset()
    lib.nusifRunSteps(C.byref(case.d), C.byref(case.s), profile_steps, trace)
    lib.nusifDeviceSynchronize()
    kern = {}
    for t in TIMERS:
        launches = C.c_int(0)
        ms = lib.nusifTimerRead(t.encode(), C.byref(launches))
        if launches.value:
            kern[t] = dict(ms_total=ms, launches=launches.value,
                           us_per_launch=1e3 * ms / launches.value,
                           bytes_per_launch=lib.nusifKernelBytes(C.byref(case.d), t.encode()))
    lib.nusifTimerDisableAll()
    # the solver region alone, as the reference rates it (src/solver-rb.c:161,
    # src/profiler.c:61): iterations x interior cells / solve time, an
    # iteration of the gmg solve being one V-cycle
    lib.nusifSolveTiming(C.byref(case.d), 1)
    lib.nusifRunSteps(C.byref(case.d), C.byref(case.s), profile_steps, trace)
    lib.nusifSolveTiming(C.byref(case.d), 0)
    solve_ms, solves = C.c_double(), C.c_longlong()
    lib.nusifSolveTimes(C.byref(case.d), C.byref(solve_ms), C.byref(solves))
    solve = dict(ms=solve_ms.value, solves=solves.value,
                 cycles=sum(trace[4 * q + 3] for q in range(profile_steps)))
    # the dominant kernel once more, 20 launches back to back on the live
    # level-0 buffers with events around the batch only (no per-launch event
    # packets): the duration the roofline uses
    sweep_us = lib.nusifTimeFinestSweep(C.byref(case.s), 20)
    log("[bench] %s: %d steps in %.3f s" % (mode, steps, t1 - t0))
    return case, dict(seconds=t1 - t0, cycles=cycles, rel=rels, kernels=kern, sweep_us=sweep_us,
                      warmstart_taken=ws, solve=solve)


def solver_rate(res, cells):
    """solver MLUP/s (V-cycles x interior cells / solve time) and the solve's
    milliseconds per step, from the solve-timed steps"""
    s = res["solve"]
    if not s["solves"] or s["ms"] <= 0:
        return {}
    return {"solver_mlups": s["cycles"] * cells / (s["ms"] * 1e-3) * 1e-6,
            "solve_ms_per_step": s["ms"] / s["solves"]}


def host_cpu():
    """model name and logical CPU count of this host (/proc/cpuinfo)"""
    model, count = "unknown", os.cpu_count()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, count


def cpu_baseline(case_gpu, n, budget_s=20.0):
    """The oracle (one host core) on the same workload: the state after the
    GPU warmup, then whole steps with its FP64 multigrid to the same 1e-6,
    with the GPU modes' smoother (omega 1.3) and first stop test after as many
    cycles as the previous solve needed (the GPU's mg_inner 0 grouping, whose
    contraction-based shrinking it leaves out)."""
    omega = float(mode_options("fp64")["mg_omega"])
    orc = A.load(A.oracle_lib_path("rb_fast"))
    gmg = A.load(A.oracle_lib_path("gmg_fast"))
    prm = cavity_params(n)
    o = A.Case(orc, **prm)
    o.set_fields(**case_gpu.fields())
    d = o.d
    cells = n ** 3
    steps, t_total, group, cycles = 0, 0.0, 1, 0
    while t_total < budget_s and steps < 50:
        t0 = time.perf_counter()
        orc.computeTimestep(C.byref(d))
        for fn in ("setBoundaryConditions", "setSpecialBoundaryCondition", "computeFG",
                   "computeRHS"):
            getattr(orc, fn)(C.byref(d))
        rel = C.c_double()
        g = d.grid
        c = gmg.orcGmgSolveGrouped(g.imax, g.jmax, g.kmax, g.dx, g.dy, g.dz, d.p, d.rhs, prm["levels"],
                                   prm["presmooth"], prm["postsmooth"], omega, 1e-6, 100, 0, 0, group,
                                   C.byref(rel))
        orc.adaptUV(C.byref(d))
        t_total += time.perf_counter() - t0
        steps += 1
        cycles += c
        group = max(1, c)
        log("[bench] cpu baseline step %d: %.2f s, %d cycles" % (steps, t_total, c))
    o.close()
    model, ncpu = host_cpu()
    return dict(value=steps / t_total, unit="time steps/s", cores=1, kind="port",
                host_cpu=model, host_logical_cpus=ncpu,
                sample="%d step(s) of the %d^3 cavity from the GPU state after warmup; oracle "
                       "C restatement (-O3 -ffast-math, one core of %s), FP64 V(2,2) 6-level multigrid "
                       "to 1e-6 relative residual, smoother omega %.1f as on the GPU, first stop test "
                       "after the previous solve's cycle count" % (steps, n, model, omega),
                cycles_per_step=cycles / steps, mlups=cells * steps / t_total * 1e-6)


def run_rb(lib, case_gpu, n, steps=3):
    """The red-black SOR solver (src/solver-rb.c, its own stop test: carried
    mean-square res < eps^2 or itermax; omega 1.9, eps 1e-3, itermax 1000) on
    the state the headline run reached: steps/s, SOR iterations per step, the
    reference's solver rate (iterations x interior cells / solve time,
    src/profiler.c:61) and the colour pass's rate."""
    c = A.Case(lib, options={"solver": "rb"}, **cavity_params(n))
    c.set_fields(**case_gpu.fields())
    trace = (C.c_double * (4 * steps))()
    lib.nusifSolveTiming(C.byref(c.d), 1)
    lib.nusifDeviceSynchronize()
    t0 = time.perf_counter()
    lib.nusifRunSteps(C.byref(c.d), C.byref(c.s), steps, trace)
    lib.nusifDeviceSynchronize()
    t1 = time.perf_counter()
    lib.nusifSolveTiming(C.byref(c.d), 0)
    ms, solves = C.c_double(), C.c_longlong()
    lib.nusifSolveTimes(C.byref(c.d), C.byref(ms), C.byref(solves))
    its = sum(trace[4 * q + 3] for q in range(steps))
    # the iteration's kernel: one fused launch (rb_fused, 24 B/cell) or a
    # launch per colour pass (rb_pass, 16 B/cell each)
    for t in (b"rb_fused", b"rb_pass"):
        lib.nusifTimerEnable(t)
    lib.nusifTimerReset()
    lib.nusifRunSteps(C.byref(c.d), C.byref(c.s), 1, trace)
    lib.nusifDeviceSynchronize()
    kern = None
    for t in (b"rb_fused", b"rb_pass"):
        launches = C.c_int(0)
        ms_t = lib.nusifTimerRead(t, C.byref(launches))
        if launches.value and kern is None:
            us = 1e3 * ms_t / launches.value
            kern = {"kernel": t.decode(), "us": us,
                    "GBs": lib.nusifKernelBytes(C.byref(c.d), t) / (us * 1e-6) / 1e9}
    lib.nusifTimerDisableAll()
    c.close()
    cells = float(n) ** 3
    out = {"value": steps / (t1 - t0), "unit": "time steps/s", "steps": steps,
           "iterations_per_step": its / steps,
           "solver_mlups": its * cells / (ms.value * 1e-3) * 1e-6,
           "solve_ms_per_step": ms.value / max(solves.value, 1),
           "stop_test": "reference: carried mean-square in-sweep residual < eps^2 = 1e-6, or itermax 1000"}
    if kern:
        out["iteration_kernel"] = kern
    log("[bench] rb: %d steps in %.3f s, %.0f iterations/step" % (steps, t1 - t0, its / steps))
    return out


def channel_params():
    """the combined configuration at scale: channel 768 x 192 x 192 (domain
    4 x 1 x 1), Re 500, inflow u = 1 with a seeded 1% uniform perturbation,
    zero-gradient outflow, no-slip walls"""
    return dict(name="canal", xlength=4.0, ylength=1.0, zlength=1.0, bcLeft=3, bcRight=3, bcBottom=1, bcTop=1,
                bcFront=1, bcBack=1, u_init=1.0, re=500.0, imax=768, jmax=192, kmax=192, tau=0.5, te=1e9,
                gamma=0.9, levels=6, presmooth=2, postsmooth=2)


def run_channel(lib, warmup=10, steps=40):
    """The channel at scale: the learned warm start (channel-tuned network)
    with FP32 V-cycles in FP64 defect correction against plain FP64 multigrid
    and against the same FP32 solver without the network, each to 1e-6
    relative residual per step, from the same start."""
    base = {"solver": "gmg", "mg_tol": 1e-6, "mg_project": 1, "u_inflow": 1.0, "inflow_noise": 0.01, "seed": 7,
            "mg_omega": 1.3, "mg_inner": 0}
    weights = os.path.join(ROOT, "cfd-bench_amd", "weights", "unet_channel.bin")
    modes = {"fp64": dict(base, precision="fp64"),
             "mixed": dict(base, precision="mixed"),
             "learned": dict(base, precision="mixed", warmstart="unet", unet_weights=weights)}
    out = {}
    cells = 768.0 * 192 * 192
    for name, opts in modes.items():
        c = A.Case(lib, options=opts, **channel_params())
        tr = (C.c_double * (4 * max(warmup, steps)))()
        lib.nusifRunSteps(C.byref(c.d), C.byref(c.s), warmup, tr)
        tried0, taken0 = C.c_longlong(), C.c_longlong()
        lib.nusifWarmstartCounts(C.byref(c.s), C.byref(tried0), C.byref(taken0))
        lib.nusifDeviceSynchronize()
        t0 = time.perf_counter()
        lib.nusifRunSteps(C.byref(c.d), C.byref(c.s), steps, tr)
        lib.nusifDeviceSynchronize()
        dt = time.perf_counter() - t0
        tried, taken = C.c_longlong(), C.c_longlong()
        lib.nusifWarmstartCounts(C.byref(c.s), C.byref(tried), C.byref(taken))
        c.close()
        out[name] = {"value": steps / dt, "mlups": cells * steps / dt * 1e-6,
                     "cycles_per_step": sum(tr[4 * q + 3] for q in range(steps)) / steps,
                     "max_rel_residual": max(tr[4 * q + 2] for q in range(steps))}
        if name == "learned":
            out[name]["warmstart_guess_taken_fraction"] = ((taken.value - taken0.value)
                                                           / max(tried.value - tried0.value, 1))
        log("[bench] channel 768x192x192 %s: %d steps in %.3f s" % (name, steps, dt))
    out["speedup_learned_vs_fp64_mg"] = out["learned"]["value"] / out["fp64"]["value"]
    out["speedup_mixed_vs_fp64_mg"] = out["mixed"]["value"] / out["fp64"]["value"]
    out["unit"] = "time steps/s"
    out["workload"] = ("channel 768x192x192, Re 500, inflow u=1 with seeded 1%% noise, pressure to 1e-6 relative "
                       "residual per step, V(2,2) 6 levels, omega 1.3; %d steps after %d" % (steps, warmup))
    return out


def cpu_rb(case_gpu, n, gpu_iters, itermax, iters=40):
    """Red-black SOR on one host core on the same state, one solve capped at
    `iters` iterations: the reference's solver rate. The reference itself
    (built from its sources with its own flags into oracle/_ref) when it is
    there and the GPU's solves of this state ran to itermax (the reference
    reports no iteration count; a solve that cannot stop before itermax runs
    exactly the cap), else the oracle's restatement."""
    ref = A.ref_lib_path("rb_fast")
    use_ref = os.path.exists(ref) and gpu_iters >= itermax
    lib = A.load(ref) if use_ref else A.load(A.oracle_lib_path("rb_fast"))
    o = A.Case(lib, **dict(cavity_params(n), itermax=iters))
    o.set_fields(**case_gpu.fields())
    d = o.d
    lib.computeTimestep(C.byref(d))
    for fn in ("setBoundaryConditions", "setSpecialBoundaryCondition", "computeFG", "computeRHS"):
        getattr(lib, fn)(C.byref(d))
    t0 = time.perf_counter()
    lib.solve(C.byref(o.s), d.p, d.rhs)
    t = time.perf_counter() - t0
    it = iters if use_ref else o.iterations()
    o.close()
    log("[bench] cpu rb (%s): %d iterations in %.2f s" % ("reference" if use_ref else "port", it, t))
    what = ("the reference's own solver-rb.c and discretization.c (-O3 -ffast-math, its flags)" if use_ref
            else "oracle C restatement (-O3 -ffast-math)")
    return {"solver_mlups": it * float(n) ** 3 / t * 1e-6, "iterations": it, "seconds": t, "cores": 1,
            "kind": "reference" if use_ref else "port",
            "sample": "one solve of the %d^3 cavity from the GPU state, capped at %d iterations; %s, one "
                      "host core" % (n, iters, what)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mode", default="learned", choices=["learned", "mixed", "fp64"])
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-compare", action="store_true", help="skip the fp64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only time a device-to-device copy of MB megabytes (counter calibration)")
    ap.add_argument("--opt", action="append", default=[], metavar="KEY=VALUE",
                    help="extra solver option (nusifSetOption), e.g. mg_graph=0")
    a = ap.parse_args()
    for kv in a.opt:
        k, v = kv.split("=", 1)
        EXTRA_OPTS[k] = v
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        sys.exit("bench.py: this build runs on one GPU only (the solver is a single-domain "
                 "MI355X path; --gpus %d is not supported)" % a.gpus)
    lib = A.load(A.gpu_lib_path("rb"))
    name = C.create_string_buffer(256)
    if lib.nusifDeviceName(name, 256) < 1:
        sys.exit("bench.py: no GPU visible to the HIP library")
    log("[bench] device", name.value.decode())
    if a.copy_probe:
        nbytes = a.copy_probe * 1 << 20
        ms = lib.nusifCopyProbe(nbytes, 20)
        print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": ms,
                          "GBs_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return
    n = a.size
    cells = float(n) ** 3
    case, res = run_mode(lib, a.mode, n, a.warmup, a.steps)
    sps = a.steps / res["seconds"]
    out = {
        "metric": METRIC, "value": sps, "unit": "time steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * res["seconds"] / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if a.mode != "fp64" else "f64", "data": "synthetic",
        "mlups": cells * sps * 1e-6,
        "config": {"workload": "lid-driven cavity %d^3 cells, Re=1000, lid u=1, tau=0.5, "
                               "pressure to 1e-6 relative residual per step, V(2,2) 6 levels, "
                               "red-black smoother omega 1.3; mode %s" % (n, a.mode),
                   "grid": [n, n, n], "mode": a.mode, "device": name.value.decode(),
                   "options": dict(EXTRA_OPTS)},
        "cycles_per_step": sum(res["cycles"]) / len(res["cycles"]),
        "max_rel_residual": max(res["rel"]),
    }
    out.update(solver_rate(res, cells))
    if res["warmstart_taken"] is not None:
        out["warmstart_guess_taken_fraction"] = res["warmstart_taken"]
    dom = DOMINANT[a.mode]
    k = res["kernels"].get(dom)
    if k and res["sweep_us"] > 0:
        ach = k["bytes_per_launch"] / (res["sweep_us"] * 1e-6) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                           "traffic": measured_traffic(dom, n),
                           "us_per_launch": res["sweep_us"],
                           "us_per_launch_event_timed_in_steps": k["us_per_launch"],
                           "algorithmic_bytes_per_launch": k["bytes_per_launch"]}
    out["kernels"] = {t: {"us": round(v["us_per_launch"], 2), "launches": v["launches"],
                          "GBs": round(v["bytes_per_launch"] / (v["us_per_launch"] * 1e-6) / 1e9, 1)}
                      for t, v in res["kernels"].items()}
    if not a.no_compare and a.mode != "fp64":
        ref_case, ref = run_mode(lib, "fp64", n, a.warmup, a.steps)
        ref_case.close()
        f64 = a.steps / ref["seconds"]
        out["fp64_mg"] = {"value": f64, "mlups": cells * f64 * 1e-6,
                          "cycles_per_step": sum(ref["cycles"]) / len(ref["cycles"])}
        out["fp64_mg"].update(solver_rate(ref, cells))
        out["speedup_vs_fp64_mg"] = sps / f64
        if a.mode == "learned":  # the same solver without the network
            mx_case, mx = run_mode(lib, "mixed", n, a.warmup, a.steps)
            mx_case.close()
            vm = a.steps / mx["seconds"]
            out["fp32_mg_without_warmstart"] = {
                "value": vm, "mlups": cells * vm * 1e-6,
                "cycles_per_step": sum(mx["cycles"]) / len(mx["cycles"]),
                "speedup_vs_fp64_mg": vm / f64}
            out["fp32_mg_without_warmstart"].update(solver_rate(mx, cells))
    if not a.no_compare:
        out["channel_768"] = run_channel(lib)
        out["rb_sor"] = run_rb(lib, case, n)
        if not a.no_cpu:
            out["rb_sor"]["cpu_baseline"] = cpu_rb(case, n, out["rb_sor"]["iterations_per_step"],
                                                   cavity_params(n)["itermax"])
    # what plain streaming kernels reach on this device (256 MiB arrays): the
    # practical ceiling under the 8 TB/s vendor peak
    ms = (C.c_double * 3)()
    lib.nusifStreamProbe(256 << 20, 20, ms)
    nb = float(256 << 20)
    out["hbm_probe"] = {"copy_GBs": 2 * nb / (ms[0] * 1e-3) / 1e9,
                        "two_reads_one_write_GBs": 3 * nb / (ms[1] * 1e-3) / 1e9,
                        "read_GBs": nb / (ms[2] * 1e-3) / 1e9, "array_bytes": int(nb)}
    if "roofline" in out:
        out["roofline"]["frac_of_probe_2r1w"] = (out["roofline"]["achieved"]
                                                 / out["hbm_probe"]["two_reads_one_write_GBs"])
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, n)
    case.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
