"""Trains the warm-start U-Net on the GPU and measures what it gains.

  python cfd-bench_amd/train_unet.py [--grid 128 [256 ...]] [--runs R] [--steps S] [--epochs E]
                                      [--out cfd-bench_amd/weights/unet_cavity.bin]
  python cfd-bench_amd/train_unet.py --channel 768 192 192 --init cfd-bench_amd/weights/unet_cavity.bin
                                      [--runs R] [--steps S] [--epochs E] [--out .../unet_channel.bin]

Training pairs are generated by the solver itself (lid-driven cavities of
grid^3 cells, run r on the r-th grid in turn, seeded Re ~ U[10, 1000], lid
speed ~ U[0.5, 2], gmg with FP32 V-cycles in FP64 defect correction to 1e-6),
then the network is trained with Adam, all inside the library
(nusifUnetTrainCavityGrids). The evaluation runs held-out cavities (Re and
lid not drawn from the training seed) on every grid with and without the
warm start and prints one JSON line: V-cycles per step, ms per step, and how
often the network's guess beat the previous-step pressure.

With --channel, the pairs come from channels of that many cells (the canal
problem on a 4 x 1 x 1 domain, inflow u = 1 with a seeded 1% perturbation, a
new seed per run, Re ~ U[100, 1000]) and training starts from --init
(fine-tuning; nusifUnetTrainChannel); the evaluation runs the channel at
Re 500 with an inflow seed no run drew.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import nusif_api as A  # noqa: E402


def evaluate(lib, weights, grid, re, lid, warm, steps, force=False, extent=64):
    prm = dict(name="dcavity", re=re, imax=grid, jmax=grid, kmax=grid, tau=0.5, te=1e9, levels=6,
               presmooth=2, postsmooth=2)
    opts = {"solver": "gmg", "precision": "mixed", "mg_tol": 1e-6, "lid": lid}
    if warm:
        opts.update({"warmstart": "unet_force" if force else "unet", "unet_weights": weights,
                     "unet_extent": extent})
    case = A.Case(lib, options=opts, **prm)
    tr = (C.c_double * (4 * (warmup + steps)))()
    lib.nusifRunSteps(C.byref(case.d), C.byref(case.s), warmup, tr)
    used, cycles = 0, 0.0
    res = (C.c_double * 4)()
    ratios = []
    lib.nusifDeviceSynchronize()
    t0 = time.perf_counter()
    for q in range(steps):
        lib.nusifRunSteps(C.byref(case.d), C.byref(case.s), 1, tr)
        cycles += tr[3]
        used += max(lib.nusifWarmstartUsed(C.byref(case.s)), 0)
        if warm:
            lib.nusifWarmstartResiduals(C.byref(case.s), res)
            ratios.append((res[1] / res[0], res[3] / res[2]))
    lib.nusifDeviceSynchronize()
    dt = time.perf_counter() - t0
    case.close()
    out = dict(cycles_per_step=cycles / steps, ms_per_step=1e3 * dt / steps,
               guess_used_fraction=used / steps if warm else None)
    if ratios:
        out["guess_over_prev_residual_fine"] = sum(r[0] for r in ratios) / len(ratios)
        out["guess_over_prev_residual_coarse"] = sum(r[1] for r in ratios) / len(ratios)
    return out


warmup = 10


def evaluate_channel(lib, weights, ext, re, seed, warm, steps, extent=64, force=False):
    prm = dict(name="canal", xlength=4.0, ylength=1.0, zlength=1.0, bcLeft=3, bcRight=3, bcBottom=1, bcTop=1,
               bcFront=1, bcBack=1, u_init=1.0, re=re, imax=ext[0], jmax=ext[1], kmax=ext[2], tau=0.5, te=1e9,
               gamma=0.9, levels=6, presmooth=2, postsmooth=2)
    opts = {"solver": "gmg", "precision": "mixed", "mg_tol": 1e-6, "mg_project": 1, "u_inflow": 1.0,
            "inflow_noise": 0.01, "seed": seed, "mg_omega": 1.3}
    if warm:
        opts.update({"warmstart": "unet_force" if force else "unet", "unet_weights": weights,
                     "unet_extent": extent})
    case = A.Case(lib, options=opts, **prm)
    tr = (C.c_double * (4 * (warmup + steps)))()
    lib.nusifRunSteps(C.byref(case.d), C.byref(case.s), warmup, tr)
    ratio = None
    if warm:  # the guess's restricted and fine residual against the previous p's, one step
        lib.nusifRunSteps(C.byref(case.d), C.byref(case.s), 1, tr)
        res = (C.c_double * 4)()
        lib.nusifWarmstartResiduals(C.byref(case.s), res)
        ratio = {"fine": res[1] / res[0], "restricted": res[3] / res[2]}
    tried0, taken0, tried, taken = C.c_longlong(), C.c_longlong(), C.c_longlong(), C.c_longlong()
    lib.nusifWarmstartCounts(C.byref(case.s), C.byref(tried0), C.byref(taken0))
    lib.nusifDeviceSynchronize()
    t0 = time.perf_counter()
    lib.nusifRunSteps(C.byref(case.d), C.byref(case.s), steps, tr)
    lib.nusifDeviceSynchronize()
    dt = time.perf_counter() - t0
    lib.nusifWarmstartCounts(C.byref(case.s), C.byref(tried), C.byref(taken))
    case.close()
    out = dict(cycles_per_step=sum(tr[4 * q + 3] for q in range(steps)) / steps, ms_per_step=1e3 * dt / steps)
    if warm:
        out["guess_used_fraction"] = (taken.value - taken0.value) / max(tried.value - tried0.value, 1)
        out["guess_over_prev_residual"] = ratio
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, nargs="+", default=[128])
    ap.add_argument("--extent", type=int, default=64, help="network grid extent (unet_extent)")
    ap.add_argument("--runs", type=int, default=16)
    ap.add_argument("--skip", type=int, default=5)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "unet_cavity.bin"))
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--eval-steps", type=int, default=40)
    ap.add_argument("--channel", type=int, nargs=3, metavar=("IMAX", "JMAX", "KMAX"),
                    help="train on channel pairs of this grid (fine-tuning from --init)")
    ap.add_argument("--init", default="", help="weights to start from (--channel)")
    a = ap.parse_args()
    lib = A.load(A.gpu_lib_path("rb"))
    if lib.nusifDeviceName(None, 0) < 1:
        sys.exit("train_unet.py: no GPU visible to the HIP library")
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    if a.channel:
        report = {"channel": a.channel, "init": os.path.basename(a.init)}
        if not a.eval_only:
            losses = (C.c_double * a.epochs)()
            t0 = time.perf_counter()
            n = lib.nusifUnetTrainChannel(a.init.encode(), a.out.encode(), a.channel[0], a.channel[1],
                                          a.channel[2], 4.0, a.extent, a.runs, a.steps, a.skip, a.epochs, a.batch,
                                          a.lr, a.seed, losses)
            report.update(pairs=n, train_seconds=time.perf_counter() - t0, loss_per_epoch=list(losses))
        # held out: an inflow seed no training run drew (they use seed * 1000 + r)
        for w, name in ((a.out, "with_unet"), (a.init, "with_init_weights")):
            if w and os.path.exists(w):
                report[name] = evaluate_channel(lib, w, a.channel, 500.0, 99991, True, a.eval_steps,
                                                extent=a.extent)
                report[name + "_always"] = evaluate_channel(lib, w, a.channel, 500.0, 99991, True, a.eval_steps,
                                                            extent=a.extent, force=True)
        report["without"] = evaluate_channel(lib, a.out, a.channel, 500.0, 99991, False, a.eval_steps)
        print(json.dumps(report), flush=True)
        return
    report = {"grid": a.grid}
    if not a.eval_only:
        losses = (C.c_double * a.epochs)()
        t0 = time.perf_counter()
        grids = (C.c_int * len(a.grid))(*a.grid)
        n = lib.nusifUnetTrainCavityGrids(a.out.encode(), grids, len(a.grid), a.extent, a.runs, a.steps,
                                          a.skip, a.epochs, a.batch, a.lr, a.seed, losses)
        report.update(pairs=n, train_seconds=time.perf_counter() - t0, loss_per_epoch=list(losses))
    # held-out case: Re and lid chosen outside the training draws
    for grid in a.grid:
        for re, lid in ((500.0, 1.0), (150.0, 1.6)):
            base = evaluate(lib, a.out, grid, re, lid, False, a.eval_steps)
            warm = evaluate(lib, a.out, grid, re, lid, True, a.eval_steps, extent=a.extent)
            force = evaluate(lib, a.out, grid, re, lid, True, a.eval_steps, force=True, extent=a.extent)
            report["eval_%d_re%g_lid%g" % (grid, re, lid)] = {"without": base, "with_unet": warm,
                                                              "always_guess": force}
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
