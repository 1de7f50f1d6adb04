"""Builds the MI355X NuSiF library in-tree with hipcc (gfx950 only).

Outputs (git-ignored, they travel with the tree):
  cfd-bench_amd/lib/libnusif_rb.so   reference API, default solver red-black SOR
  cfd-bench_amd/lib/libnusif_mg.so   reference API, default solver MG (reference semantics)
  cfd-bench_amd/bin/nusif            driver (reference main.c loop + nsteps), links libnusif_rb
Usage: python cfd-bench_amd/build.py [-j N] [--verbose]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "build")
LIB = os.path.join(HERE, "lib")
BIN = os.path.join(HERE, "bin")
INC = [os.path.join(ROOT, "include"), CSRC]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu99", "-Wall", "-Wno-unused-result"]

HIP_SOURCES = ["context.hip", "discretization.hip", "sor.hip", "gmg.hip", "unet.hip", "warmstart.hip"]
SOLVER_VARIANTS = {"rb": "SOLVER_RB", "mg": "SOLVER_MG"}


def run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise SystemExit("build failed: " + " ".join(cmd[:3]) + " ... " + cmd[-1])
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if out and verbose:
        print(out)
    return r


def hip_obj(src, out, defines, verbose):
    inc = sum((["-I", d] for d in INC), [])
    run([HIPCC] + HIP_FLAGS + inc + defines + ["-c", os.path.join(CSRC, src), "-o", out], verbose)
    return out


def c_obj(src, out, verbose):
    inc = sum((["-I", d] for d in INC), [])
    run(["gcc"] + C_FLAGS + inc + ["-c", os.path.join(CSRC, src), "-o", out], verbose)
    return out


def build(jobs=8, verbose=False):
    for d in (OBJ, LIB, BIN):
        os.makedirs(d, exist_ok=True)
    tasks = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for s in HIP_SOURCES:
            tasks.append(ex.submit(hip_obj, s, os.path.join(OBJ, s + ".o"), [], verbose))
        for v, kind in SOLVER_VARIANTS.items():
            tasks.append(ex.submit(hip_obj, "solver.hip", os.path.join(OBJ, "solver_%s.o" % v),
                                   ["-DNUSIF_DEFAULT_SOLVER=" + kind], verbose))
        tasks.append(ex.submit(c_obj, "host_api.c", os.path.join(OBJ, "host_api.o"), verbose))
        tasks.append(ex.submit(c_obj, "nusif_main.c", os.path.join(OBJ, "nusif_main.o"), verbose))
        for t in tasks:
            t.result()
    common = [os.path.join(OBJ, s + ".o") for s in HIP_SOURCES] + [os.path.join(OBJ, "host_api.o")]
    for v in SOLVER_VARIANTS:
        out = os.path.join(LIB, "libnusif_%s.so" % v)
        tmp = out + ".tmp"
        run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", tmp]
            + common + [os.path.join(OBJ, "solver_%s.o" % v)], verbose)
        os.replace(tmp, out)
    exe = os.path.join(BIN, "nusif")
    run([HIPCC, "-o", exe + ".tmp", os.path.join(OBJ, "nusif_main.o"), "-L" + LIB, "-lnusif_rb",
         "-Wl,-rpath,$ORIGIN/../lib"], verbose)
    os.replace(exe + ".tmp", exe)
    return [os.path.join(LIB, "libnusif_%s.so" % v) for v in SOLVER_VARIANTS]


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    for p in build(a.j, a.verbose):
        print("built", os.path.relpath(p, ROOT))
