"""The drop-in, end to end: the reference's own main.c, compiled unchanged
and linked against libnusif_<solver>.so (oracle/Makefile target `dropin`),
runs a .par file on the GPU; the reference executable built from its own
sources runs the same file on the CPU. Their per-step output (main.c:80-87
TIME lines), residual.dat (progress.c:63-66) and the VTK result
(comm.c:436-476 + vtkWriter-seq.c) must agree.

Both executables live in oracle/_ref (built here from the reference sources,
never committed); the test is skipped where they were not built.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")

# values are followed by a space or a comment, as in the reference's own
# .par files: the parser splits on ' ' only, so "name dcavity\n" would keep
# the newline in the problem name (and then apply no lid)
PAR = """name dcavity        # problem
bcLeft 1
bcRight 1
bcBottom 1
bcTop 1
bcFront 1
bcBack 1
gx 0.0
gy 0.0
gz 0.0
re {re}
u_init 0.0
v_init 0.0
w_init 0.0
p_init 0.0
xlength 1.0
ylength 1.0
zlength 1.0
imax {n}
jmax {n}
kmax {n}
te {te}
dt 0.02
tau 0.5
levels {levels}
presmooth 5
postsmooth 5
itermax 500
eps 0.001
omg 1.7
gamma 0.9
"""


def _run(exe, workdir, par):
    os.makedirs(workdir, exist_ok=True)
    with open(os.path.join(workdir, "case.par"), "w") as fh:
        fh.write(par)
    print("running", os.path.basename(exe), flush=True)
    r = subprocess.run([exe, "case.par"], cwd=workdir, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _vtk(path):
    with open(path) as fh:
        lines = fh.read().splitlines()
    head = [l for l in lines if l[:1].isalpha() or l.startswith("#")]
    nums = np.array([float(x) for l in lines if not (l[:1].isalpha() or l.startswith("#"))
                     for x in l.split()])
    return head, nums


@pytest.mark.gpu
@pytest.mark.parametrize("solver,n,re,te,levels", [("rb", 32, 10.0, 0.02, 1),
                                                  ("mg", 16, 100.0, 0.12, 2)])
def test_reference_main_on_gpu_library(solver, n, re, te, levels, tmp_path, gpu_ok):
    cpu = os.path.join(REF, "NusifSolver-%s" % solver)
    gpu = os.path.join(REF, "NusifSolver-gpu-%s" % solver)
    if not (os.path.exists(cpu) and os.path.exists(gpu)):
        pytest.skip("reference executables not built (oracle/Makefile ref, dropin)")
    par = PAR.format(n=n, re=re, te=te, levels=levels)
    out_c = _run(cpu, str(tmp_path / "cpu"), par)
    out_g = _run(gpu, str(tmp_path / "gpu"), par)
    t_c = [l for l in out_c.splitlines() if l.startswith("TIME")]
    t_g = [l for l in out_g.splitlines() if l.startswith("TIME")]
    assert len(t_c) > 5 and t_c == t_g
    assert os.path.exists(str(tmp_path / "cpu" / "dcavity.vtk")), os.listdir(str(tmp_path / "cpu"))
    res_c = np.loadtxt(str(tmp_path / "cpu" / "residual.dat"), delimiter=",", ndmin=2)
    res_g = np.loadtxt(str(tmp_path / "gpu" / "residual.dat"), delimiter=",", ndmin=2)
    assert res_c.shape == res_g.shape
    np.testing.assert_allclose(res_g, res_c, rtol=0, atol=2e-6)
    h_c, v_c = _vtk(str(tmp_path / "cpu" / "dcavity.vtk"))
    h_g, v_g = _vtk(str(tmp_path / "gpu" / "dcavity.vtk"))
    assert h_c == h_g
    assert v_c.shape == v_g.shape == (4 * n ** 3,)
    # %f output: six decimals, so values equal to 1e-10 may differ by one unit
    np.testing.assert_allclose(v_g, v_c, rtol=0, atol=1.5e-6)
