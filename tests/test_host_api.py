"""Host side of the C API, without a GPU: exported symbols, the .par parser,
result collection and the VTK writer.

Each check compares the HIP library's host functions with fixed expectations
and, where the reference has been compiled from its own sources into
oracle/_ref, with the reference's functions on the same input:
  parameter.c:13-100 (initParameter/readParameter), comm.c:436-476
  (sequential commCollectResult), vtkWriter-seq.c:12-146 (VTK legacy writer,
  ASCII and big-endian BINARY).
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import nusif_api as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gpu_lib(kind="rb"):
    path = A.gpu_lib_path(kind)
    if not os.path.exists(path):
        pytest.fail("HIP library %s is not built (run __graft_entry__.build())" % path)
    return A.load(path)


def _ref_lib():
    path = A.ref_lib_path("rb")
    if not os.path.exists(path):
        pytest.skip("reference build oracle/_ref absent")
    return A.load(path)


def _declared_functions():
    names = set()
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
        text = re.sub(r"typedef[^;]*;", " ", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(([^;{}()]*)\)\s*;", text):
            names.add(m.group(1))
    return names


@pytest.mark.parametrize("kind", ["rb", "mg"])
def test_library_exports_every_declared_function(kind):
    declared = _declared_functions()
    assert {"initDiscretization", "computeFG", "solve", "initSolver", "readParameter",
            "vtkOpen", "commCollectResult", "nusifRunSteps"} <= declared
    path = A.gpu_lib_path(kind)
    assert os.path.exists(path), path
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = sorted(declared - exported)
    assert not missing, "declared in include/ but not exported: %s" % missing
    lib = C.CDLL(path)  # loads without a GPU
    for name in declared:
        getattr(lib, name)


PAR = """# parser corner cases
name dcavity        # trailing comment after the value
imax 12\t\t\t# tabs after the value
jmax\t10
kmax 8
reynolds 250.0
re 100.0
te 2.5
dtx 0.25
tau 0.5
gamma 0.9
omg 1.7
eps 1e-4
itermax 40
levels 3
presmooth 2
postsmooth 1
gx 0.1
gy -0.2
gz 0.0
bcLeft 3
bcRight 3
bcBottom 1
bcTop 2
bcFront 1
bcBack 4
u_init 1.0
v_init 0
w_init 0
p_init 0.5
xlength 4.0
ylength 1.0
zlength 1.0
nsteps 7
lid 1.5
"""

EXPECTED = dict(imax=12, kmax=8, re=100.0, te=2.5, dt=0.25, tau=0.5, gamma=0.9, omg=1.7,
                eps=1e-4, itermax=40, levels=3, presmooth=2, postsmooth=1, gx=0.1, gy=-0.2,
                gz=0.0, bcLeft=3, bcRight=3, bcBottom=1, bcTop=2, bcFront=1, bcBack=4,
                u_init=1.0, v_init=0.0, w_init=0.0, p_init=0.5, xlength=4.0, ylength=1.0,
                zlength=1.0, name=b"dcavity")


def _parse(lib, path):
    prm = A.Parameter()
    lib.initParameter(C.byref(prm))
    lib.readParameter(C.byref(prm), path.encode())
    return prm


def test_parameter_parser_corner_cases(tmp_path):
    """Prefix keys (reynolds -> re, dtx -> dt), the later line winning, '#'
    comments, split on ' ' only (a tab-separated line has no value and is
    skipped, so jmax keeps its default), and the added exact-match keys."""
    f = tmp_path / "case.par"
    f.write_text(PAR)
    lib = _gpu_lib()
    default = A.Parameter()
    lib.initParameter(C.byref(default))
    prm = _parse(lib, str(f))
    for k, v in EXPECTED.items():
        assert getattr(prm, k) == v, k
    assert prm.jmax == default.jmax
    assert lib.nusifStepLimit(C.byref(prm)) == 7


def test_parameter_parser_matches_reference(tmp_path):
    f = tmp_path / "case.par"
    f.write_text(PAR)
    ref = _ref_lib()
    ours = _parse(_gpu_lib(), str(f))
    theirs = _parse(ref, str(f))
    for name, _ in A.Parameter._fields_:
        assert getattr(ours, name) == getattr(theirs, name), name


def _padded_fields(rng, imax, jmax, kmax):
    n = (imax + 2) * (jmax + 2) * (kmax + 2)
    return [rng.uniform(-3.0, 3.0, n) for _ in range(4)]


def _collect(lib, fields, imax, jmax, kmax):
    comm = A.CommType()
    lib.commInit(C.byref(comm), 0, None)
    lib.commPartition(C.byref(comm), kmax, jmax, imax)
    nout = imax * jmax * kmax
    outs = [np.zeros(nout) for _ in range(4)]
    dp = lambda a: a.ctypes.data_as(A._DP)  # noqa: E731
    u, v, w, p = fields
    lib.commCollectResult(C.byref(comm), dp(outs[0]), dp(outs[1]), dp(outs[2]), dp(outs[3]),
                          dp(u), dp(v), dp(w), dp(p), kmax, jmax, imax)
    return outs


def test_collect_result_averages_faces_to_centres():
    imax, jmax, kmax = 5, 4, 3
    rng = np.random.default_rng(3)
    fields = _padded_fields(rng, imax, jmax, kmax)
    ug, vg, wg, pg = _collect(_gpu_lib(), fields, imax, jmax, kmax)
    u, v, w, p = (a.reshape(kmax + 2, jmax + 2, imax + 2) for a in fields)
    I = slice(1, imax + 1)
    J = slice(1, jmax + 1)
    K = slice(1, kmax + 1)
    np.testing.assert_array_equal(pg, p[K, J, I].ravel())
    np.testing.assert_array_equal(ug, ((u[K, J, I] + u[K, J, 0:imax]) / 2.0).ravel())
    np.testing.assert_array_equal(vg, ((v[K, J, I] + v[K, 0:jmax, I]) / 2.0).ravel())
    np.testing.assert_array_equal(wg, ((w[K, J, I] + w[0:kmax, J, I]) / 2.0).ravel())
    ref = os.path.exists(A.ref_lib_path("rb"))
    if ref:
        theirs = _collect(_ref_lib(), fields, imax, jmax, kmax)
        for a, b in zip((ug, vg, wg, pg), theirs):
            np.testing.assert_array_equal(a, b)


class VtkOptions(C.Structure):
    _fields_ = [("grid", C.POINTER(A.Grid)), ("fh", C.c_void_p), ("fmt", C.c_int),
                ("comm", A.CommType)]


class VtkVector(C.Structure):
    _fields_ = [("u", A._DP), ("v", A._DP), ("w", A._DP)]


def _write_vtk(lib, problem, fmt, grid, s, u, v, w):
    lib.vtkOpen.argtypes = [C.POINTER(VtkOptions), C.c_char_p]
    lib.vtkScalar.argtypes = [C.POINTER(VtkOptions), C.c_char_p, A._DP]
    lib.vtkVector.argtypes = [C.POINTER(VtkOptions), C.c_char_p, VtkVector]
    lib.vtkClose.argtypes = [C.POINTER(VtkOptions)]
    o = VtkOptions()
    o.grid = C.pointer(grid)
    o.fmt = fmt
    dp = lambda a: a.ctypes.data_as(A._DP)  # noqa: E731
    lib.vtkOpen(C.byref(o), problem.encode())
    lib.vtkScalar(C.byref(o), b"pressure", dp(s))
    lib.vtkVector(C.byref(o), b"velocity", VtkVector(dp(u), dp(v), dp(w)))
    lib.vtkClose(C.byref(o))
    with open(problem + ".vtk", "rb") as fh:
        return fh.read()


@pytest.mark.parametrize("fmt", [0, 1], ids=["ascii", "binary"])
def test_vtk_writer(fmt, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    imax, jmax, kmax = 4, 3, 2
    grid = A.Grid(0.25, 1.0 / 3.0, 0.5, imax, jmax, kmax, 1.0, 1.0, 1.0)
    rng = np.random.default_rng(11)
    n = imax * jmax * kmax
    s, u, v, w = (rng.uniform(-2.0, 2.0, n) for _ in range(4))
    ours = _write_vtk(_gpu_lib(), "ours", fmt, grid, s, u, v, w)
    head = ours.split(b"SCALARS")[0].decode()
    assert head == ("# vtk DataFile Version 3.0\nPAMPI cfd solver output\n%s\n"
                    "DATASET STRUCTURED_POINTS\nDIMENSIONS 4 3 2\n"
                    "ORIGIN 0.125000 0.166667 0.250000\n"
                    "SPACING 0.250000 0.333333 0.500000\nPOINT_DATA 24\n"
                    % ("ASCII" if fmt == 0 else "BINARY"))
    body = ours[ours.index(b"LOOKUP_TABLE default\n") + len(b"LOOKUP_TABLE default\n"):]
    if fmt == 0:
        vals = np.array([float(x) for x in body.split(b"VECTORS")[0].split()])
        np.testing.assert_allclose(vals, s, atol=5e-7)
    else:
        vals = np.frombuffer(body[:8 * n], dtype=">f8")
        np.testing.assert_array_equal(vals, s)
        vec = body[8 * n + 1:]
        assert vec.startswith(b"VECTORS velocity double\n")
        vv = np.frombuffer(vec[len(b"VECTORS velocity double\n"):][:24 * n], dtype=">f8")
        np.testing.assert_array_equal(vv.reshape(n, 3), np.stack([u, v, w], axis=1))
    if os.path.exists(A.ref_lib_path("rb")):
        theirs = _write_vtk(_ref_lib(), "theirs", fmt, grid, s, u, v, w)
        assert ours == theirs


def test_set_option_accepts_documented_keys():
    """nusifSetOption (host code only) accepts every key include/nusif.h
    documents, with a valid value, rejects bad values (-2) and unknown keys
    (-1); nusifStepLimit reads nsteps back."""
    lib = _gpu_lib()
    doc = open(os.path.join(ROOT, "include", "nusif.h")).read()
    block = doc[doc.index("Options beyond the reference"):doc.index("int nusifSetOption")]
    keys = set(re.findall(r"^ \*   ([a-z_0-9]+)\s", block, flags=re.M))
    valid = {"nsteps": "7", "solver": "gmg", "precision": "mixed", "mg_tol": "1e-8",
             "mg_maxcycles": "50", "mg_omega": "1.3", "mg_project": "1", "mg_tail": "0", "mg_fuse_prolong": "0", "bc_fused": "0",
             "mg_graph": "0", "mg_tile_rows": "8", "mg_inner": "0", "mg_speculate": "0",
             "lid": "1.5", "u_inflow": "1.0", "inflow_noise": "0.01", "seed": "3",
             "warmstart": "none", "unet_weights": "/tmp/w.bin", "unet_extent": "32", "fg_rhs": "0",
             "rb_fused": "0", "mg_graph_cache": "4", "verbose": "2", "unet_accept_level": "1",
             "unet_graph": "1", "warmstart_skip": "2"}
    assert keys == set(valid), sorted(keys ^ set(valid))
    prm = A.Parameter()
    lib.initParameter(C.byref(prm))
    for k, v in valid.items():
        assert lib.nusifSetOption(C.byref(prm), k.encode(), v.encode()) == 0, k
    assert lib.nusifStepLimit(C.byref(prm)) == 7
    bad = {"solver": "cg", "precision": "fp16", "mg_tile_rows": "5", "mg_inner": "-1",
           "warmstart": "maybe", "mg_tol": "tight", "verbose": "3", "mg_graph_cache": "0",
           "unet_accept_level": "-1", "warmstart_skip": "-1"}
    for k, v in bad.items():
        assert lib.nusifSetOption(C.byref(prm), k.encode(), v.encode()) == -2, k
    assert lib.nusifSetOption(C.byref(prm), b"no_such_key", b"1") == -1
