"""The allocation-free red-black solve (nusifSolveRBNoAlloc) against the
reference-API solve() with the same kernels (rb_fused 0): two calls in a row
on caller-owned device arrays, the second on a smaller grid (the fixed
scratch keeps the first call's contents, which the second must not read),
each equal to solve() bit for bit: p (every padded cell), the returned
residual and the iteration count. Device memory comes from the HIP runtime
directly (hipMalloc through ctypes), as such a caller's would."""
import ctypes as C
import os

import numpy as np
import pytest

import make_golden as G
import nusif_api as A
from conftest import lib

pytestmark = pytest.mark.gpu


def _hip():
    lib("gpu_rb")  # the HIP runtime the library itself loaded
    for name in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.fail("libamdhip64.so not found")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def _to_device(hip, a):
    ptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(ptr), a.nbytes) == 0
    assert hip.hipMemcpy(ptr, a.ctypes.data, a.nbytes, H2D) == 0
    return ptr


def _solve_pair(shape, hip):
    """solve() on the state before a cavity step's solve, and the
    allocation-free solve on copies of the same p and rhs"""
    prm = dict(G.TRAJ["c1_rb"][1], imax=shape[0], jmax=shape[1], kmax=shape[2])
    c = A.Case(lib("gpu_rb"), options={"rb_fused": 0}, **prm)
    c.run(3)
    for fn in ("computeTimestep", "setBoundaryConditions", "setSpecialBoundaryCondition", "computeFG",
               "computeRHS"):
        c.call(fn)
    f = c.fields()
    p0, rhs0 = np.ascontiguousarray(f["p"]), np.ascontiguousarray(f["rhs"])
    res_ref = c.lib.solve(C.byref(c.s), c.d.p, c.d.rhs)
    it_ref = c.iterations()
    p_ref = c.fields()["p"]
    g = c.d.grid
    args = (g.dx, g.dy, g.dz, c.prm.omg, c.prm.eps, c.prm.itermax)
    c.close()
    dp, dr = _to_device(hip, p0), _to_device(hip, rhs0)
    res, it = C.c_double(), C.c_int()
    rc = lib("gpu_rb").nusifSolveRBNoAlloc(dp, dr, shape[0], shape[1], shape[2], *args, C.byref(res),
                                           C.byref(it))
    p = np.empty_like(p0)
    assert hip.hipMemcpy(p.ctypes.data, dp, p.nbytes, D2H) == 0
    hip.hipFree(dp)
    hip.hipFree(dr)
    return rc, (p, res.value, it.value), (p_ref, res_ref, it_ref)


def test_no_alloc_solve_equals_solve(gpu_ok):
    hip = _hip()
    for shape in ((40, 36, 28), (24, 20, 16)):  # the second call smaller
        print("no-alloc rb solve", shape, flush=True)
        rc, (p, res, it), (p_ref, res_ref, it_ref) = _solve_pair(shape, hip)
        assert rc == 0
        print("iterations %d, residual %.6e" % (it, res))
        assert it == it_ref and 0 < it
        assert res == res_ref
        np.testing.assert_array_equal(p, p_ref)


def test_no_alloc_rejects_what_does_not_fit(gpu_ok):
    """checked before any device work: a grid beyond the fixed scratch, and
    bad arguments"""
    f = lib("gpu_rb").nusifSolveRBNoAlloc
    res, it = C.c_double(), C.c_int()
    dummy = C.c_void_p(16)  # never dereferenced: the checks come first
    assert f(dummy, dummy, 1024, 512, 512, 1.0, 1.0, 1.0, 1.7, 1e-3, 10, C.byref(res), C.byref(it)) == -2
    assert f(dummy, dummy, 0, 8, 8, 1.0, 1.0, 1.0, 1.7, 1e-3, 10, C.byref(res), C.byref(it)) == -1
    assert f(None, dummy, 8, 8, 8, 1.0, 1.0, 1.0, 1.7, 1e-3, 10, C.byref(res), C.byref(it)) == -1
