"""ctypes binding of the NuSiF C API (reference src/discretization.h:14-41,
src/solver.h:12-36, src/parameter.h:8-24, src/comm.h:27-38).

The same binding drives three libraries that export that API:
  * the GPU library   cfd-bench_amd/lib/libnusif_{rb,mg}.so  (HIP, gfx950)
  * the CPU oracle    oracle/lib/liborc_{rb,mg}.so            (tests only)
  * the reference     oracle/_ref/libnusif_ref_{rb,mg}.so     (tests only)

`Case` runs the reference's time loop (src/main.c:60-88) through whichever
library it is given, so parity tests compare like with like.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(HERE, "lib")


class Grid(C.Structure):
    _fields_ = [("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("imax", C.c_int), ("jmax", C.c_int), ("kmax", C.c_int),
                ("xlength", C.c_double), ("ylength", C.c_double), ("zlength", C.c_double)]


class CommType(C.Structure):
    _fields_ = [("rank", C.c_int), ("size", C.c_int), ("neighbours", C.c_int * 6),
                ("coords", C.c_int * 3), ("dims", C.c_int * 3),
                ("imaxLocal", C.c_int), ("jmaxLocal", C.c_int), ("kmaxLocal", C.c_int)]


class Parameter(C.Structure):
    _fields_ = [("imax", C.c_int), ("jmax", C.c_int), ("kmax", C.c_int),
                ("xlength", C.c_double), ("ylength", C.c_double), ("zlength", C.c_double),
                ("itermax", C.c_int), ("eps", C.c_double), ("omg", C.c_double),
                ("re", C.c_double), ("tau", C.c_double), ("gamma", C.c_double),
                ("te", C.c_double), ("dt", C.c_double),
                ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double),
                ("name", C.c_char_p),
                ("bcLeft", C.c_int), ("bcRight", C.c_int), ("bcBottom", C.c_int),
                ("bcTop", C.c_int), ("bcFront", C.c_int), ("bcBack", C.c_int),
                ("u_init", C.c_double), ("v_init", C.c_double), ("w_init", C.c_double),
                ("p_init", C.c_double),
                ("levels", C.c_int), ("presmooth", C.c_int), ("postsmooth", C.c_int)]


_DP = C.POINTER(C.c_double)


class Discretization(C.Structure):
    _fields_ = [("grid", Grid),
                ("p", _DP), ("rhs", _DP), ("f", _DP), ("g", _DP), ("h", _DP),
                ("u", _DP), ("v", _DP), ("w", _DP),
                ("eps", C.c_double), ("omega", C.c_double),
                ("re", C.c_double), ("tau", C.c_double), ("gamma", C.c_double),
                ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double),
                ("itermax", C.c_int), ("dt", C.c_double), ("te", C.c_double),
                ("dtBound", C.c_double), ("problem", C.c_char_p),
                ("bcLeft", C.c_int), ("bcRight", C.c_int), ("bcBottom", C.c_int),
                ("bcTop", C.c_int), ("bcFront", C.c_int), ("bcBack", C.c_int),
                ("comm", CommType)]


class Solver(C.Structure):
    _fields_ = [("grid", C.POINTER(Grid)),
                ("p", _DP), ("rhs", _DP), ("f", _DP), ("g", _DP), ("h", _DP),
                ("u", _DP), ("v", _DP), ("w", _DP),
                ("eps", C.c_double), ("omega", C.c_double),
                ("re", C.c_double), ("tau", C.c_double), ("gamma", C.c_double),
                ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double),
                ("itermax", C.c_int), ("dt", C.c_double), ("te", C.c_double),
                ("dtBound", C.c_double), ("problem", C.c_char_p),
                ("bcLeft", C.c_int), ("bcRight", C.c_int), ("bcBottom", C.c_int),
                ("bcTop", C.c_int), ("bcFront", C.c_int), ("bcBack", C.c_int),
                ("r", C.POINTER(_DP)), ("e", C.POINTER(_DP)),
                ("levels", C.c_int), ("presmooth", C.c_int), ("postsmooth", C.c_int),
                ("comm", C.POINTER(CommType))]


FIELDS = ("p", "rhs", "f", "g", "h", "u", "v", "w")
STEP_FUNCS = ("computeRHS", "normalizePressure", "computeTimestep", "setBoundaryConditions",
              "setSpecialBoundaryCondition", "computeFG", "adaptUV")

# Default parameter values of initParameter (reference src/parameter.c:13-30)
# plus the keys initParameter leaves unset (src/parameter.c, App. Q17).
DEFAULTS = dict(xlength=1.0, ylength=1.0, zlength=1.0, imax=100, jmax=100, kmax=100,
                itermax=1000, eps=0.0001, omg=1.7, re=100.0, gamma=0.9, tau=0.5,
                levels=5, presmooth=5, postsmooth=5, te=1.0, dt=0.02, gx=0.0, gy=0.0,
                gz=0.0, name="dcavity", bcLeft=1, bcRight=1, bcBottom=1, bcTop=1,
                bcFront=1, bcBack=1, u_init=0.0, v_init=0.0, w_init=0.0, p_init=0.0)


def load(path):
    """Open a library exporting the NuSiF API and declare its prototypes."""
    lib = C.CDLL(path)
    if not hasattr(lib, "initDiscretization"):  # helper library (oracle gmg / unet)
        _declare_optional(lib)
        return lib
    for fn in STEP_FUNCS + ("initDiscretization",):
        getattr(lib, fn).restype = None
    lib.initDiscretization.argtypes = [C.POINTER(Discretization), C.POINTER(Parameter)]
    for fn in STEP_FUNCS:
        getattr(lib, fn).argtypes = [C.POINTER(Discretization)]
    lib.initSolver.argtypes = [C.POINTER(Solver), C.POINTER(Discretization), C.POINTER(Parameter)]
    lib.initSolver.restype = None
    lib.solve.argtypes = [C.POINTER(Solver), _DP, _DP]
    lib.solve.restype = C.c_double
    if hasattr(lib, "commCollectResult"):
        lib.commCollectResult.argtypes = [C.POINTER(CommType)] + [_DP] * 8 + [C.c_int] * 3
        lib.commCollectResult.restype = None
    _declare_optional(lib)
    return lib


def _declare_optional(lib):
    opt = {
        # GPU library extras (include/nusif.h)
        "nusifSyncHost": ([C.POINTER(Discretization)], None),
        "nusifSyncDevice": ([C.POINTER(Discretization)], None),
        "nusifFreeDiscretization": ([C.POINTER(Discretization)], None),
        "nusifFreeSolver": ([C.POINTER(Solver)], None),
        "nusifSetOption": ([C.POINTER(Parameter), C.c_char_p, C.c_char_p], C.c_int),
        "nusifVerboseLevel": ([C.POINTER(Parameter)], C.c_int),
        "nusifSolverIterations": ([C.POINTER(Solver)], C.c_int),
        "nusifSolverFirstGroup": ([C.POINTER(Solver)], C.c_int),
        "nusifSolverRelResidual": ([C.POINTER(Solver)], C.c_double),
        "nusifDeviceName": ([C.c_char_p, C.c_int], C.c_int),
        "nusifTimerEnable": ([C.c_char_p], C.c_int),
        "nusifTimerDisableAll": ([], None),
        "nusifTimeFinestSweep": ([C.POINTER(Solver), C.c_int], C.c_double),
        "nusifTimerRead": ([C.c_char_p, C.POINTER(C.c_int)], C.c_double),
        "nusifTimerReset": ([], None),
        "nusifKernelBytes": ([C.POINTER(Discretization), C.c_char_p], C.c_double),
        "nusifDeviceSynchronize": ([], None),
        "nusifCopyProbe": ([C.c_size_t, C.c_int], C.c_double),
        "nusifStreamProbe": ([C.c_size_t, C.c_int, _DP], None),
        "nusifWarmstartUsed": ([C.POINTER(Solver)], C.c_int),
        "nusifWarmstartResiduals": ([C.POINTER(Solver), _DP], None),
        "nusifWarmstartCounts": ([C.POINTER(Solver), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)],
                                 None),
        "nusifUnetTrainCavity": ([C.c_char_p] + [C.c_int] * 6 + [C.c_double, C.c_int, _DP], C.c_int),
        "nusifUnetTrainCavityGrids": ([C.c_char_p, C.POINTER(C.c_int)] + [C.c_int] * 7 +
                                      [C.c_double, C.c_int, _DP], C.c_int),
        "nusifUnetTrainChannel": ([C.c_char_p, C.c_char_p] + [C.c_int] * 3 + [C.c_double] + [C.c_int] * 6 +
                                  [C.c_double, C.c_int, _DP], C.c_int),
        "nusifUnetLossGrad": ([C.POINTER(C.c_int)] + [C.c_int] * 3 + [_DP, _DP, _DP, _DP, _DP], C.c_double),
        "nusifUnetPredict": ([C.POINTER(C.c_int)] + [C.c_int] * 3 + [_DP, _DP, _DP], None),
        "nusifRestrictResidualCheck": ([C.c_int] * 3 + [C.c_double] * 3 + [_DP, _DP, C.POINTER(C.c_float),
                                       C.POINTER(C.c_float), _DP], None),
        "nusifProlongSetCheck": ([C.c_int] * 3 + [C.POINTER(C.c_float), C.POINTER(C.c_float)], None),
        "nusifRunSteps": ([C.POINTER(Discretization), C.POINTER(Solver), C.c_int,
                           C.POINTER(C.c_double)], C.c_int),
        "nusifSolveTiming": ([C.POINTER(Discretization), C.c_int], None),
        "nusifSolveRBNoAlloc": ([C.c_void_p, C.c_void_p] + [C.c_int] * 3 + [C.c_double] * 5 + [C.c_int, _DP,
                                C.POINTER(C.c_int)], C.c_int),
        "nusifSolveTimes": ([C.POINTER(Discretization), _DP, C.POINTER(C.c_longlong)], None),
        # oracle extras
        "orcSetProblemExtras": ([C.c_double, C.c_double], None),
        "orcLastIterations": ([], C.c_int),
        "orcFreeDiscretization": ([C.POINTER(Discretization)], None),
        "orcGmgLevels": ([C.c_int] * 4, C.c_int),
        "orcGmgSolve": ([C.c_int] * 3 + [C.c_double] * 3 + [_DP, _DP] + [C.c_int] * 3
                        + [C.c_double] * 2 + [C.c_int] * 3 + [_DP], C.c_int),
        "orcGmgSolveGrouped": ([C.c_int] * 3 + [C.c_double] * 3 + [_DP, _DP] + [C.c_int] * 3
                               + [C.c_double] * 2 + [C.c_int] * 4 + [_DP], C.c_int),
        "orcGmgCoarseSolve": ([C.c_int] * 3 + [C.c_double] * 3 + [_DP, _DP], None),
        "orcUnetParamCount": ([C.POINTER(C.c_int)], C.c_long),
        "orcUnetLoss": ([C.POINTER(C.c_int), C.c_int, _DP, _DP, _DP, _DP, _DP, C.c_int], C.c_double),
        "orcUnetLossBox": ([C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int, _DP, _DP, _DP, _DP, _DP, C.c_int],
                           C.c_double),
        "orcUnetTransfer": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _DP, _DP], None),
        "orcBf16": ([C.c_double], C.c_double),
        "orcGmgProlong": ([C.c_int] * 3 + [_DP, _DP], None),
        "orcGmgRestrict": ([C.c_int] * 3 + [C.c_double] * 3 + [_DP] * 3, None),
    }
    for name, (args, res) in opt.items():
        if hasattr(lib, name):
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res


def make_parameter(**kw):
    """A Parameter filled from DEFAULTS overridden by keyword values."""
    vals = dict(DEFAULTS)
    vals.update(kw)
    prm = Parameter()
    keep = []
    for name, _ in Parameter._fields_:
        v = vals[name]
        if name == "name":
            b = v.encode() if isinstance(v, str) else v
            keep.append(b)
            setattr(prm, name, b)
        else:
            setattr(prm, name, v)
    prm._keep = keep
    return prm


class Case:
    """One problem instance on one library: Parameter, Discretization, Solver."""

    def __init__(self, lib, options=None, **params):
        self.lib = lib
        self.prm = make_parameter(**params)
        self.options = dict(options or {})
        if hasattr(lib, "nusifSetOption"):
            # the library prints the reference's -DVERBOSE lines by default;
            # a caller driving it from Python reads the results instead
            self.options.setdefault("verbose", 0)
        if self.options:
            if not hasattr(lib, "nusifSetOption"):
                raise RuntimeError("library has no nusifSetOption")
            for k, v in self.options.items():
                rc = lib.nusifSetOption(C.byref(self.prm), k.encode(), str(v).encode())
                if rc != 0:
                    raise ValueError("option %s=%s rejected" % (k, v))
        self.d = Discretization()
        self.s = Solver()
        c = self.d.comm
        c.rank, c.size = 0, 1
        c.imaxLocal, c.jmaxLocal, c.kmaxLocal = self.prm.imax, self.prm.jmax, self.prm.kmax
        lib.initDiscretization(C.byref(self.d), C.byref(self.prm))
        lib.initSolver(C.byref(self.s), C.byref(self.d), C.byref(self.prm))
        self.t = 0.0
        self.nt = 0
        self.trace = []  # (t, dt, res) per step, as writeResidual records t before t += dt

    # ---- host views of the padded fields -------------------------------
    @property
    def shape(self):
        g = self.d.grid
        return (g.kmax + 2, g.jmax + 2, g.imax + 2)

    def view(self, name):
        n = int(np.prod(self.shape))
        ptr = getattr(self.d, name)
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self.shape)

    def pull(self):
        """Make host arrays current (device -> host) when the library has devices."""
        if hasattr(self.lib, "nusifSyncHost"):
            self.lib.nusifSyncHost(C.byref(self.d))

    def push(self):
        if hasattr(self.lib, "nusifSyncDevice"):
            self.lib.nusifSyncDevice(C.byref(self.d))

    def fields(self):
        self.pull()
        return {k: self.view(k).copy() for k in FIELDS}

    def set_fields(self, **arrays):
        self.pull()
        for k, a in arrays.items():
            self.view(k)[...] = a
        self.push()

    def call(self, fn):
        getattr(self.lib, fn)(C.byref(self.d))

    # ---- the reference time step (src/main.c:60-88) --------------------
    def step(self):
        d = self.d
        if d.tau > 0.0:
            self.lib.computeTimestep(C.byref(d))
        self.lib.setBoundaryConditions(C.byref(d))
        self.lib.setSpecialBoundaryCondition(C.byref(d))
        self.lib.computeFG(C.byref(d))
        self.lib.computeRHS(C.byref(d))
        if self.nt % 100 == 0:
            self.lib.normalizePressure(C.byref(d))
        res = self.lib.solve(C.byref(self.s), d.p, d.rhs)
        self.lib.adaptUV(C.byref(d))
        self.trace.append((self.t, d.dt, res))
        self.t += d.dt
        self.nt += 1
        return res

    def run(self, nsteps):
        for _ in range(nsteps):
            self.step()
        return self

    def iterations(self):
        if hasattr(self.lib, "nusifSolverIterations"):
            return self.lib.nusifSolverIterations(C.byref(self.s))
        if hasattr(self.lib, "orcLastIterations"):
            return self.lib.orcLastIterations()
        return None

    def close(self):
        if hasattr(self.lib, "nusifFreeSolver"):
            self.lib.nusifFreeSolver(C.byref(self.s))
        if hasattr(self.lib, "nusifFreeDiscretization"):
            self.lib.nusifFreeDiscretization(C.byref(self.d))
        elif hasattr(self.lib, "orcFreeDiscretization"):
            self.lib.orcFreeDiscretization(C.byref(self.d))


def gpu_lib_path(solver="rb"):
    return os.path.join(LIBDIR, "libnusif_%s.so" % solver)


def oracle_lib_path(kind="rb"):
    return os.path.join(ROOT, "oracle", "lib", "liborc_%s.so" % kind)


def ref_lib_path(kind="rb"):
    return os.path.join(ROOT, "oracle", "_ref", "libnusif_ref_%s.so" % kind)


def rel_maxnorm(a, b):
    """max|a-b| / max|b| (0 when both are zero)."""
    den = float(np.max(np.abs(b)))
    num = float(np.max(np.abs(a - b)))
    if den == 0.0:
        return num
    return num / den
