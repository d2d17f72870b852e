"""ctypes driver of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker. ``Oracle(scene, precision=64)`` holds an FP64 copy of a
scene and runs the restated reference stages (mpm_oracle.hpp); precision=32 with threads is
the CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "liboracle.so")
HEADER = os.path.join(ROOT, "include", "snow.h")


class _Material(C.Structure):
    _fields_ = [("young", C.c_float), ("poisson", C.c_float), ("hardening", C.c_float),
                ("crit_comp", C.c_float), ("crit_stretch", C.c_float), ("flip_alpha", C.c_float),
                ("gravity", C.c_float * 3), ("grad_scale", C.c_float)]


class _Grid(C.Structure):
    _fields_ = [("dim", C.c_int32 * 3), ("origin", C.c_float * 3), ("spacing", C.c_float)]


def _scene_struct(T):
    P = C.POINTER(T)

    class S(C.Structure):
        _fields_ = [("n", C.c_int64), ("x", P), ("v", P), ("m", P), ("fe", P), ("fp", P), ("rho0", P),
                    ("grid", _Grid), ("material", _Material), ("num_colliders", C.c_int32), ("colliders", P),
                    ("gm", P), ("gmv", P), ("gf", P), ("gvstar", P), ("nthreads", C.c_int32)]
    return S


Scene64 = _scene_struct(C.c_double)
Scene32 = _scene_struct(C.c_float)
_lib = None


def build() -> str:
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    for p, T, S in (("oracle64", C.c_double, Scene64), ("oracle32", C.c_float, Scene32)):
        SP = C.POINTER(S)
        TP = C.POINTER(T)
        for fn in ("init", "grid_mass", "particle_volume", "p2g"):
            getattr(lib, f"{p}_{fn}").argtypes = [SP]
            getattr(lib, f"{p}_{fn}").restype = C.c_int
        for fn in ("rigid_advance", "grid_update", "g2p_advance"):
            getattr(lib, f"{p}_{fn}").argtypes = [SP, C.c_double]
            getattr(lib, f"{p}_{fn}").restype = C.c_int
        getattr(lib, f"{p}_step").argtypes = [SP, C.c_double, C.c_int32]
        getattr(lib, f"{p}_step").restype = C.c_int
        for fn in ("weighting", "weighting_gradient"):
            getattr(lib, f"{p}_{fn}").argtypes = [T]
            getattr(lib, f"{p}_{fn}").restype = T
        for fn in ("mu", "lambda"):
            getattr(lib, f"{p}_{fn}").argtypes = [T, T, T, T]
            getattr(lib, f"{p}_{fn}").restype = T
        getattr(lib, f"{p}_svd").argtypes = [TP, TP, TP, TP]
        getattr(lib, f"{p}_svd").restype = None
        getattr(lib, f"{p}_polar").argtypes = [TP, TP]
        getattr(lib, f"{p}_polar").restype = None
        getattr(lib, f"{p}_getijk").argtypes = [C.c_int32, C.POINTER(C.c_int32)]
        getattr(lib, f"{p}_getijk").restype = None
        getattr(lib, f"{p}_collide").argtypes = [TP, TP, TP, C.c_int32, TP]
        getattr(lib, f"{p}_collide").restype = None
    DP = C.POINTER(C.c_double)
    SP = C.POINTER(Scene64)
    lib.oracle64_fc_stress.argtypes = [DP, C.c_double, C.c_double, DP]
    lib.oracle64_fc_hessian.argtypes = [DP, C.c_double, C.c_double, DP, DP]
    lib.oracle64_polar_exact.argtypes = [DP, DP]
    for fn in ("fc_stress", "fc_hessian", "polar_exact"):
        getattr(lib, f"oracle64_{fn}").restype = None
    for fn in ("implicit_df", "implicit_force"):
        getattr(lib, f"oracle64_{fn}").argtypes = [SP, DP, DP]
        getattr(lib, f"oracle64_{fn}").restype = C.c_int
    lib.oracle64_implicit_apply.argtypes = [SP, C.c_double, C.c_double, DP, DP]
    lib.oracle64_implicit_apply.restype = C.c_int
    lib.oracle64_implicit_solve.argtypes = [SP, C.c_double, C.c_double, C.c_double, C.c_int32,
                                            C.POINTER(C.c_int32), DP]
    lib.oracle64_implicit_solve.restype = C.c_int
    lib.oracle64_step_semi_implicit.argtypes = [SP, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                                C.POINTER(C.c_int32), DP]
    lib.oracle64_step_semi_implicit.restype = C.c_int
    _lib = lib
    return lib


def _p(a, T):
    return None if a is None else a.ctypes.data_as(C.POINTER(T))


class Oracle:
    """FP64 (precision=64) oracle or FP32 (precision=32) CPU baseline for one PhysicalScene."""

    def __init__(self, scene, precision: int = 64, nthreads: int = 1):
        self.lib = load()
        self.dtype = np.float64 if precision == 64 else np.float32
        self.T = C.c_double if precision == 64 else C.c_float
        self.pre = "oracle64" if precision == 64 else "oracle32"
        ps = scene.particles
        dt = self.dtype
        self.x = np.array(ps.x, dt)
        self.v = np.array(ps.v, dt)
        self.m = np.array(ps.mass, dt)
        self.fe = np.array(ps.fe, dt)
        self.fp = np.array(ps.fp, dt)
        self.rho0 = np.zeros(len(ps), dt)
        g = scene.grid
        self.dims = (g.dimx, g.dimy, g.dimz)
        nn = g.num_nodes
        self.gm = np.zeros(nn, dt)
        self.gmv = np.zeros((nn, 3), dt)
        self.gf = np.zeros((nn, 3), dt)
        self.gvstar = np.zeros((nn, 3), dt)
        cols = [c.as_array() for c in scene.colliders.colliders]
        self.colliders = np.array(cols, dt).reshape(-1, 12)
        S = Scene64 if precision == 64 else Scene32
        s = S()
        s.n = len(ps)
        T = self.T
        s.x, s.v, s.m, s.fe, s.fp, s.rho0 = (_p(a, T) for a in (self.x, self.v, self.m, self.fe, self.fp, self.rho0))
        s.grid.dim[0], s.grid.dim[1], s.grid.dim[2] = g.dimx, g.dimy, g.dimz
        s.grid.origin[0], s.grid.origin[1], s.grid.origin[2] = g.x_off, g.y_off, g.z_off
        s.grid.spacing = g.h
        mt = scene.material
        s.material.young, s.material.poisson, s.material.hardening = mt.young, mt.poisson, mt.hardening
        s.material.crit_comp, s.material.crit_stretch, s.material.flip_alpha = mt.crit_comp, mt.crit_stretch, mt.flip_alpha
        for k in range(3):
            s.material.gravity[k] = mt.gravity[k]
        s.material.grad_scale = mt.grad_scale
        s.num_colliders = len(self.colliders)
        s.colliders = _p(self.colliders, T) if len(self.colliders) else None
        s.gm, s.gmv, s.gf, s.gvstar = (_p(a, T) for a in (self.gm, self.gmv, self.gf, self.gvstar))
        s.nthreads = nthreads
        self.s = s

    def _call(self, name, *args):
        rc = getattr(self.lib, f"{self.pre}_{name}")(C.byref(self.s), *args)
        if rc:
            raise RuntimeError(f"oracle {name} failed ({rc})")

    def init(self):
        self._call("init")

    def grid_mass(self):
        self._call("grid_mass")

    def particle_volume(self):
        self._call("particle_volume")

    def rigid_advance(self, dt):
        self._call("rigid_advance", float(np.float32(dt)))

    def p2g(self):
        self._call("p2g")

    def grid_update(self, dt):
        self._call("grid_update", float(np.float32(dt)))

    def g2p_advance(self, dt):
        self._call("g2p_advance", float(np.float32(dt)))

    def step(self, dt, nsteps: int = 1):
        self._call("step", float(np.float32(dt)), int(nsteps))

    # ---- semi-implicit update (FP64 oracle only) -------------------------------------------
    def _need64(self):
        if self.pre != "oracle64":
            raise RuntimeError("the semi-implicit update is in the FP64 oracle only")

    def implicit_df(self, du):
        """df(du) per node (the Hessian action of the node force, no dt), from the current particles."""
        self._need64()
        du = np.ascontiguousarray(du, np.float64).reshape(-1, 3)
        out = np.empty_like(du)
        self._call("implicit_df", _p(du, C.c_double), _p(out, C.c_double))
        return out

    def implicit_force(self, du):
        """The node force of the energy with the nodes displaced by du (exact polar factor)."""
        self._need64()
        du = np.ascontiguousarray(du, np.float64).reshape(-1, 3)
        out = np.empty_like(du)
        self._call("implicit_force", _p(du, C.c_double), _p(out, C.c_double))
        return out

    def implicit_apply(self, dt, beta, v):
        """K v = M v - beta dt df(dt v) on the nodes with m > 0 (uses the current grid mass)."""
        self._need64()
        v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        out = np.empty_like(v)
        self._call("implicit_apply", float(dt), float(beta), _p(v, C.c_double), _p(out, C.c_double))
        return out

    def implicit_solve(self, dt, beta, tol=1e-6, max_iter=50):
        """Replaces gvstar by the solution of K v = M v*; returns (iterations, relative residual)."""
        self._need64()
        it, rr = C.c_int32(0), C.c_double(0.0)
        self._call("implicit_solve", float(np.float32(dt)), float(beta), float(tol), int(max_iter), C.byref(it),
                   C.byref(rr))
        return it.value, rr.value

    def step_semi_implicit(self, dt, beta, tol=1e-6, max_iter=50, nsteps: int = 1):
        self._need64()
        it, rr = C.c_int32(0), C.c_double(0.0)
        self._call("step_semi_implicit", float(np.float32(dt)), float(beta), float(tol), int(max_iter),
                   int(nsteps), C.byref(it), C.byref(rr))
        return it.value, rr.value


def svd(A, precision=64):
    lib = load()
    dt, T, pre = (np.float64, C.c_double, "oracle64") if precision == 64 else (np.float32, C.c_float, "oracle32")
    A = np.ascontiguousarray(np.asarray(A, dt).T).reshape(9)  # column-major for the oracle
    W, S, V = np.empty(9, dt), np.empty(9, dt), np.empty(9, dt)
    getattr(lib, f"{pre}_svd")(_p(A, T), _p(W, T), _p(S, T), _p(V, T))
    # stored column-major: element (r, c) at c*3 + r -> transpose of a row-major reshape
    return W.reshape(3, 3).T, S.reshape(3, 3).T, V.reshape(3, 3).T


def polar(A, precision=64):
    lib = load()
    dt, T, pre = (np.float64, C.c_double, "oracle64") if precision == 64 else (np.float32, C.c_float, "oracle32")
    Acm = np.ascontiguousarray(np.asarray(A, dt).T).reshape(9)
    R = np.empty(9, dt)
    getattr(lib, f"{pre}_polar")(_p(Acm, T), _p(R, T))
    return R.reshape(3, 3).T


def weighting(x):
    return load().oracle64_weighting(float(x))


def weighting_gradient(x):
    return load().oracle64_weighting_gradient(float(x))


def getijk(s):
    out = (C.c_int32 * 3)()
    load().oracle64_getijk(s, out)
    return tuple(out)


def collide(x, v, colliders):
    lib = load()
    x = np.ascontiguousarray(x, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    cs = np.ascontiguousarray(colliders, np.float64).reshape(-1, 12)
    out = np.empty(3, np.float64)
    lib.oracle64_collide(_p(x, C.c_double), _p(v, C.c_double), _p(cs, C.c_double), len(cs), _p(out, C.c_double))
    return out


def _cm(A):
    return np.ascontiguousarray(np.asarray(A, np.float64).T).reshape(9)


def fc_stress(F, mu, la):
    """First Piola-Kirchhoff stress 2 mu (F - R) + lambda (J - 1) cof F, exact polar factor."""
    out = np.empty(9)
    load().oracle64_fc_stress(_p(_cm(F), C.c_double), float(mu), float(la), _p(out, C.c_double))
    return out.reshape(3, 3).T


def fc_hessian(F, mu, la, dF):
    """d2Psi/dF2 (F) : dF of the same energy."""
    out = np.empty(9)
    load().oracle64_fc_hessian(_p(_cm(F), C.c_double), float(mu), float(la), _p(_cm(dF), C.c_double),
                               _p(out, C.c_double))
    return out.reshape(3, 3).T


def polar_exact(F):
    out = np.empty(9)
    load().oracle64_polar_exact(_p(_cm(F), C.c_double), _p(out, C.c_double))
    return out.reshape(3, 3).T
