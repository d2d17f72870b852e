"""snow_amd - MI355X-native explicit snow MPM timestep.

Host-side mirror of the reference's simulation interface over the C ABI of
``include/snow.h`` (the HIP library ``libsnow_amd.so``):

* ``ParticleSystem``, ``Grid``, ``Collider``/``CollisionObjects``, ``PhysicalScene`` follow
  src/snow/object/{particlesystem,grid,collider,collisionObjects}.hpp and src/snow/scene/Scene.hpp;
* ``ExplicitTimeUpdate`` (``init()``, ``update(dt)``) follows
  src/snow/simulation/timeUpdate/explicittimeupdate.hpp:11-30 and timeUpdate.hpp:10-18;
* ``MPMPhysicEngine`` follows src/snow/simulation/MPMPhysicEngine.hpp:7-16.

All compute runs in the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._capi import (FLAG_COMPENSATED_G2P, FLAG_KEEP_GRID, FLAG_MIXED, Collider as _CCollider, Grid as _CGrid,
                    Material as _CMaterial, SceneDesc as _CSceneDesc, SnowError, load_library, ptr)

__all__ = [
    "Material", "Grid", "Collider", "CollisionObjects", "ParticleSystem", "PhysicalScene", "ExplicitTimeUpdate",
    "MPMPhysicEngine", "SnowContext", "SnowError", "FLAG_COMPENSATED_G2P", "FLAG_MIXED", "FLAG_KEEP_GRID",
    "HALF_PLANE", "SPHERE", "BOX", "default_boundary_colliders", "read_ply",
]

HALF_PLANE, SPHERE, BOX = 0, 1, 2


@dataclass
class Material:
    """Material constants; defaults are src/snow/utils/defines.hpp:50-54 and the shader literals."""
    young: float = 6e6
    poisson: float = 0.2
    hardening: float = 30.0
    crit_comp: float = 1.0e-2
    crit_stretch: float = 1e-3
    flip_alpha: float = 0.95
    gravity: tuple = (0.0, -9.81, 0.0)
    grad_scale: float = 1.0  # 1 = reference (no 1/h in cubic.include.glsl:25-40); 0 = physical 1/h

    def to_c(self) -> _CMaterial:
        m = _CMaterial()
        m.young, m.poisson, m.hardening = self.young, self.poisson, self.hardening
        m.crit_comp, m.crit_stretch, m.flip_alpha = self.crit_comp, self.crit_stretch, self.flip_alpha
        for k in range(3):
            m.gravity[k] = self.gravity[k]
        m.grad_scale = self.grad_scale
        return m


@dataclass
class Grid:
    """Dense node grid, Grid(x, y, z, spacing, pos_x, pos_y, pos_z) as grid.hpp:20-28."""
    dimx: int = 201
    dimy: int = 201
    dimz: int = 201
    h: float = 0.05
    x_off: float = 0.5125
    y_off: float = 0.5125
    z_off: float = 1.5125

    @property
    def dims(self):
        return (self.dimx, self.dimy, self.dimz)

    @property
    def num_nodes(self) -> int:
        return self.dimx * self.dimy * self.dimz

    def to_c(self) -> _CGrid:
        g = _CGrid()
        g.dim[0], g.dim[1], g.dim[2] = self.dimx, self.dimy, self.dimz
        g.origin[0], g.origin[1], g.origin[2] = self.x_off, self.y_off, self.z_off
        g.spacing = self.h
        return g

    def node_positions(self) -> np.ndarray:
        """Node positions x0 + h*(i,j,k), x fastest (grid.cpp:98-108)."""
        k, j, i = np.meshgrid(np.arange(self.dimz), np.arange(self.dimy), np.arange(self.dimx), indexing="ij")
        pos = np.stack([self.x_off + self.h * i, self.y_off + self.h * j, self.z_off + self.h * k], -1)
        return pos.reshape(-1, 3).astype(np.float32)


@dataclass
class Collider:
    """Collision object in the cx/cv/cn encoding of collisionObjects.cpp:25-66."""
    position: tuple
    type: int = HALF_PLANE
    velocity: tuple = (0.0, 0.0, 0.0)
    radius: float = 0.0
    normal: tuple = (0.0, 0.0, 0.0)  # box: half-extents
    friction: float = 0.2

    def to_c(self) -> _CCollider:
        c = _CCollider()
        for k in range(3):
            c.pos[k], c.vel[k], c.normal[k] = self.position[k], self.velocity[k], self.normal[k]
        c.type, c.radius, c.friction = float(self.type), self.radius, self.friction
        return c

    def as_array(self) -> np.ndarray:
        return np.array([*self.position, float(self.type), *self.velocity, self.radius, *self.normal, self.friction],
                        dtype=np.float64)


@dataclass
class CollisionObjects:
    colliders: list = field(default_factory=list)


def default_boundary_colliders(grid: Grid, offset_cells: int = 4, friction: float = 0.2) -> list:
    """The six inward half-planes launchSnow adds at GRID_COLLISION_PLANE_OFFSET cells from the grid faces
    (explicit.cpp:9-67)."""
    lo = (grid.x_off + offset_cells * grid.h, grid.y_off + offset_cells * grid.h, grid.z_off + offset_cells * grid.h)
    hi = (grid.x_off + (grid.dimx - offset_cells) * grid.h, grid.y_off + (grid.dimy - offset_cells) * grid.h,
          grid.z_off + (grid.dimz - offset_cells) * grid.h)
    out = []
    for n in ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)):
        out.append(Collider(lo, HALF_PLANE, normal=n, friction=friction))
    for n in ((-1.0, 0.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, -1.0)):
        out.append(Collider(hi, HALF_PLANE, normal=n, friction=friction))
    return out


class ParticleSystem:
    """Particle set (particlesystem.hpp:15-59). Arrays are float32: x [N,3], v [N,3] (m/s), mass [N],
    fe/fp [N,9] column-major (identity by default, particle.hpp:8-11)."""

    def __init__(self, x=None, v=None, mass=None, fe=None, fp=None):
        self.x = np.zeros((0, 3), np.float32) if x is None else np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        n = len(self.x)
        self.v = np.zeros((n, 3), np.float32) if v is None else np.ascontiguousarray(v, np.float32).reshape(-1, 3)
        self.mass = (np.full(n, 1.25e-3, np.float32) if mass is None
                     else np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float32), (n,))))
        eye = np.eye(3, dtype=np.float32).reshape(-1)
        self.fe = np.tile(eye, (n, 1)) if fe is None else np.ascontiguousarray(fe, np.float32).reshape(-1, 9)
        self.fp = np.tile(eye, (n, 1)) if fp is None else np.ascontiguousarray(fp, np.float32).reshape(-1, 9)

    def __len__(self):
        return len(self.x)

    @classmethod
    def from_file(cls, filename: str, pos=(0.0, 0.0, 0.0), scale=(1.0, 1.0, 1.0), mass: float = 6.25e-3):
        """ParticleSystem::initParticlesFromFile (particlesystem.cpp:14-38): CSV lines "x,y,z", each
        scaled and offset, velocity 0, mass 6.25e-3."""
        pts = []
        with open(filename) as f:
            for line in f:
                parts = line.strip().split(",")
                if len(parts) < 3 or not parts[0]:
                    continue
                pts.append([float(parts[0]), float(parts[1]), float(parts[2])])
        x = np.asarray(pts, np.float32).reshape(-1, 3) * np.asarray(scale, np.float32) + np.asarray(pos, np.float32)
        return cls(x=x, mass=np.full(len(x), mass, np.float32))

    @classmethod
    def from_reference_buffers(cls, pxm: np.ndarray, pv: np.ndarray, pfe=None, pfp=None):
        """Decode the reference SSBO encodings (pxm vec4, pv ivec4 x1e8, mat4 column-major)."""
        lib = load_library()
        n = len(pxm)
        pxm = np.ascontiguousarray(pxm, np.float32).reshape(n, 4)
        pv = np.ascontiguousarray(pv, np.int32).reshape(n, 4)
        pfe = None if pfe is None else np.ascontiguousarray(pfe, np.float32).reshape(n, 16)
        pfp = None if pfp is None else np.ascontiguousarray(pfp, np.float32).reshape(n, 16)
        x, v = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32)
        m, fe, fp = np.empty(n, np.float32), np.empty((n, 9), np.float32), np.empty((n, 9), np.float32)
        rc = lib.snow_import_reference_particles(n, ptr(pxm), ptr(pv, C.c_int32), ptr(pfe), ptr(pfp), ptr(x), ptr(v),
                                                 ptr(m), ptr(fe), ptr(fp))
        if rc:
            raise SnowError(rc, "import of reference buffers failed")
        return cls(x, v, m, fe, fp)


@dataclass
class PhysicalScene:
    """PhysicalScene{particleSys, grid, colliderSys} (Scene.hpp:7-11) plus the material and step size."""
    particles: ParticleSystem
    grid: Grid
    colliders: CollisionObjects = field(default_factory=CollisionObjects)
    material: Material = field(default_factory=Material)
    dt: float = 1e-3

    def save(self, filename: str) -> None:
        """Write the scene in the SNOWSCN1 format read by the C++ driver (snow_amd/csrc/snow_run.cpp):
        the grid, material, colliders, dt and the particles' x, v, mass, F_E, F_P."""
        g, m = self.grid.to_c(), self.material.to_c()
        cols = [c.to_c() for c in self.colliders.colliders]
        ps = self.particles
        with open(filename, "wb") as f:
            f.write(b"SNOWSCN1")
            f.write(np.int64(len(ps)).tobytes())
            f.write(bytes(g))  # dim[3], origin[3], spacing
            f.write(bytes(m))
            f.write(np.int32(len(cols)).tobytes())
            for c in cols:
                f.write(bytes(c))
            f.write(np.float32(self.dt).tobytes())
            for a in (ps.x, ps.v, ps.mass, ps.fe, ps.fp):
                f.write(np.ascontiguousarray(a, np.float32).tobytes())

    @classmethod
    def load(cls, filename: str) -> "PhysicalScene":
        """Read a SNOWSCN1 scene file (the format of save() and of the C++ driver's --dump-scene)."""
        with open(filename, "rb") as f:
            data = f.read()
        if data[:8] != b"SNOWSCN1":
            raise ValueError(f"{filename}: not a SNOWSCN1 scene file")
        off = 8

        def take(dtype, count):
            nonlocal off
            a = np.frombuffer(data, dtype, count, off)
            off += a.nbytes
            return a

        n = int(take(np.int64, 1)[0])
        dim = take(np.int32, 3)
        origin, h = take(np.float32, 3), float(take(np.float32, 1)[0])
        mt = take(np.float32, 10)
        ncol = int(take(np.int32, 1)[0])
        cols = take(np.float32, 12 * ncol).reshape(ncol, 12)
        dt = float(take(np.float32, 1)[0])
        x, v, mass = take(np.float32, 3 * n), take(np.float32, 3 * n), take(np.float32, n)
        fe, fp = take(np.float32, 9 * n), take(np.float32, 9 * n)
        grid = Grid(int(dim[0]), int(dim[1]), int(dim[2]), h, float(origin[0]), float(origin[1]), float(origin[2]))
        material = Material(float(mt[0]), float(mt[1]), float(mt[2]), float(mt[3]), float(mt[4]), float(mt[5]),
                            (float(mt[6]), float(mt[7]), float(mt[8])), float(mt[9]))
        colliders = [Collider(tuple(map(float, c[0:3])), int(c[3]), tuple(map(float, c[4:7])), float(c[7]),
                              tuple(map(float, c[8:11])), float(c[11])) for c in cols]
        ps = ParticleSystem(x.reshape(n, 3), v.reshape(n, 3), mass, fe.reshape(n, 9), fp.reshape(n, 9))
        return cls(ps, grid, CollisionObjects(colliders), material, dt)


def read_ply(filename: str) -> dict:
    """Read a binary little-endian PLY point set of float properties (the frames written by
    SnowContext.export_ply and the C++ driver's --ply): {property name: array [N]}."""
    with open(filename, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    header = data[:end].decode("ascii").split("\n")
    if header[0] != "ply" or header[1] != "format binary_little_endian 1.0":
        raise ValueError(f"{filename}: not a binary little-endian PLY file")
    n, props = 0, []
    for line in header[2:]:
        w = line.split()
        if w[:2] == ["element", "vertex"]:
            n = int(w[2])
        elif w[:2] == ["property", "float"]:
            props.append(w[2])
    a = np.frombuffer(data, np.float32, n * len(props), end).reshape(n, len(props))
    return {p: a[:, i].copy() for i, p in enumerate(props)}


def load_run_output(filename: str) -> dict:
    """Read the SNOWOUT1 file written by the C++ driver: x, v [N,3], fe, fp [N,9]."""
    with open(filename, "rb") as f:
        if f.read(8) != b"SNOWOUT1":
            raise ValueError(f"{filename}: not a SNOWOUT1 file")
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        out = {}
        for k, w in (("x", 3), ("v", 3), ("fe", 9), ("fp", 9)):
            out[k] = np.frombuffer(f.read(4 * w * n), np.float32).reshape(n, w).copy()
    return out


class SnowContext:
    """Owns one ``snow_ctx`` (one HIP stream) holding a batch of scenes."""

    def __init__(self, scenes, flags: int = 0):
        if isinstance(scenes, PhysicalScene):
            scenes = [scenes]
        self.lib = load_library()
        self.scenes = list(scenes)
        self._keep = []
        descs = (_CSceneDesc * len(self.scenes))()
        for d, s in zip(descs, self.scenes):
            ps = s.particles
            arrs = [ps.x, ps.v, ps.mass, ps.fe, ps.fp]
            self._keep.extend(arrs)
            d.num_particles = len(ps)
            d.x, d.v, d.mass, d.fe, d.fp = (ptr(a) for a in arrs)
            d.grid = s.grid.to_c()
            d.material = s.material.to_c()
            cols = (_CCollider * max(1, len(s.colliders.colliders)))()
            for k, c in enumerate(s.colliders.colliders):
                cols[k] = c.to_c()
            self._keep.append(cols)
            d.num_colliders = len(s.colliders.colliders)
            d.colliders = C.cast(cols, C.POINTER(_CCollider))
            d.dt = s.dt
        h = C.c_void_p()
        rc = self.lib.snow_create_batch(descs, len(self.scenes), flags, C.byref(h))
        self._keep = []
        if rc:
            raise SnowError(rc, self.lib.snow_last_error(None).decode())
        self.h = h
        self.flags = flags

    # -- plumbing
    def _chk(self, rc):
        if rc:
            raise SnowError(rc, self.lib.snow_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.snow_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- step interface
    def init(self):
        self._chk(self.lib.snow_init(self.h))

    def step(self, dt: float | None = None):
        self._chk(self.lib.snow_step_batch(self.h) if dt is None else self.lib.snow_step(self.h, dt))

    def run_graph(self, steps: int):
        self._chk(self.lib.snow_run_graph(self.h, steps))

    def step_semi_implicit(self, dt: float | None = None, beta: float = 1.0, tol: float = 1e-5,
                           max_iter: int = 100):
        """One semi-implicit step (snow_step_semi_implicit): dt None takes each scene's own dt."""
        self._chk(self.lib.snow_step_semi_implicit(self.h, 0.0 if dt is None else dt, beta, tol, max_iter))

    def implicit_stats(self, scene: int = 0) -> tuple:
        """(iterations, relative residual) of the scene's last semi-implicit solve."""
        it, rr = C.c_int32(), C.c_double()
        self._chk(self.lib.snow_implicit_stats(self.h, scene, C.byref(it), C.byref(rr)))
        return int(it.value), float(rr.value)

    def autotune(self, steps: int = 5) -> dict:
        g, p, q, b = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(self.lib.snow_autotune(self.h, steps, C.byref(g), C.byref(p), C.byref(q), C.byref(b)))
        return {"g2p_block": g.value, "p2g_chunk": p.value, "grid_blocks": q.value, "sort_bits": b.value}

    def sync(self):
        self._chk(self.lib.snow_sync(self.h))

    def check_finite(self):
        self._chk(self.lib.snow_check_finite(self.h))

    # -- stages
    def stage(self, name: str, dt: float | None = None):
        fn = getattr(self.lib, "snow_stage_" + name)
        self._chk(fn(self.h) if dt is None else fn(self.h, dt))

    # -- state transfer
    def export_ply(self, filename: str, scene: int = 0) -> None:
        """Write one frame of a scene as a binary little-endian PLY point set (x, y, z, mass): the
        vec4 (x, m) buffer the reference renders as GL points (particlesystem.cpp:4-13), in creation
        order."""
        pxm, _, _, _ = self.export_reference(scene)
        pxm = np.ascontiguousarray(pxm, "<f4").reshape(-1, 4)
        with open(filename, "wb") as f:
            f.write(("ply\nformat binary_little_endian 1.0\nelement vertex %d\nproperty float x\nproperty float y\n"
                     "property float z\nproperty float mass\nend_header\n" % len(pxm)).encode("ascii"))
            f.write(pxm.tobytes())

    def read_particles(self, scene: int = 0) -> dict:
        n = self.lib.snow_num_particles(self.h, scene)
        out = {"x": np.empty((n, 3), np.float32), "v": np.empty((n, 3), np.float32),
               "fe": np.empty((n, 9), np.float32), "fp": np.empty((n, 9), np.float32),
               "rho0": np.empty(n, np.float32)}
        self._chk(self.lib.snow_read_particles(self.h, scene, ptr(out["x"]), ptr(out["v"]), ptr(out["fe"]),
                                               ptr(out["fp"]), ptr(out["rho0"])))
        return out

    def write_particles(self, scene: int = 0, x=None, v=None, fe=None, fp=None, rho0=None):
        conv = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (x, v, fe, fp, rho0)]
        self._chk(self.lib.snow_write_particles(self.h, scene, *(ptr(a) for a in conv)))

    def read_grid(self, scene: int = 0, fields=("m", "mv", "f", "vstar")) -> dict:
        g = self.scenes[scene].grid
        nn = g.num_nodes
        out = {}
        shapes = {"m": (nn,), "mv": (nn, 3), "f": (nn, 3), "vstar": (nn, 3)}
        for k in fields:
            out[k] = np.empty(shapes[k], np.float32)
        self._chk(self.lib.snow_read_grid(self.h, scene, ptr(out.get("m")), ptr(out.get("mv")), ptr(out.get("f")),
                                          ptr(out.get("vstar"))))
        return out

    def write_grid(self, scene: int = 0, m=None, mv=None, f=None, vstar=None):
        conv = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (m, mv, f, vstar)]
        self._chk(self.lib.snow_write_grid(self.h, scene, *(ptr(a) for a in conv)))

    def read_colliders(self, scene: int = 0) -> np.ndarray:
        nc = len(self.scenes[scene].colliders.colliders)
        arr = (_CCollider * max(1, nc))()
        self._chk(self.lib.snow_read_colliders(self.h, scene, arr))
        return np.array([[*c.pos, c.type, *c.vel, c.radius, *c.normal, c.friction] for c in arr[:nc]],
                        dtype=np.float32).reshape(nc, 12)

    def export_reference(self, scene: int = 0):
        n = self.lib.snow_num_particles(self.h, scene)
        pxm, pv = np.empty((n, 4), np.float32), np.empty((n, 4), np.int32)
        pfe, pfp = np.empty((n, 16), np.float32), np.empty((n, 16), np.float32)
        self._chk(self.lib.snow_export_reference_particles(self.h, scene, ptr(pxm), ptr(pv, C.c_int32), ptr(pfe),
                                                           ptr(pfp)))
        return pxm, pv, pfe, pfp

    # -- timing
    def enable_stage_timers(self, on: bool = True):
        self._chk(self.lib.snow_enable_stage_timers(self.h, 1 if on else 0))

    def stage_stats(self) -> dict:
        """Per-stage {name: (max_ns, min_ns, avg_ns, stdev_ns, calls)} of the timed calls."""
        ns = self.lib.snow_num_stages()
        a = [np.zeros(ns, np.float64) for _ in range(4)]
        calls = np.zeros(ns, np.int64)
        self._chk(self.lib.snow_stage_stats(self.h, *(x.ctypes.data_as(C.POINTER(C.c_double)) for x in a),
                                            calls.ctypes.data_as(C.POINTER(C.c_int64))))
        return {self.lib.snow_stage_name(i).decode(): (a[0][i], a[1][i], a[2][i], a[3][i], int(calls[i]))
                for i in range(ns)}

    def write_benchstats(self, path: str = "benchstats.csv"):
        """Per-stage times in the reference's benchstats.csv layout."""
        self._chk(self.lib.snow_write_benchstats(self.h, path.encode()))

    def positions_view(self, scene: int = 0) -> tuple:
        """(device address, stride in bytes, count) of the scene's float4 (x, y, z, m) particle data."""
        p, stride, n = C.c_void_p(), C.c_int64(), C.c_int64()
        self._chk(self.lib.snow_positions_view(self.h, scene, C.byref(p), C.byref(stride), C.byref(n)))
        return p.value, stride.value, n.value

    def stage_times(self) -> dict:
        ns = self.lib.snow_num_stages()
        ms, calls = np.zeros(ns, np.float64), np.zeros(ns, np.int64)
        self._chk(self.lib.snow_stage_times(self.h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                            calls.ctypes.data_as(C.POINTER(C.c_int64))))
        return {self.lib.snow_stage_name(i).decode(): (float(ms[i]), int(calls[i])) for i in range(ns)}

    def reset_stage_times(self):
        self._chk(self.lib.snow_reset_stage_times(self.h))

    def active_nodes(self, scene: int = 0) -> int:
        v = C.c_int64()
        self._chk(self.lib.snow_active_nodes(self.h, scene, C.byref(v)))
        return int(v.value)

    def fp64_particles(self) -> int:
        v = C.c_int64()
        self._chk(self.lib.snow_fp64_particles(self.h, C.byref(v)))
        return int(v.value)


class ExplicitTimeUpdate:
    """TimeUpdate-shaped driver (timeUpdate.hpp:10-18): ``init()`` and ``update(dt)``."""

    def __init__(self, scene: PhysicalScene, flags: int = 0):
        self.scene = scene
        self.ctx = SnowContext([scene], flags)

    def init(self):
        self.ctx.init()

    def update(self, dt: float):
        self.ctx.step(dt)


class SemiImplicitTimeUpdate:
    """SemiImplicitTimeUpdate (semiimplicittimeupdate.hpp:6-13, empty in the reference): ``init()`` and
    ``update(dt)`` with the semi-implicit grid velocity solve of snow_step_semi_implicit."""

    def __init__(self, scene: PhysicalScene, flags: int = 0, beta: float = 1.0, tol: float = 1e-5,
                 max_iter: int = 100):
        self.scene = scene
        self.ctx = SnowContext([scene], flags)
        self.beta, self.tol, self.max_iter = beta, tol, max_iter

    def init(self):
        self.ctx.init()

    def update(self, dt: float):
        self.ctx.step_semi_implicit(dt, self.beta, self.tol, self.max_iter)


class MPMPhysicEngine:
    """physicEngine / MPMPhysicEngine (physicEngine.hpp:7-16, MPMPhysicEngine.cpp:3-7)."""

    def __init__(self, integration: ExplicitTimeUpdate):
        self.integration = integration

    def init(self) -> bool:
        self.integration.init()
        return True

    def update(self, dt: float):
        self.integration.update(dt)
