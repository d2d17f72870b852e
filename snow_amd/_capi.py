"""ctypes binding of ``libsnow_amd.so`` (declared in ``include/snow.h``).

The library is the product: nothing here falls back to a CPU path. Loading fails loudly
when the shared library is missing; build it with ``python -m snow_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsnow_amd.so")

SNOW_OK = 0
SNOW_E_ARG = -1
SNOW_E_HIP = -2
SNOW_E_NONFINITE = -3
SNOW_E_CAPACITY = -4
SNOW_E_NODEVICE = -5
SNOW_E_STATE = -6

FLAG_COMPENSATED_G2P = 1 << 0
FLAG_MIXED = 1 << 1
FLAG_KEEP_GRID = 1 << 2

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)


class Material(C.Structure):
    _fields_ = [
        ("young", C.c_float), ("poisson", C.c_float), ("hardening", C.c_float),
        ("crit_comp", C.c_float), ("crit_stretch", C.c_float), ("flip_alpha", C.c_float),
        ("gravity", C.c_float * 3), ("grad_scale", C.c_float),
    ]


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int32 * 3), ("origin", C.c_float * 3), ("spacing", C.c_float)]


class Collider(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3), ("type", C.c_float), ("vel", C.c_float * 3), ("radius", C.c_float),
        ("normal", C.c_float * 3), ("friction", C.c_float),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_particles", C.c_int64), ("x", f32p), ("v", f32p), ("mass", f32p), ("fe", f32p), ("fp", f32p),
        ("grid", Grid), ("material", Material), ("num_colliders", C.c_int32),
        ("colliders", C.POINTER(Collider)), ("dt", C.c_float),
    ]


# (name, restype, argtypes) of every exported symbol of include/snow.h
SIGNATURES = [
    ("snow_material_defaults", None, [C.POINTER(Material)]),
    ("snow_grid_defaults", None, [C.POINTER(Grid)]),
    ("snow_create", C.c_int, [C.POINTER(SceneDesc), C.c_uint32, C.POINTER(C.c_void_p)]),
    ("snow_create_batch", C.c_int, [C.POINTER(SceneDesc), C.c_int32, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("snow_destroy", C.c_int, [C.c_void_p]),
    ("snow_last_error", C.c_char_p, [C.c_void_p]),
    ("snow_num_scenes", C.c_int, [C.c_void_p]),
    ("snow_num_particles", C.c_int64, [C.c_void_p, C.c_int32]),
    ("snow_init", C.c_int, [C.c_void_p]),
    ("snow_step", C.c_int, [C.c_void_p, C.c_double]),
    ("snow_step_batch", C.c_int, [C.c_void_p]),
    ("snow_run_graph", C.c_int, [C.c_void_p, C.c_int32]),
    ("snow_step_semi_implicit", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int32]),
    ("snow_implicit_stats", C.c_int, [C.c_void_p, C.c_int32, i32p, f64p]),
    ("snow_autotune", C.c_int, [C.c_void_p, C.c_int32, i32p, i32p, i32p, i32p]),
    ("snow_stage_grid_mass", C.c_int, [C.c_void_p]),
    ("snow_stage_particle_volume", C.c_int, [C.c_void_p]),
    ("snow_stage_rigid_advance", C.c_int, [C.c_void_p, C.c_double]),
    ("snow_stage_grid_reset", C.c_int, [C.c_void_p]),
    ("snow_stage_p2g", C.c_int, [C.c_void_p]),
    ("snow_stage_grid_update", C.c_int, [C.c_void_p, C.c_double]),
    ("snow_stage_g2p", C.c_int, [C.c_void_p]),
    ("snow_stage_particle_advance", C.c_int, [C.c_void_p, C.c_double]),
    ("snow_stage_g2p_advance", C.c_int, [C.c_void_p, C.c_double]),
    ("snow_read_particles", C.c_int, [C.c_void_p, C.c_int32, f32p, f32p, f32p, f32p, f32p]),
    ("snow_read_particles_noalloc", C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_int64, f32p, f32p, f32p, f32p,
                                             f32p]),
    ("snow_write_particles", C.c_int, [C.c_void_p, C.c_int32, f32p, f32p, f32p, f32p, f32p]),
    ("snow_read_grid", C.c_int, [C.c_void_p, C.c_int32, f32p, f32p, f32p, f32p]),
    ("snow_write_grid", C.c_int, [C.c_void_p, C.c_int32, f32p, f32p, f32p, f32p]),
    ("snow_read_colliders", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Collider)]),
    ("snow_import_reference_particles", C.c_int,
     [C.c_int64, f32p, i32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p]),
    ("snow_export_reference_particles", C.c_int, [C.c_void_p, C.c_int32, f32p, i32p, f32p, f32p]),
    ("snow_sync", C.c_int, [C.c_void_p]),
    ("snow_check_finite", C.c_int, [C.c_void_p]),
    ("snow_enable_stage_timers", C.c_int, [C.c_void_p, C.c_int32]),
    ("snow_num_stages", C.c_int, []),
    ("snow_stage_name", C.c_char_p, [C.c_int32]),
    ("snow_stage_times", C.c_int, [C.c_void_p, f64p, i64p]),
    ("snow_reset_stage_times", C.c_int, [C.c_void_p]),
    ("snow_stage_stats", C.c_int, [C.c_void_p, f64p, f64p, f64p, f64p, i64p]),
    ("snow_write_benchstats", C.c_int, [C.c_void_p, C.c_char_p]),
    ("snow_positions_view", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), i64p, i64p]),
    ("snow_active_nodes", C.c_int, [C.c_void_p, C.c_int32, i64p]),
    ("snow_fp64_particles", C.c_int, [C.c_void_p, i64p]),
]

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library (cached). Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m snow_amd.build`")
    lib = C.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a: np.ndarray | None, ctype=C.c_float):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class SnowError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"snow error {code}: {msg}")
        self.code = code
