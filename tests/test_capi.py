"""C-ABI boundary checks that need no GPU: the library loads, exports every function
include/snow.h declares, validates descriptors before touching a device, and the host-only
helpers (defaults, reference-encoding import) behave as the reference defines them."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from snow_amd import _capi
from snow_amd._capi import SceneDesc, load_library, ptr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "snow.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(snow_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = load_library()
    names = declared_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (snow_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(lib, n)
    # every declared function has a ctypes signature in the binding
    assert set(names) == {s[0] for s in _capi.SIGNATURES}


def test_library_has_gfx950_code_object():
    # the offload bundle of our own build names its target triple
    data = open(_capi.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_defaults_match_defines_hpp():
    lib = load_library()
    m = _capi.Material()
    lib.snow_material_defaults(C.byref(m))
    assert m.young == pytest.approx(6e6) and m.poisson == pytest.approx(0.2) and m.hardening == pytest.approx(30.0)
    assert m.crit_comp == pytest.approx(1e-2) and m.crit_stretch == pytest.approx(1e-3)
    assert m.flip_alpha == pytest.approx(0.95) and tuple(m.gravity) == pytest.approx((0.0, -9.81, 0.0))
    assert m.grad_scale == 1.0
    g = _capi.Grid()
    lib.snow_grid_defaults(C.byref(g))
    assert tuple(g.dim) == (201, 201, 201) and g.spacing == pytest.approx(0.05)
    assert tuple(g.origin) == pytest.approx((0.5125, 0.5125, 1.5125))


def test_stage_names_follow_reference():
    lib = load_library()
    names = [lib.snow_stage_name(i).decode() for i in range(lib.snow_num_stages())]
    for ref in ("#Sim:Advance Rigids", "#Sim:Scratch Grid", "#Sim:Transfer to Grid", "#Sim:Compute Derivatives",
                "#Sim:Transfer to Particles", "#Sim:Advance Particles"):
        assert ref in names
    assert lib.snow_stage_name(99) is None


def _desc(n=4, dims=(8, 8, 8), h=0.125):
    x = np.full((n, 3), 0.5, np.float32)
    v = np.zeros((n, 3), np.float32)
    m = np.ones(n, np.float32)
    d = SceneDesc()
    d.num_particles = n
    d.x, d.v, d.mass = ptr(x), ptr(v), ptr(m)
    d.grid.dim[0], d.grid.dim[1], d.grid.dim[2] = dims
    d.grid.spacing = h
    lib = load_library()
    lib.snow_material_defaults(C.byref(d.material))
    d.dt = 1e-3
    return d, (x, v, m)


@pytest.mark.parametrize("mutate", ["dims", "spacing", "poisson", "missing_x", "neg_n", "bad_collider", "dt"])
def test_create_rejects_bad_descriptors(mutate):
    lib = load_library()
    d, keep = _desc()
    if mutate == "dims":
        d.grid.dim[1] = 0
    elif mutate == "spacing":
        d.grid.spacing = 0.0
    elif mutate == "poisson":
        d.material.poisson = 0.5
    elif mutate == "missing_x":
        d.x = None
    elif mutate == "neg_n":
        d.num_particles = -1
    elif mutate == "bad_collider":
        d.num_colliders = 1
        d.colliders = None
    elif mutate == "dt":
        d.dt = -1.0
    h = C.c_void_p()
    rc = lib.snow_create(C.byref(d), 0, C.byref(h))
    assert rc == _capi.SNOW_E_ARG
    assert h.value is None
    assert lib.snow_last_error(None).decode()


def test_create_batch_arg_checks():
    lib = load_library()
    h = C.c_void_p()
    assert lib.snow_create_batch(None, 1, 0, C.byref(h)) == _capi.SNOW_E_ARG
    d, keep = _desc()
    assert lib.snow_create_batch(C.byref(d), 0, 0, C.byref(h)) == _capi.SNOW_E_ARG
    assert lib.snow_create_batch(C.byref(d), 1, 0, None) == _capi.SNOW_E_ARG


def test_null_context_calls_fail_cleanly():
    lib = load_library()
    assert lib.snow_init(None) == _capi.SNOW_E_ARG
    assert lib.snow_step(None, 1e-3) == _capi.SNOW_E_ARG
    assert lib.snow_destroy(None) == _capi.SNOW_E_ARG
    assert lib.snow_sync(None) == _capi.SNOW_E_ARG
    assert lib.snow_stage_p2g(None) == _capi.SNOW_E_ARG


def test_import_reference_encodings():
    """pxm vec4 (x, m), pv ivec4 (v*1e8, rho0*1e6), mat4 column-major (particlesystem.cpp:101-228)."""
    lib = load_library()
    n = 3
    pxm = np.array([[0.1, 0.2, 0.3, 1e-3], [1.0, 2.0, 3.0, 2e-3], [0.0, 0.0, 0.0, 5.0]], np.float32)
    pv = np.array([[10 * 10 ** 8 // 10, -12 * 10 ** 7, 0, 0], [0, 1, -1, 0], [5, 0, 0, 0]], np.int32)
    F = np.arange(16, dtype=np.float32).reshape(4, 4)  # m[col][row]
    pfe = np.tile(F.reshape(1, 16), (n, 1))
    x, v, m = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32), np.empty(n, np.float32)
    fe, fp = np.empty((n, 9), np.float32), np.empty((n, 9), np.float32)
    rc = lib.snow_import_reference_particles(n, ptr(pxm), ptr(pv, C.c_int32), ptr(pfe), None, ptr(x), ptr(v), ptr(m),
                                             ptr(fe), ptr(fp))
    assert rc == 0
    np.testing.assert_array_equal(x, pxm[:, :3])
    np.testing.assert_array_equal(m, pxm[:, 3])
    np.testing.assert_allclose(v[0], [1.0, -1.2, 0.0], rtol=1e-6)
    # column c, row r of the mat4 -> fe[c*3 + r]
    for c in range(3):
        for r in range(3):
            assert fe[0, c * 3 + r] == F[c, r]
    np.testing.assert_array_equal(fp, np.tile(np.eye(3, dtype=np.float32).reshape(1, 9), (n, 1)))


def test_cpp_driver_built_and_checks_input(tmp_path):
    """The C++ driver over include/snow_time_update.hpp is built and rejects bad input cleanly."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "snow_amd", "snow_run")
    assert os.access(exe, os.X_OK)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2 and "usage" in res.stderr
    bad = tmp_path / "bad.scn"
    bad.write_bytes(b"NOTASCENE" + b"\0" * 64)
    res = subprocess.run([exe, str(bad), "1", str(tmp_path / "o")], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2 and "SNOWSCN1" in res.stderr


def test_scene_file_roundtrip(tmp_path):
    """PhysicalScene.save writes the SNOWSCN1 layout the C++ driver reads (header sizes, order)."""
    import numpy as np
    from snow_amd import scenes
    sc = scenes.tiny_block(n_side=3, dims=16)
    p = tmp_path / "t.scn"
    sc.save(str(p))
    raw = p.read_bytes()
    n = len(sc.particles)
    ncol = len(sc.colliders.colliders)
    assert raw[:8] == b"SNOWSCN1"
    assert int(np.frombuffer(raw[8:16], np.int64)[0]) == n
    assert len(raw) == 8 + 8 + 28 + 40 + 4 + 48 * ncol + 4 + 4 * 25 * n
    dims = np.frombuffer(raw[16:28], np.int32)
    assert tuple(dims) == (sc.grid.dimx, sc.grid.dimy, sc.grid.dimz)
    x = np.frombuffer(raw[len(raw) - 100 * n:len(raw) - 88 * n], np.float32).reshape(n, 3)
    assert np.array_equal(x, sc.particles.x)
    fp = np.frombuffer(raw[len(raw) - 36 * n:], np.float32).reshape(n, 9)
    assert np.array_equal(fp, sc.particles.fp)
