"""The semi-implicit step (snow_step_semi_implicit) against the FP64 oracle's restatement of the
published method (Stomakhin et al. 2013, eq. 9; the reference's SemiImplicitTimeUpdate is empty,
semiimplicittimeupdate.cpp:3-5), plus its invariants: beta = 0 is the explicit step bit for bit, runs
repeat bit for bit, a scene gives the same bits alone or in a batch.

Tolerances: the device solves in FP32 with the rotation of the reference SVD (V rounded to 1e-5),
the oracle in FP64 with the exact polar factor; both stop at the same relative residual. Differences
are therefore of the order of the FP32 rounding of the explicit step plus 1e-5 of the implicit
correction.
"""
import numpy as np
import pytest

from snow_amd import SnowContext, SnowError
from snow_amd import scenes

pytestmark = pytest.mark.gpu


def deformed(scene, amp=0.03, seed=5):
    rng = np.random.default_rng(seed)
    n = len(scene.particles)
    fe = np.eye(3)[None] + amp * rng.uniform(-1, 1, (n, 3, 3))
    fp = np.eye(3)[None] + 0.3 * amp * rng.uniform(-1, 1, (n, 3, 3))
    scene.particles.fe = np.ascontiguousarray(fe.transpose(0, 2, 1).reshape(n, 9), np.float32)
    scene.particles.fp = np.ascontiguousarray(fp.transpose(0, 2, 1).reshape(n, 9), np.float32)
    return scene


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def test_beta_zero_is_the_explicit_step():
    sc = deformed(scenes.snowball_drop())
    with SnowContext(sc) as a, SnowContext(sc) as b:
        a.init()
        b.init()
        for _ in range(4):
            a.step(sc.dt)
            b.step_semi_implicit(sc.dt, beta=0.0)
            assert b.implicit_stats()[0] == 0
        pa, pb = a.read_particles(), b.read_particles()
    for k in ("x", "v", "fe", "fp"):
        assert np.array_equal(pa[k], pb[k]), k


@pytest.mark.parametrize("beta,grad_scale", [(1.0, 1.0), (0.5, 1.0), (1.0, 0.0), (0.5, 0.0)])
def test_one_step_teacher_forced(oracle_mod, beta, grad_scale):
    """From the same deformed state, one step at ten times the scene's dt: the GPU against the oracle,
    and the implicit correction (semi-implicit minus explicit) well above that difference. With the
    reference's weight gradient (grad_scale 1) the stiffness term is h^2 smaller than with the
    physical 1/h (grad_scale 0), whose solve takes several iterations."""
    sc = deformed(scenes.snowball_drop(grad_scale=grad_scale))
    sc.particles.v[:] = np.random.default_rng(3).normal(0, 0.5, sc.particles.v.shape).astype(np.float32)
    dt, tol = 1e-3, 1e-6
    o = oracle_mod.Oracle(sc)
    o.init()
    o.rho0[:] = o.rho0.astype(np.float32)
    print(f"\nscene: snowball drop, {len(sc.particles)} particles, dt {dt}, beta {beta}, grad_scale {grad_scale}",
          flush=True)
    with SnowContext(sc) as ctx, SnowContext(sc) as ex:
        for c in (ctx, ex):
            c.init()
            c.write_particles(rho0=o.rho0.astype(np.float32))
        ctx.step_semi_implicit(dt, beta=beta, tol=tol, max_iter=200)
        it_gpu, rr_gpu = ctx.implicit_stats()
        g = ctx.read_particles()
        ex.step(dt)
        e = ex.read_particles()
    it_o, rr_o = o.step_semi_implicit(dt, beta, tol, 200)
    print(f"iterations GPU {it_gpu} (residual {rr_gpu:.2e}), oracle {it_o} (residual {rr_o:.2e})")
    dx, dv = np.abs(g["x"] - o.x).max(), rel_err(g["v"], o.v)
    dfe, dfp = np.abs(g["fe"] - o.fe).max(), np.abs(g["fp"] - o.fp).max()
    corr = rel_err(e["v"], g["v"])
    print(f"x {dx:.2e}  v rel {dv:.2e}  F_E {dfe:.2e}  F_P {dfp:.2e}  implicit correction of v (rel) {corr:.2e}")
    assert rr_gpu <= tol and rr_o <= tol
    assert 0 < it_gpu < 200 and 0 < it_o < 200
    assert dx < 1e-6
    assert dv < 1e-4
    assert dfe < 1e-4 and dfp < 1e-4
    assert corr > 20 * dv


@pytest.mark.parametrize("grad_scale", [1.0, 0.0])
def test_free_running_against_oracle(oracle_mod, grad_scale):
    """20 semi-implicit steps of the deformed snowball at 5x its dt, free running."""
    sc = deformed(scenes.snowball_drop(grad_scale=grad_scale), amp=0.02)
    dt, steps = 5e-4, 20
    o = oracle_mod.Oracle(sc, nthreads=8)
    o.init()
    o.rho0[:] = o.rho0.astype(np.float32)
    with SnowContext(sc) as ctx:
        ctx.init()
        ctx.write_particles(rho0=o.rho0.astype(np.float32))
        for _ in range(steps):
            ctx.step_semi_implicit(dt, beta=1.0, tol=1e-6, max_iter=200)
        ctx.check_finite()
        ctx_it = ctx.implicit_stats()[0]
        g = ctx.read_particles()
    for _ in range(steps):
        o.step_semi_implicit(dt, 1.0, 1e-6, 200)
    dx, dv = np.abs(g["x"] - o.x).max(), rel_err(g["v"], o.v)
    dfe = np.abs(g["fe"] - o.fe).max()
    print(f"\n{steps} steps (grad_scale {grad_scale}, last solve {ctx_it} iterations): x {dx:.2e}  v rel {dv:.2e}  "
          f"F_E {dfe:.2e}")
    assert dx < 1e-5
    # FP32 against FP64 (and the reference SVD's rotation against the exact polar factor) drift apart
    # over the free-running steps; with the physical gradient the system is stiffer and the drift
    # of v larger
    assert dv < (1e-3 if grad_scale == 1.0 else 5e-3)
    assert dfe < 1e-3


def test_deterministic_and_batch_equals_single():
    a_sc = deformed(scenes.snowball_drop(), seed=1)
    b_sc = deformed(scenes.tiny_block(n_side=16, dims=32, velocity=(0.3, -0.5, 0.2)), seed=2)
    dt = 5e-4
    out = []
    for scs in ([a_sc], [a_sc], [a_sc, b_sc]):
        with SnowContext(scs) as ctx:
            ctx.init()
            for _ in range(3):
                ctx.step_semi_implicit(dt, beta=1.0, tol=1e-6, max_iter=100)
            out.append([ctx.read_particles(s) for s in range(len(scs))])
    with SnowContext([b_sc]) as ctx:
        ctx.init()
        for _ in range(3):
            ctx.step_semi_implicit(dt, beta=1.0, tol=1e-6, max_iter=100)
        single_b = ctx.read_particles()
    for k in ("x", "v", "fe", "fp"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k  # run to run
        assert np.array_equal(out[0][0][k], out[2][0][k]), k  # scene a alone / in the batch
        assert np.array_equal(single_b[k], out[2][1][k]), k  # scene b alone / in the batch


def test_argument_checks():
    sc = scenes.tiny_block(n_side=4, dims=8)
    with SnowContext(sc) as ctx:
        with pytest.raises(SnowError):
            ctx.implicit_stats()  # no solve yet
        ctx.init()
        for kw in ({"beta": -0.1}, {"beta": 1.5}, {"tol": 0.0}, {"max_iter": -1}):
            with pytest.raises(SnowError):
                ctx.step_semi_implicit(1e-4, **kw)
        ctx.step_semi_implicit(1e-4)
        it, rr = ctx.implicit_stats()
        assert it >= 0 and rr >= 0.0


def test_cpp_driver_semi_implicit_matches_python_binding(tmp_path):
    """snow_run --semi-implicit (snow_amd::SnowSemiImplicitTimeUpdate over the C ABI) gives the same
    bits as the ctypes binding for init + 5 semi-implicit steps of the deformed 64^3 snowball."""
    import os
    import subprocess

    from snow_amd import load_run_output
    sc = deformed(scenes.snowball_drop())
    sc.dt = 5e-4
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "snow_amd", "snow_run")
    scene_file, out_file = str(tmp_path / "c1.scn"), str(tmp_path / "c1.out")
    sc.save(scene_file)
    print(f"\n[cpp driver] {exe} {scene_file} 5 --semi-implicit 1.0", flush=True)
    res = subprocess.run([exe, scene_file, "5", out_file, "--semi-implicit", "1.0"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    got = load_run_output(out_file)
    with SnowContext(sc) as ctx:
        ctx.init()
        for _ in range(5):
            ctx.step_semi_implicit(sc.dt, beta=1.0, tol=1e-5, max_iter=100)
        assert ctx.implicit_stats()[0] > 0
        ref = ctx.read_particles()
    for k in ("x", "v", "fe", "fp"):
        assert np.array_equal(got[k], ref[k]), k


def test_empty_scene_in_a_batch():
    """A scene without particles has no node block with mass: its solve is done at its first residual
    (0 iterations), and the other scene of the batch gives the same bits as alone."""
    from snow_amd import ParticleSystem
    empty = scenes.tiny_block(n_side=4, dims=16)
    empty.particles = ParticleSystem()
    full = deformed(scenes.snowball_drop(grad_scale=0.0), seed=4)
    dt = 5e-4
    with SnowContext(full) as ctx:
        ctx.init()
        for _ in range(3):
            ctx.step_semi_implicit(dt, beta=1.0, tol=1e-6, max_iter=100)
        single = ctx.read_particles()
        it_single = ctx.implicit_stats()[0]
    with SnowContext([empty, full]) as ctx:
        ctx.init()
        for _ in range(3):
            ctx.step_semi_implicit(dt, beta=1.0, tol=1e-6, max_iter=100)
        ctx.check_finite()
        assert ctx.implicit_stats(0)[0] == 0
        assert ctx.implicit_stats(1)[0] == it_single > 0
        batch = ctx.read_particles(1)
    for k in ("x", "v", "fe", "fp"):
        assert np.array_equal(single[k], batch[k]), k


def test_partial_tiles_against_oracle(oracle_mod):
    """A 37^3 grid (the last node tiles are partial) with the block near the grid's edge, physical
    gradient: one step from a deformed state against the oracle, as the teacher-forced test."""
    sc = deformed(scenes.tiny_block(n_side=14, dims=37, center=(0.5, 0.25, 0.8), velocity=(0.4, -1.5, 0.3),
                                    grad_scale=0.0), seed=9)
    dt, tol = 1e-3, 1e-6
    o = oracle_mod.Oracle(sc)
    o.init()
    o.rho0[:] = o.rho0.astype(np.float32)
    with SnowContext(sc) as ctx:
        ctx.init()
        ctx.write_particles(rho0=o.rho0.astype(np.float32))
        ctx.step_semi_implicit(dt, beta=1.0, tol=tol, max_iter=200)
        it_gpu, rr_gpu = ctx.implicit_stats()
        g = ctx.read_particles()
    it_o, rr_o = o.step_semi_implicit(dt, 1.0, tol, 200)
    dx, dv = np.abs(g["x"] - o.x).max(), rel_err(g["v"], o.v)
    print(f"\niterations GPU {it_gpu} ({rr_gpu:.2e}), oracle {it_o} ({rr_o:.2e}); x {dx:.2e}  v rel {dv:.2e}")
    assert rr_gpu <= tol and rr_o <= tol
    assert it_gpu > 0 and abs(it_gpu - it_o) <= 1
    assert dx < 1e-6
    assert dv < 1e-4
