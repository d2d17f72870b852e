"""Parity of the HIP path (libsnow_amd.so, through the C ABI) with the FP64 CPU oracle.

Every stage is checked teacher-forced (GPU and oracle start from the same state), then the
fused step over many steps, then size-independent properties at the 1M-particle size.
Tolerances are stated per quantity; the FP32 device path and the FP64 oracle follow the same
algorithm, so differences come from FP32 rounding, including the 1e-5 rounding of V in the SVD
(SVD.include.glsl:354-358), which can flip between FP32 and FP64 for one entry.
"""
import numpy as np
import pytest

from snow_amd import (FLAG_COMPENSATED_G2P, FLAG_KEEP_GRID, FLAG_MIXED, SPHERE, Collider, CollisionObjects, Material,
                      ParticleSystem, SnowContext)
from snow_amd import scenes

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() / scale


def deformed(scene, amp=0.02, seed=5):
    """Give every particle a random F_E near I and F_P near I (so forces are not rounding noise)."""
    rng = np.random.default_rng(seed)
    n = len(scene.particles)
    fe = np.eye(3)[None] + amp * rng.uniform(-1, 1, (n, 3, 3))
    fp = np.eye(3)[None] + 0.5 * amp * rng.uniform(-1, 1, (n, 3, 3))
    # column-major per particle: element (r, c) at c*3 + r
    scene.particles.fe = np.ascontiguousarray(fe.transpose(0, 2, 1).reshape(n, 9), np.float32)
    scene.particles.fp = np.ascontiguousarray(fp.transpose(0, 2, 1).reshape(n, 9), np.float32)
    return scene


def clamp_deformed(scene, seed=11, spread=0.01):
    """F_E = U diag(s) V^T with random rotations U, V and singular values s spread uniformly over
    [1 - theta_c - spread, 1 + theta_s + spread]: the update's singular values then fall on both sides
    of the clamp bounds (particleAdvance.glsl:112-117), and some within the FP64 path's 2e-4 of them.
    F_P = I."""
    rng = np.random.default_rng(seed)
    n = len(scene.particles)
    mat = scene.material
    lo, hi = 1.0 - mat.crit_comp - spread, 1.0 + mat.crit_stretch + spread

    def rotations(k):
        q, r = np.linalg.qr(rng.normal(size=(k, 3, 3)))
        q = q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]
        q[np.linalg.det(q) < 0, :, 0] *= -1.0
        return q
    u, v = rotations(n), rotations(n)
    sv = rng.uniform(lo, hi, (n, 3))
    fe = u @ (sv[:, :, None] * v.transpose(0, 2, 1))
    scene.particles.fe = np.ascontiguousarray(fe.transpose(0, 2, 1).reshape(n, 9), np.float32)
    scene.particles.fp = np.tile(np.eye(3, dtype=np.float32).reshape(1, 9), (n, 1))
    return scene


def with_colliders(scene):
    scene.colliders = CollisionObjects(scene.colliders.colliders + [
        Collider((0.5, 0.45, 0.5), SPHERE, velocity=(0.0, 0.5, 0.0), radius=0.08, friction=0.3)])
    return scene


@pytest.fixture(scope="module")
def c1_scene():
    return scenes.snowball_drop()


def test_init_rest_density(oracle_mod, c1_scene):
    o = oracle_mod.Oracle(c1_scene)
    o.init()
    with SnowContext(c1_scene) as ctx:
        ctx.init()
        g = ctx.read_particles()
    assert rel_err(g["rho0"], o.rho0) < 1e-5  # FP32 sums of 64 terms vs FP64


@pytest.mark.parametrize("which", ["tiny", "c1"])
def test_p2g_teacher_forced(oracle_mod, c1_scene, which):
    sc = deformed(scenes.tiny_block(n_side=10, dims=16) if which == "tiny" else c1_scene)
    o = oracle_mod.Oracle(sc)
    o.init()
    o.p2g()
    with SnowContext(sc) as ctx:
        ctx.write_particles(rho0=o.rho0.astype(np.float32))
        ctx.stage("p2g")
        g = ctx.read_grid()
    assert rel_err(g["m"], o.gm) < 2e-6
    assert rel_err(g["mv"], o.gmv) < 2e-6
    # force: F_E-dependent stress through the FP32 SVD; tolerance covers V-rounding flips
    assert rel_err(g["f"], o.gf) < 2e-3
    assert np.abs(g["f"].sum(0)).max() < 1e-4 * np.abs(g["f"]).sum()


def test_p2g_out_of_grid(oracle_mod):
    """Particles whose stencils leave the grid (or lie far outside) drop those nodes (p2gTransfer.glsl:87)."""
    sc = scenes.tiny_block(n_side=6, dims=12, center=(0.04, 0.5, 0.97))
    extra = np.array([[5.0, 5.0, 5.0], [-0.2, 0.5, 0.5], [0.5, -0.15, 0.5]], np.float32)
    ps = sc.particles
    ps.x = np.concatenate([ps.x, extra])
    ps.v = np.concatenate([ps.v, np.zeros_like(extra)])
    ps.mass = np.concatenate([ps.mass, np.full(3, ps.mass[0], np.float32)])
    ps.fe = np.concatenate([ps.fe, np.tile(ps.fe[:1], (3, 1))])
    ps.fp = np.concatenate([ps.fp, np.tile(ps.fp[:1], (3, 1))])
    o = oracle_mod.Oracle(sc)
    o.init()
    o.p2g()
    with SnowContext(sc) as ctx:
        ctx.write_particles(rho0=o.rho0.astype(np.float32))
        ctx.stage("p2g")
        g = ctx.read_grid()
    assert o.gm.sum() < ps.mass.sum() * 0.999
    assert rel_err(g["m"], o.gm) < 2e-6
    assert rel_err(g["mv"], o.gmv) < 2e-6


def test_p2g_bound_follows_host_writes(oracle_mod):
    """The P2G's fixed-point scale comes from bounds published by the kernels that wrote the records.
    After fused steps, a host write that makes every momentum ~1e5 times larger must not reuse the
    published bound (the sums would leave the fixed-point range): the P2G must match the oracle,
    and graph replays from the written state must equal eager steps bit for bit."""
    sc = deformed(scenes.tiny_block(n_side=10, dims=16))
    dt = sc.dt
    with SnowContext(sc) as a, SnowContext(sc) as b:
        for ctx in (a, b):
            ctx.init()
            for _ in range(3):
                ctx.step(dt)
        st = a.read_particles()
        v_big = (st["v"] * 1000.0 + np.array([200.0, -50.0, 30.0])).astype(np.float32)
        for ctx in (a, b):
            ctx.write_particles(v=v_big)
        a.stage("p2g")
        g = a.read_grid(fields=("m", "mv"))
        for _ in range(2):
            a.step(dt)
        b.run_graph(2)
        pa, pb = a.read_particles(), b.read_particles()
    sc2 = scenes.tiny_block(n_side=10, dims=16)
    sc2.particles.x, sc2.particles.v = st["x"].copy(), v_big.copy()
    sc2.particles.fe, sc2.particles.fp = st["fe"].copy(), st["fp"].copy()
    o = oracle_mod.Oracle(sc2)
    o.rho0[:] = st["rho0"]
    o.p2g()
    assert rel_err(g["m"], o.gm) < 2e-6
    assert rel_err(g["mv"], o.gmv) < 2e-6
    for k in ("x", "v", "fe", "fp"):
        assert np.array_equal(pa[k], pb[k]), k


def test_grid_update_teacher_forced(oracle_mod, c1_scene):
    sc = with_colliders(deformed(scenes.snowball_drop()))
    o = oracle_mod.Oracle(sc)
    o.init()
    o.p2g()
    dt = 1e-4
    with SnowContext(sc) as ctx:
        ctx.write_grid(m=o.gm, mv=o.gmv, f=o.gf)
        ctx.stage("grid_update", dt)
        g = ctx.read_grid(fields=("vstar",))
    o.grid_update(dt)
    assert rel_err(g["vstar"], o.gvstar) < 1e-5


def test_g2p_advance_teacher_forced(oracle_mod):
    sc = with_colliders(deformed(scenes.snowball_drop(), amp=0.01))
    sc.particles.v[:] = np.random.default_rng(3).normal(0, 0.5, sc.particles.v.shape).astype(np.float32)
    o = oracle_mod.Oracle(sc)
    o.init()
    o.rigid_advance(1e-4)
    o.p2g()
    dt = 1e-4
    o.grid_update(dt)
    gm, gmv, gvs = o.gm.copy(), o.gmv.copy(), o.gvstar.copy()
    with SnowContext(sc) as ctx:
        ctx.write_particles(rho0=o.rho0.astype(np.float32))
        ctx.stage("rigid_advance", dt)
        ctx.write_grid(m=gm, mv=gmv, vstar=gvs)
        ctx.stage("g2p")
        ctx.stage("particle_advance", dt)
        g = ctx.read_particles()
    o.g2p_advance(dt)
    assert np.abs(g["x"] - o.x).max() < 1e-6
    assert rel_err(g["v"], o.v) < 1e-5
    assert np.abs(g["fe"] - o.fe).max() < 1e-4
    assert np.abs(g["fp"] - o.fp).max() < 1e-4


def test_grid_update_production_kernel(oracle_mod):
    """The tiled grid merge + update that snow_step runs (k_grid<false, true>), from a GPU P2G,
    against the oracle's gForceCompute on exactly the node fields that P2G produced."""
    sc = with_colliders(clamp_deformed(scenes.snowball_drop()))
    sc.particles.v[:] = np.random.default_rng(4).normal(0, 0.5, sc.particles.v.shape).astype(np.float32)
    dt = 1e-4
    o = oracle_mod.Oracle(sc)
    o.init()
    with SnowContext(sc) as ctx:
        ctx.write_particles(rho0=o.rho0.astype(np.float32))
        ctx.stage("p2g")
        g = ctx.read_grid(fields=("m", "mv", "f"))
        ctx.stage("grid_update", dt)
        gv = ctx.read_grid(fields=("vstar",))
    o.gm[:], o.gmv[:], o.gf[:] = g["m"], g["mv"], g["f"]
    o.grid_update(dt)
    assert (g["m"] > 0).sum() > 1000
    assert rel_err(gv["vstar"], o.gvstar) < 1e-5  # FP32 vs FP64 on the same input


def _g2p_advance_production(oracle_mod, flags):
    """One fused G2P + particle advance (the kernel snow_step runs) from the oracle's grid and the
    same particle state, singular values of F_E around the clamp bounds. Returns the GPU state, the
    oracle and the number of particles that took the FP64 path."""
    sc = with_colliders(clamp_deformed(scenes.snowball_drop()))
    sc.particles.v[:] = np.random.default_rng(3).normal(0, 0.5, sc.particles.v.shape).astype(np.float32)
    dt = 1e-4
    o = oracle_mod.Oracle(sc)
    o.init()
    o.p2g()
    o.grid_update(dt)
    gm, gmv, gvs = o.gm.copy(), o.gmv.copy(), o.gvstar.copy()
    with SnowContext(sc, flags) as ctx:
        ctx.write_particles(rho0=o.rho0.astype(np.float32))
        ctx.write_grid(m=gm, mv=gmv, vstar=gvs)
        ctx.stage("g2p_advance", dt)
        g = ctx.read_particles()
        fp64 = ctx.fp64_particles()
    o.g2p_advance(dt)
    return g, o, fp64


@pytest.mark.parametrize("flags", [0, FLAG_COMPENSATED_G2P, FLAG_MIXED, FLAG_MIXED | FLAG_COMPENSATED_G2P],
                         ids=["plain", "compensated", "mixed", "mixed_compensated"])
def test_g2p_advance_production_kernel(oracle_mod, flags):
    g, o, fp64 = _g2p_advance_production(oracle_mod, flags)
    # the stored F_E is W S' V^T with V rounded to 1e-5 (SVD.include.glsl:354-358), so the singular
    # values of a clamped F_E lie within a few 1e-5 of the bound
    sv = np.linalg.svd(o.fe.reshape(-1, 3, 3), compute_uv=False)
    mat = scenes.snowball_drop().material
    at_clamp = (np.abs(sv - (1 - mat.crit_comp)) < 5e-5) | (np.abs(sv - (1 + mat.crit_stretch)) < 5e-5)
    assert at_clamp.any(-1).mean() > 0.3  # the clamp is active for many particles
    assert np.abs(g["x"] - o.x).max() < 1e-6
    assert rel_err(g["v"], o.v) < 1e-5
    assert np.abs(g["fe"] - o.fe).max() < 1e-4
    assert np.abs(g["fp"] - o.fp).max() < 1e-4
    if flags & FLAG_MIXED:
        assert fp64 > 0
    else:
        assert fp64 == 0


def test_mixed_fp64_path_not_worse_teacher_forced(oracle_mod):
    """From the same state, the mixed form (FP64 SVD and force matrix for particles near the clamp)
    has no larger F_E / F_P error against the FP64 oracle than the plain FP32 form: the particles it
    does not send to the FP64 path are computed as in the plain form."""
    gp, o, _ = _g2p_advance_production(oracle_mod, 0)
    gm, _, fp64 = _g2p_advance_production(oracle_mod, FLAG_MIXED)
    assert fp64 > 0
    print(f"\nFP64 path: {fp64} of {len(o.x)} particles")
    for k in ("fe", "fp"):
        ep = np.sqrt(((gp[k] - getattr(o, k)) ** 2).sum(-1))
        em = np.sqrt(((gm[k] - getattr(o, k)) ** 2).sum(-1))
        print(f"{k}: plain max {ep.max():.3e} rms {np.sqrt((ep ** 2).mean()):.3e}; "
              f"mixed max {em.max():.3e} rms {np.sqrt((em ** 2).mean()):.3e}")
        assert np.sqrt((em ** 2).mean()) <= np.sqrt((ep ** 2).mean())
        assert em.max() <= ep.max() + 1e-6
    assert np.array_equal(gp["x"], gm["x"])  # velocity and advection do not depend on the form


def test_fused_step_matches_stage_chain(oracle_mod):
    sc = with_colliders(deformed(scenes.tiny_block(n_side=12, dims=24), amp=0.01))
    dt = 1e-4
    with SnowContext(sc, FLAG_KEEP_GRID) as a, SnowContext(sc) as b:
        a.init()
        b.init()
        for _ in range(3):
            a.step(dt)
            b.stage("rigid_advance", dt)
            b.stage("p2g")
            b.stage("grid_update", dt)
            b.stage("g2p")
            b.stage("particle_advance", dt)
        pa, pb = a.read_particles(), b.read_particles()
    # the fused step permutes particles (cell order ties broken by the previous slot) while the stage
    # chain keeps creation order, so per-node sums may run in a different order: FP32-level agreement
    assert np.abs(pa["x"] - pb["x"]).max() <= 1e-6
    assert rel_err(pa["v"], pb["v"]) <= 1e-5
    assert np.abs(pa["fe"] - pb["fe"]).max() <= 1e-4
    assert np.abs(pa["fp"] - pb["fp"]).max() <= 1e-4


def test_free_running_100_steps(oracle_mod, c1_scene):
    """Plain MPM on the 64^3 snowball drop: 100 steps FP32 GPU vs FP64 oracle."""
    sc = c1_scene
    dt = sc.dt
    o = oracle_mod.Oracle(sc, nthreads=8)
    o.init()
    o.step(dt, 100)
    with SnowContext(sc) as ctx:
        ctx.init()
        for _ in range(100):
            ctx.step(dt)
        ctx.check_finite()
        g = ctx.read_particles()
    err_x = np.abs(g["x"] - o.x).max()  # domain size 1
    dv = np.abs(g["v"] - o.v).max(axis=1) / np.abs(o.v).max()
    worst = int(np.argmax(dv))
    print(f"\n100 steps C1: max|dx|={err_x:.3e} rms={np.sqrt(((g['x'] - o.x) ** 2).mean()):.3e} "
          f"v={rel_err(g['v'], o.v):.3e} v99.9%={np.quantile(dv, 0.999):.3e} v_rms={np.sqrt((dv ** 2).mean()):.3e} "
          f"worst particle {worst} at x={o.x[worst]} v={o.v[worst]} gpu v={g['v'][worst]} "
          f"fe={np.abs(g['fe'] - o.fe).max():.3e} fp={np.abs(g['fp'] - o.fp).max():.3e}")
    assert err_x < 1e-4  # positions within 1e-4 of the domain size
    # Velocities: the singular-value clamp and the stick/slide switch of the ground friction are
    # discontinuous, so FP32 roundoff moves a few particles across them during 100 steps; those
    # particles differ by ~1e-3 of max|v|, the bulk by far less.
    assert rel_err(g["v"], o.v) < 2e-3
    assert np.quantile(dv, 0.999) < 1e-3
    assert np.abs(g["fe"] - o.fe).max() < 1e-3


def test_deterministic_bitwise(c1_scene):
    outs = []
    for _ in range(2):
        with SnowContext(c1_scene, FLAG_KEEP_GRID) as ctx:
            ctx.init()
            for _ in range(5):
                ctx.step(c1_scene.dt)
            outs.append((ctx.read_particles(), ctx.read_grid()))
    for k in ("x", "v", "fe", "fp"):
        assert np.array_equal(outs[0][0][k], outs[1][0][k]), k
    for k in ("m", "mv", "f", "vstar"):
        assert np.array_equal(outs[0][1][k], outs[1][1][k]), k


def test_batch_equals_single():
    sc = [scenes.tiny_block(n_side=8, dims=16, seed=s, velocity=(0.2 * s, -1.0, 0.1)) for s in range(3)]
    sc[1] = scenes.tiny_block(n_side=5, dims=24, seed=9)
    singles = []
    for s in sc:
        with SnowContext(s) as ctx:
            ctx.init()
            for _ in range(4):
                ctx.step(1e-4)
            singles.append(ctx.read_particles())
    with SnowContext(sc) as ctx:
        ctx.init()
        for _ in range(4):
            ctx.step(1e-4)
        batch = [ctx.read_particles(i) for i in range(3)]
    for a, b in zip(singles, batch):
        for k in ("x", "v", "fe", "fp"):
            assert np.array_equal(a[k], b[k]), k


def test_active_nodes_count_the_nodes_with_mass():
    """snow_active_nodes sums the per-scene count slots the grid update adds to; it must equal the
    nodes with m > 0 of every scene, alone and in a batch, step after step, and on the 2^20-particle
    scene whose grid update spreads the adds over all slots."""
    sc = [scenes.tiny_block(n_side=8, dims=16, seed=s, velocity=(0.2 * s, -1.0, 0.1)) for s in range(3)]
    sc[1] = scenes.tiny_block(n_side=5, dims=24, seed=9)
    for batch in (sc[:1], sc):
        with SnowContext(batch) as ctx:
            ctx.init()
            for _ in range(3):
                ctx.step(1e-4)
                for i in range(len(batch)):
                    m = ctx.read_grid(i, fields=("m",))["m"]
                    assert ctx.active_nodes(i) == int(np.count_nonzero(m > 0)), i
    big = scenes.two_snowballs()
    with SnowContext(big) as ctx:
        ctx.init()
        for _ in range(2):
            ctx.step(big.dt)
        m = ctx.read_grid(0, fields=("m",))["m"]
        assert ctx.active_nodes(0) == int(np.count_nonzero(m > 0)) > 10000


def test_readback_roundtrip_and_reference_export():
    sc = scenes.tiny_block(n_side=6, dims=16)
    with SnowContext(sc) as ctx:
        ctx.init()
        ctx.step(1e-4)
        p = ctx.read_particles()
        ctx.write_particles(x=p["x"], v=p["v"], fe=p["fe"], fp=p["fp"], rho0=p["rho0"])
        q = ctx.read_particles()
        pxm, pv, pfe, pfp = ctx.export_reference()
    for k in p:
        assert np.array_equal(p[k], q[k]), k
    np.testing.assert_array_equal(pxm[:, :3], p["x"])
    np.testing.assert_array_equal(pxm[:, 3], sc.particles.mass)
    np.testing.assert_array_equal(pv[:, :3], (p["v"] * np.float32(1e8)).astype(np.int32))
    np.testing.assert_array_equal(pfe.reshape(-1, 4, 4)[:, :3, :3].reshape(-1, 9), p["fe"])


def test_rigid_colliders_advance():
    sc = with_colliders(scenes.tiny_block())
    with SnowContext(sc) as ctx:
        ctx.init()
        for _ in range(10):
            ctx.step(1e-3)
        cols = ctx.read_colliders()
    assert cols[1, 1] == pytest.approx(0.45 + 10 * 1e-3 * 0.5, abs=1e-6)


def test_empty_scene_steps_and_advances_colliders():
    """A scene without particles steps: alone (the step start then runs as its own launch) its
    colliders advance and nothing else happens; batched with a scene that has particles (the step
    start rides on the sort) both scenes' colliders advance and the other scene's particles match
    its own context bit for bit."""
    empty = with_colliders(scenes.tiny_block())
    empty.particles = ParticleSystem()
    with SnowContext(empty) as ctx:
        ctx.init()
        for _ in range(10):
            ctx.step(1e-3)
        cols = ctx.read_colliders()
        assert ctx.read_particles()["x"].shape == (0, 3)
    assert cols[1, 1] == pytest.approx(0.45 + 10 * 1e-3 * 0.5, abs=1e-6)
    full = with_colliders(scenes.tiny_block(n_side=6, dims=16, seed=3))
    with SnowContext(full) as ctx:
        ctx.init()
        for _ in range(4):
            ctx.step(1e-4)
        single = ctx.read_particles()
    with SnowContext([empty, full]) as ctx:
        ctx.init()
        for _ in range(4):
            ctx.step(1e-4)
        batch = ctx.read_particles(1)
        cols0, cols1 = ctx.read_colliders(0), ctx.read_colliders(1)
    for k in ("x", "v", "fe", "fp"):
        assert np.array_equal(single[k], batch[k]), k
    for c in (cols0, cols1):
        assert c[1, 1] == pytest.approx(0.45 + 4 * 1e-4 * 0.5, abs=1e-6)


def test_nonfinite_detected():
    sc = scenes.tiny_block()
    with SnowContext(sc) as ctx:
        ctx.init()
        fe = np.tile(np.eye(3, dtype=np.float32).reshape(1, 9), (len(sc.particles), 1))
        fe[3, 0] = np.nan
        ctx.write_particles(fe=fe)
        with pytest.raises(Exception):
            ctx.check_finite()


def test_two_snowballs_1m_properties(oracle_mod):
    """2^20 particles on 128^3: mass on the grid equals the particle mass, state stays finite, one step
    agrees with the FP64 oracle, and the run is reproducible."""
    sc = scenes.two_snowballs()
    dt = sc.dt
    with SnowContext(sc, FLAG_KEEP_GRID) as ctx:
        ctx.init()
        ctx.step(dt)
        g = ctx.read_grid(fields=("m",))
        p = ctx.read_particles()
        assert g["m"].astype(np.float64).sum() == pytest.approx(float(sc.particles.mass.astype(np.float64).sum()),
                                                                 rel=1e-5)
        for _ in range(20):
            ctx.step(dt)
        ctx.check_finite()
    o = oracle_mod.Oracle(sc, nthreads=16)
    o.init()
    o.step(dt, 1)
    assert np.abs(p["x"] - o.x).max() < 1e-6
    assert rel_err(p["v"], o.v) < 1e-4


def test_mixed_precision_not_worse(oracle_mod):
    sc = deformed(scenes.snowball_drop(), amp=0.03)
    dt = sc.dt
    o = oracle_mod.Oracle(sc, nthreads=8)
    o.init()
    o.step(dt, 20)
    errs = {}
    for flags in (0, FLAG_MIXED):
        with SnowContext(sc, flags) as ctx:
            ctx.init()
            for _ in range(20):
                ctx.step(dt)
            g = ctx.read_particles()
        errs[flags] = (np.abs(g["x"] - o.x).max(), np.abs(g["fe"] - o.fe).max(), np.abs(g["fp"] - o.fp).max())
    print("\nplain", errs[0], "mixed", errs[FLAG_MIXED])
    assert errs[FLAG_MIXED][0] <= 2 * errs[0][0] + 1e-7


def _sweep_small():
    sc = scenes.parameter_sweep(num_scenes=6, nmin=2000, nmax=20000, dims_choices=(32, 48))
    return sc


def test_graph_replay_matches_eager():
    """Replays of the captured step equal eager batched steps bit for bit, with eager work in between."""
    sc = _sweep_small()
    with SnowContext(sc) as a, SnowContext(sc) as b:
        a.init()
        b.init()
        for _ in range(7):
            a.step()
        b.run_graph(2)
        b.step()
        b.run_graph(3)
        b.sync()
        b.run_graph(1)
        for i in range(len(sc)):
            pa, pb = a.read_particles(i), b.read_particles(i)
            for k in ("x", "v", "fe", "fp"):
                assert np.array_equal(pa[k], pb[k]), (i, k)


def test_autotune_keeps_state_and_results():
    sc = _sweep_small()
    with SnowContext(sc) as a, SnowContext(sc) as b:
        a.init()
        b.init()
        for ctx in (a, b):
            ctx.step()
        before = b.read_particles(2)
        shape = b.autotune(steps=2)
        assert shape["g2p_block"] in (64, 128, 256)
        assert shape["sort_bits"] in (8, 9, 10, 11)  # a different radix digit must not change the bits below
        after = b.read_particles(2)
        for k in ("x", "v", "fe", "fp"):
            assert np.array_equal(before[k], after[k]), k
        for _ in range(3):
            a.step()
        b.run_graph(3)
        for i in range(len(sc)):
            pa, pb = a.read_particles(i), b.read_particles(i)
            for k in ("x", "v", "fe", "fp"):
                assert np.array_equal(pa[k], pb[k]), (i, k)


def test_autotune_then_new_state_matches_untuned():
    """After autotuning (whose trials step a snapshot and rewind the step counter), a host write of
    new positions and steps with another dt give the same bits as a context that never tuned: no
    cell tile keeps a stamp from a trial step."""
    sc = _sweep_small()
    with SnowContext(sc) as a, SnowContext(sc) as b:
        a.init()
        b.init()
        for ctx in (a, b):
            ctx.step()
        b.autotune(steps=3)
        for i in range(len(sc)):
            st = a.read_particles(i)
            x2 = (st["x"] + np.float32(0.004)).astype(np.float32)
            for ctx in (a, b):
                ctx.write_particles(i, x=x2)
        for ctx in (a, b):
            for _ in range(3):
                ctx.step(7e-5)
        for i in range(len(sc)):
            pa, pb = a.read_particles(i), b.read_particles(i)
            for k in ("x", "v", "fe", "fp"):
                assert np.array_equal(pa[k], pb[k]), (i, k)


def test_batched_sweep_matches_oracle(oracle_mod):
    """Per-scene parameters and dt in one batch: each scene agrees with its own FP64 oracle."""
    sc = _sweep_small()[:3]
    with SnowContext(sc) as ctx:
        ctx.init()
        ctx.run_graph(10)
        got = [ctx.read_particles(i) for i in range(len(sc))]
    for s, g in zip(sc, got):
        o = oracle_mod.Oracle(s, nthreads=8)
        o.init()
        o.step(s.dt, 10)
        assert np.abs(g["x"] - o.x).max() < 1e-6
        assert rel_err(g["v"], o.v) < 1e-4


def test_cpp_driver_matches_python_binding(tmp_path, c1_scene):
    """The C++ step interface (include/snow_time_update.hpp, driven by snow_amd/snow_run) gives the
    same bits as the ctypes binding for init + 10 steps of the 64^3 snowball drop."""
    import os
    import subprocess

    from snow_amd import load_run_output
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "snow_amd", "snow_run")
    scene_file, out_file = str(tmp_path / "c1.scn"), str(tmp_path / "c1.out")
    c1_scene.save(scene_file)
    print(f"\n[cpp driver] {exe} {scene_file} 10", flush=True)
    res = subprocess.run([exe, scene_file, "10", out_file], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    got = load_run_output(out_file)
    with SnowContext(c1_scene) as ctx:
        ctx.init()
        for _ in range(10):
            ctx.step(c1_scene.dt)
        ref = ctx.read_particles()
    for k in ("x", "v", "fe", "fp"):
        assert np.array_equal(got[k], ref[k]), k


def test_benchstats_csv_layout(tmp_path, c1_scene):
    """Stage timers give per-stage statistics, written in the reference's benchstats.csv layout."""
    with SnowContext(c1_scene) as ctx:
        ctx.init()
        ctx.enable_stage_timers(True)
        ctx.reset_stage_times()
        for _ in range(5):
            ctx.step(c1_scene.dt)
        stats = ctx.stage_stats()
        path = str(tmp_path / "benchstats.csv")
        ctx.write_benchstats(path)
    lines = open(path).read().splitlines()
    assert lines[0] == "BENCHMARKER STATS: (all times in ns)" and lines[1] == ""
    assert lines[2] == "Name:;Max:;Min:;Avg:;Stdev:;#Calls:"
    rows = {r.split(";")[0]: r.split(";") for r in lines[3:]}
    assert "#Sim:Transfer to Grid" in rows and "#Sim:Transfer to Particles" in rows
    for name, r in rows.items():
        mx, mn, avg, sd, calls = float(r[1]), float(r[2]), float(r[3]), float(r[4]), int(r[5])
        assert calls == 5 and mx >= avg >= mn > 0 and sd >= 0, name
        assert stats[name][4] == 5


def test_positions_view_is_the_particle_state(c1_scene):
    """The zero-copy view holds (x, y, z, m) of every particle of the scene (in sorted order)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    with SnowContext(c1_scene) as ctx:
        ctx.init()
        for _ in range(3):
            ctx.step(c1_scene.dt)
        ctx.sync()
        ptr, stride, n = ctx.positions_view(0)
        assert n == len(c1_scene.particles) and stride == 32
        raw = np.empty((n, stride // 4), np.float32)
        assert hip.hipMemcpy(raw.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(n * stride), 2) == 0
        p = ctx.read_particles()
    got = raw[:, :4]
    want = np.concatenate([p["x"], c1_scene.particles.mass[:, None]], 1).astype(np.float32)
    order_g = np.lexsort(got.T[::-1])
    order_w = np.lexsort(want.T[::-1])
    assert np.array_equal(got[order_g], want[order_w])


def _storage_ids(ctx, x_creation):
    """Creation index of every particle in the device's storage order, matched by exact position
    (the zero-copy view holds x, y, z, m in storage order)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    ctx.sync()
    ptr, stride, n = ctx.positions_view(0)
    raw = np.empty((n, stride // 4), np.float32)
    assert hip.hipMemcpy(raw.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(n * stride), 2) == 0
    xs = np.ascontiguousarray(raw[:, :3]).view(np.uint32)
    xc = np.ascontiguousarray(x_creation, np.float32).view(np.uint32)
    ks = np.lexsort(xs.T[::-1])
    kc = np.lexsort(xc.T[::-1])
    assert np.array_equal(xs[ks], xc[kc])  # the same set of positions
    assert not (xc[kc][1:] == xc[kc][:-1]).all(axis=1).any()  # all distinct, so the match is unique
    ids = np.empty(n, np.int64)
    ids[ks] = kc
    return ids


@pytest.mark.parametrize("name,n", [("snowball_drop", None), ("two_snowballs", 1 << 20),
                                    ("two_snowballs", 1536000), ("two_snowballs", 4300000)])
def test_storage_order_is_stable_cell_sort(name, n):
    """A step sorts the particles by the cell key of their positions at the start of the step (cell
    tiles of 4^3, tile-major, cells inside a tile x-fastest), stably: particles of one cell keep the
    order they had. Checked on the storage order before and after one step, for every particle
    whose cell is unambiguous (u = (x - origin)/h further than 1e-4 from an integer). The sizes
    cover sort tiles of 4096 keys that fit one workgroup per CU (16 ranking waves each), more tiles
    than CUs (4 waves each), and more tiles than the column scan holds in registers."""
    sc = getattr(scenes, name)() if n is None else getattr(scenes, name)(n_particles=n)
    with SnowContext(sc) as ctx:
        ctx.init()
        for _ in range(2):
            ctx.step(sc.dt)
        x0 = ctx.read_particles()["x"]
        before = _storage_ids(ctx, x0)
        ctx.step(sc.dt)
        after = _storage_ids(ctx, ctx.read_particles()["x"])
    g = sc.grid
    origin = np.array([g.x_off, g.y_off, g.z_off], np.float64)
    u = (x0.astype(np.float64) - origin) / float(g.h)
    sure = (np.abs(u - np.rint(u)) > 1e-4).all(axis=1)
    c = np.floor(u).astype(np.int64) + 2
    ct = (np.array(g.dims) + 3 + 3) // 4
    assert (c >= 0).all() and (c < 4 * ct).all()  # every particle inside the stencil range here
    tile = (c[:, 0] >> 2) + ct[0] * ((c[:, 1] >> 2) + ct[1] * (c[:, 2] >> 2))
    key = tile * 64 + (c[:, 0] & 3) + 4 * (c[:, 1] & 3) + 16 * (c[:, 2] & 3)
    rank_before = np.empty(len(before), np.int64)
    rank_before[before] = np.arange(len(before))
    seq = after[sure[after]]  # storage order after the step, unambiguous particles only
    k, r = key[seq], rank_before[seq]
    print(f"\n[{name} {n}] {len(seq)} of {len(after)} particles checked, {len(np.unique(k))} cells", flush=True)
    assert (np.diff(k) >= 0).all()
    same = np.diff(k) == 0
    assert (np.diff(r)[same] > 0).all()
