"""The benchmark configurations on the GPU, checked against the FP64 CPU oracle.

* two snowballs, 2^20 particles on 128^3, dt 5e-5: 100 free-running steps compared with the oracle
  after every step (positions within 1e-4 of the domain size at step 100), and the configuration's
  1000 steps by properties;
* the slab impact on 256^3 (box collider, sphere impactor, physical weight gradient), plain and
  mixed precision: 2^20 particles against the oracle, and the full 2^24 particles by properties;
* the 64-scene parameter sweep at full size (graph steps, autotuned launch shapes): four scenes
  against the oracle and against their own single-scene contexts, bit for bit;
* the 513-scene batch (512 small scenes plus one 2^25-particle 512^3 scene in one graph): finite
  state, per-scene mass on the grid, small scenes against the oracle, the large scene against its
  own context bit for bit.

Oracle threads: at most 16 (the host share of one GPU).
"""
import os
import time

import numpy as np
import pytest

from snow_amd import FLAG_KEEP_GRID, FLAG_MIXED, SnowContext, scenes

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def errors(p, g, o):
    """Max errors of one GPU state against the oracle: x, F_E, F_P absolute (per-particle vector /
    Frobenius norm), v and the grid fields relative to the largest oracle value."""
    def norm_max(a, b):
        return float(np.sqrt(((np.asarray(a, np.float64) - b) ** 2).sum(-1)).max())
    out = {"x": norm_max(p["x"], o.x), "v": rel_err(p["v"], o.v), "fe": norm_max(p["fe"], o.fe),
           "fp": norm_max(p["fp"], o.fp)}
    if g is not None:
        out["grid_m"] = rel_err(g["m"], o.gm)
        out["grid_mv"] = rel_err(g["mv"], o.gmv)
    return out


def rms(p, o, k):
    return float(np.sqrt(((np.asarray(p[k], np.float64) - getattr(o, k)) ** 2).sum(-1).mean()))


def test_two_snowballs_1m_100_steps(oracle_mod):
    """The 2^20-particle headline scene stepped 100 times, compared with the FP64 oracle after every
    step (explicittimeupdate.cpp:36-63 per step): GPU and oracle advance in lockstep and the GPU state
    is read back after each step."""
    sc = scenes.two_snowballs()
    steps = 100
    h = sc.grid.h
    # Bounds at steps 1, 10 and 100: x (domain size 1), v (relative), F_E, F_P (Frobenius); between
    # them the bound of the next checkpoint applies. The node fields follow the particles: a position
    # error dx moves a node's share of a particle's mass by about |dN/du| dx / h, so the grid bounds
    # (relative to the largest node value) are 100 x_err / h.
    bounds = {1: dict(x=1e-6, v=1e-4, fe=1e-4, fp=1e-4), 10: dict(x=1e-5, v=1e-3, fe=1e-3, fp=1e-3),
              100: dict(x=1e-4, v=2e-3, fe=2e-3, fp=2e-3)}
    log(f"two snowballs: {len(sc.particles)} particles, 128^3, {steps} steps in lockstep with the oracle")
    o = oracle_mod.Oracle(sc, nthreads=THREADS)
    o.init()
    worst = {}
    with SnowContext(sc, FLAG_KEEP_GRID) as ctx:
        ctx.init()
        for s in range(1, steps + 1):
            ctx.step(sc.dt)
            o.step(sc.dt, 1)
            e = errors(ctx.read_particles(), ctx.read_grid(fields=("m", "mv")), o)
            b = dict(bounds[next(c for c in (1, 10, 100) if s <= c)])
            b["grid_m"] = b["grid_mv"] = 100 * e["x"] / h + 1e-6
            for k, lim in b.items():
                assert e[k] < lim, (s, k, e[k], lim)
                worst[k] = max(worst.get(k, 0.0), e[k])
            if s in (1, 10, 100):
                log(f"step {s}: " + " ".join(f"{k}={v:.3e}" for k, v in e.items()))
        ctx.check_finite()
    log("largest over the 100 steps: " + " ".join(f"{k}={v:.3e}" for k, v in worst.items()))
    assert e["x"] < 1e-4  # positions within 1e-4 of the domain size after 100 steps


def test_two_snowballs_1m_1000_steps_properties():
    """The configuration's 1000 steps (the balls meet and spread on the ground): finite state, every
    particle inside the grid, the particles' mass all on the grid (P2G conserves it), and the mean
    velocity only changed by gravity and the ground."""
    sc = scenes.two_snowballs()
    m = np.asarray(sc.particles.mass, np.float64)
    with SnowContext(sc, FLAG_KEEP_GRID) as ctx:
        ctx.init()
        for s in range(1000):
            ctx.step(sc.dt)
        ctx.check_finite()
        p = ctx.read_particles()
        g = ctx.read_grid(fields=("m",))
    x = p["x"]
    assert np.all((x > 0.0) & (x < 1.0))
    assert np.all(x[:, 1] > 0.05 - 2 * sc.grid.h)  # on or above the ground plane
    assert abs(g["m"].sum() - m.sum()) <= 1e-5 * m.sum()
    mom = (m[:, None] * p["v"]).sum(0) / m.sum()
    log(f"1000 steps: mean velocity {mom}, grid mass / particle mass - 1 = {g['m'].sum() / m.sum() - 1:.2e}")
    assert abs(mom[0]) < 0.05 and abs(mom[2]) < 0.05  # the two balls' +-3 m/s cancel along x


def _slab_forms(oracle_mod, steps):
    sc = scenes.slab_impact(n_particles=1 << 20, grad_scale=0.0, impactor_y=0.8)
    log(f"slab impact: {len(sc.particles)} particles, 256^3, {steps} steps, oracle")
    o = oracle_mod.Oracle(sc, nthreads=THREADS)
    o.init()
    o.step(sc.dt, steps)
    out = {}
    for name, flags in (("plain", 0), ("mixed", FLAG_MIXED)):
        with SnowContext(sc, flags | FLAG_KEEP_GRID) as ctx:
            ctx.init()
            for _ in range(steps):
                ctx.step(sc.dt)
            ctx.check_finite()
            out[name] = (ctx.read_particles(), ctx.read_grid(fields=("m", "mv")),
                         ctx.fp64_particles() if flags else 0)
    return sc, o, out


def test_slab_impact_256_plain_and_mixed(oracle_mod):
    """The mixed-precision stress scene on its 256^3 grid with 2^20 particles: the sphere impactor
    (10 m/s) starts in contact with the slab, which rests above the box collider; weight gradients
    with 1/h so that particles reach the plastic clamp (particleAdvance.glsl:112-117). 20 steps,
    both forms against the FP64 oracle. The mixed form takes particles near the clamp through the
    FP64 path and its error stays at the plain form's: 1-ulp differences are amplified at the
    clamp and friction switches in either form, so the comparison allows 25 % on the maxima and
    5 % on the RMS."""
    steps = 20
    sc, o, out = _slab_forms(oracle_mod, steps)
    mat = sc.material
    sv = np.linalg.svd(o.fe.reshape(-1, 3, 3), compute_uv=False)
    clamped = ((np.abs(sv - (1 - mat.crit_comp)) < 1e-6) | (np.abs(sv - (1 + mat.crit_stretch)) < 1e-6)).any(-1)
    log(f"clamped fraction {clamped.mean():.2e}, FP64 path in the last step: {out['mixed'][2]} particles")
    assert clamped.sum() > 0
    assert out["mixed"][2] > 0
    e = {k: errors(out[k][0], out[k][1], o) for k in out}
    r = {k: {q: rms(out[k][0], o, q) for q in ("x", "v", "fe", "fp")} for k in out}
    for k in out:
        log(f"{k}: " + " ".join(f"{q}={v:.3e}" for q, v in e[k].items()) + " | rms " +
            " ".join(f"{q}={v:.3e}" for q, v in r[k].items()))
        assert e[k]["x"] < 1e-4
        assert e[k]["grid_m"] < 1e-4
    for q in ("x", "v", "fe", "fp"):
        assert e["mixed"][q] <= 1.25 * e["plain"][q] + 1e-7, q
        assert r["mixed"][q] <= 1.05 * r["plain"][q] + 1e-9, q


def test_slab_impact_full_size_properties():
    """The full 2^24-particle slab on 256^3, plain and mixed, 10 steps: finite state, the P2G puts
    the particle mass on the grid, the FP64 path is taken, and the two forms stay together."""
    sc = scenes.slab_impact(grad_scale=0.0, impactor_y=0.8)
    n = len(sc.particles)
    assert n == 1 << 24
    mass = float(sc.particles.mass.astype(np.float64).sum())
    xs = {}
    for name, flags in (("plain", 0), ("mixed", FLAG_MIXED)):
        log(f"slab impact full size, {name}")
        with SnowContext(sc, flags) as ctx:
            ctx.init()
            ctx.step(sc.dt)
            m = ctx.read_grid(fields=("m",))["m"]
            assert m.astype(np.float64).sum() == pytest.approx(mass, rel=1e-5)
            for _ in range(9):
                ctx.step(sc.dt)
            ctx.check_finite()
            if flags:
                assert ctx.fp64_particles() > 0
            xs[name] = ctx.read_particles()["x"]
    d = np.abs(xs["plain"] - xs["mixed"]).max()
    log(f"max |x_plain - x_mixed| = {d:.3e}")
    assert d < 1e-5


def test_parameter_sweep_64_scenes(oracle_mod):
    """All 64 scenes (4k-256k particles, grids 32^3-96^3, per-scene material and dt) in one
    context: autotuned launch shapes, 10 graph-replayed steps. Four scenes (the largest, a 96^3 one,
    a box and a sphere) agree with the FP64 oracle and bit for bit with their own contexts."""
    sc = scenes.parameter_sweep()
    n = [len(s.particles) for s in sc]
    big = int(np.argmax(n))
    g96 = next(i for i, s in enumerate(sc) if s.grid.dimx == 96 and i != big)
    box = next(i for i, s in enumerate(sc) if i not in (big, g96) and s.shape == "box")
    sph = next(i for i, s in enumerate(sc) if i not in (big, g96, box) and s.shape == "sphere")
    pick = [big, g96, box, sph]
    log(f"sweep: {sum(n)} particles in 64 scenes; checking scenes {pick} "
        f"({[n[i] for i in pick]} particles, grids {[sc[i].grid.dimx for i in pick]})")
    steps = 10
    with SnowContext(sc) as ctx:
        ctx.init()
        shape = ctx.autotune(steps=2)
        log(f"autotuned launch shape {shape}")
        ctx.run_graph(steps)
        ctx.check_finite()
        batch = {i: ctx.read_particles(i) for i in pick}
    for i in pick:
        with SnowContext(sc[i]) as ctx:
            ctx.init()
            for _ in range(steps):
                ctx.step(sc[i].dt)
            alone = ctx.read_particles()
        for k in ("x", "v", "fe", "fp"):
            assert np.array_equal(batch[i][k], alone[k]), (i, k)
        o = oracle_mod.Oracle(sc[i], nthreads=THREADS)
        o.init()
        o.step(sc[i].dt, steps)
        e = errors(batch[i], None, o)
        log(f"scene {i}: " + " ".join(f"{k}={v:.3e}" for k, v in e.items()))
        assert e["x"] < 1e-6
        assert e["v"] < 1e-4
        assert e["fe"] < 1e-3 and e["fp"] < 1e-3


def deposited_mass(x, mass, grid):
    """Mass the P2G puts on the grid from particles at x: each particle's mass times the cubic
    B-spline weights of its in-grid stencil nodes (out-of-grid nodes are dropped, p2gTransfer.glsl:87),
    which factor per axis."""
    u = (np.asarray(x, np.float64) - np.array([grid.x_off, grid.y_off, grid.z_off])) / grid.h
    base = np.floor(u)
    f = u - base
    w = np.stack([(1 - f) ** 3 / 6, (3 * f ** 3 - 6 * f ** 2 + 4) / 6, (-3 * f ** 3 + 3 * f ** 2 + 3 * f + 1) / 6,
                  f ** 3 / 6], -1)  # [N, 3 axes, 4 offsets], nodes base - 1 + o
    node = base[..., None] + np.arange(-1, 3)
    dims = np.array(grid.dims)[None, :, None]
    inside = (node >= 0) & (node < dims)
    share = (w * inside).sum(-1).prod(-1)
    return float((np.asarray(mass, np.float64) * share).sum())


def test_large_batch_513_scenes(oracle_mod):
    """512 small scenes (1k-64k particles, 16^3-64^3) and one 2^25-particle 512^3 slab in one step
    graph, 3 steps: finite state; every small scene's grid holds the mass its particles' stencils
    deposit inside the grid; three small scenes against the FP64 oracle; the large scene bit for
    bit against its own context."""
    sc = scenes.large_batch()
    assert len(sc) == 513 and len(sc[-1].particles) == 1 << 25
    log(f"large batch: {sum(len(s.particles) for s in sc)} particles in {len(sc)} scenes")
    steps = 3
    sample = [0, 100, 511]
    with SnowContext(sc) as ctx:
        ctx.init()
        ctx.run_graph(steps - 1)
        x_before = [ctx.read_particles(i)["x"] for i in range(512)]
        ctx.run_graph(1)
        ctx.check_finite()
        log("large batch stepped")
        for i in range(512):
            m = ctx.read_grid(i, fields=("m",))["m"].astype(np.float64).sum()
            want = deposited_mass(x_before[i], sc[i].particles.mass, sc[i].grid)
            assert m == pytest.approx(want, rel=1e-5), i
        small = {i: ctx.read_particles(i) for i in sample}
        big_batch = ctx.read_particles(512)
    log("per-scene mass checked")
    for i in sample:
        o = oracle_mod.Oracle(sc[i], nthreads=THREADS)
        o.init()
        o.step(sc[i].dt, steps)
        e = errors(small[i], None, o)
        log(f"small scene {i} ({len(sc[i].particles)} particles): " + " ".join(f"{k}={v:.3e}" for k, v in e.items()))
        assert e["x"] < 1e-6 and e["v"] < 1e-4
    with SnowContext(sc[-1]) as ctx:
        ctx.init()
        ctx.run_graph(steps)
        alone = ctx.read_particles()
    for k in ("x", "v", "fe", "fp"):
        assert np.array_equal(big_batch[k], alone[k]), k
    log("large scene alone = in batch")


def dense_and_sparse(seed=1010, dims=128):
    """A 27-particle-per-cell ball (tiles of ~1700 particles, longer than the reach of a chunk bound,
    so they are split between chunks) next to 20,000 particles scattered over the whole grid (about
    one per tile: runs of one or two particles). The chunk plan then has to lower its run cost to
    keep every window within the P2G's run lists."""
    h = 1.0 / dims
    dense = scenes.stratified_sphere(seed, 0, (0.35, 0.4, 0.5), 0.06, h, ppc_axis=3)
    sparse = scenes.rejection_sample(seed, 1, 20000, (0.05, 0.1, 0.05), (0.95, 0.95, 0.95), lambda p: np.ones(len(p), bool))
    x = np.concatenate([dense, sparse]).astype(np.float32)
    v = np.zeros_like(x)
    v[: len(dense), 0] = 2.0
    m = np.full(len(x), 400.0 * h ** 3 / 27.0, np.float32)
    ps = scenes.ParticleSystem(x=x, v=v, mass=m)
    return scenes.PhysicalScene(ps, scenes.Grid(dims, dims, dims, h, 0.0, 0.0, 0.0),
                                scenes.CollisionObjects([scenes.ground()]), scenes.snow_material(), dt=5e-5)


def test_chunk_plan_dense_and_sparse(oracle_mod):
    """Planned P2G chunks on a scene of very uneven tile occupancy: 30 graph steps equal 30 eager
    steps bit for bit, the grid holds exactly the mass the particles deposit (no particle lost to a
    chunk window), and 5 steps agree with the FP64 oracle."""
    sc = dense_and_sparse()
    log(f"dense and sparse: {len(sc.particles)} particles")
    with SnowContext(sc) as a, SnowContext(sc) as b:
        for ctx in (a, b):
            ctx.init()
        for _ in range(5):
            a.step(sc.dt)
        p5 = a.read_particles()
        for _ in range(24):
            a.step(sc.dt)
        b.run_graph(29)
        x_before = a.read_particles()["x"]
        a.step(sc.dt)
        b.run_graph(1)
        a.check_finite()
        g = a.read_grid(fields=("m",))["m"].astype(np.float64).sum()
        pa, pb = a.read_particles(), b.read_particles()
    for k in ("x", "v", "fe", "fp"):
        assert np.array_equal(pa[k], pb[k]), k
    want = deposited_mass(x_before, sc.particles.mass, sc.grid)
    log(f"grid mass {g:.9e}, deposited {want:.9e}")
    assert g == pytest.approx(want, rel=1e-5)
    o = oracle_mod.Oracle(sc, nthreads=THREADS)
    o.init()
    o.step(sc.dt, 5)
    e = errors(p5, None, o)
    log("5 steps: " + " ".join(f"{k}={v:.3e}" for k, v in e.items()))
    # the isolated particles carry the largest velocity error: the FP32 build of the oracle itself
    # is 9.7e-5 (relative to the largest speed) from the FP64 one on this scene after 5 steps
    assert e["x"] < 1e-6 and e["v"] < 5e-4
