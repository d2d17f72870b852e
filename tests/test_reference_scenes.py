"""The reference's scene programs (src/exe/*.cpp) built by the C++ scene builders
(include/snow_scenes.hpp, through snow_run --scene ... --dump-scene).

The particle sets are checked bit for bit against an independent restatement here: the Microsoft C
runtime's rand() (pinned by its well-known first outputs for seed 1) and the samplers' float32
arithmetic, with the double promotions of the mass rules, in numpy. Grid, material, colliders and
step size are checked against defines.hpp and launchSnow (explicit.cpp:6-67)."""
import numpy as np
import pytest

from snow_amd import BOX, HALF_PLANE, SPHERE, ParticleSystem, PhysicalScene, scenes

F = np.float32


class MsvcRand:
    """x <- 214013 x + 2531011 (mod 2^32), output bits 16..30, seed 1."""

    def __init__(self, seed=1):
        self.state = seed

    def next(self):
        self.state = (self.state * 214013 + 2531011) & 0xFFFFFFFF
        return (self.state >> 16) & 0x7FFF

    def sample(self, extent):
        return (F(self.next()) / F(32727.0)) * F(extent)


def sample_ball(rng, n, corner, width, iv, mass_rule):
    width = F(width)
    radius = width / F(2.0)
    xs, ms = [], []
    while len(xs) < n:
        r1, r2, r3 = rng.sample(width), rng.sample(width), rng.sample(width)
        r4 = rng.sample(width) if mass_rule[0] else F(0)
        d2 = (r1 - radius) * (r1 - radius) + (r2 - radius) * (r2 - radius) + (r3 - radius) * (r3 - radius)
        if not d2 < radius * radius:
            continue
        if mass_rule[0]:
            scale, other = mass_rule[1], mass_rule[2]
            m = F((((np.float64(r4) - 0.5) * np.float64(F(10.0))) + np.float64(F(1.0))) * np.float64(F(scale))) \
                if r4 > 0.5 else F(other)
        else:
            m = F(mass_rule[1])
        xs.append([F(corner[0]) + r1, F(corner[1]) + r2, F(corner[2]) + r3])
        ms.append(m)
    v = np.tile(np.asarray(iv, np.int64).astype(F) * F(1e-8), (n, 1))
    return np.asarray(xs, F), v.astype(F), np.asarray(ms, F)


@pytest.fixture(scope="module")
def built():
    return {name: scenes.reference(name) for name in scenes.REFERENCE_SCENES if name != "snowBunnyScene"}


def test_msvc_rand_known_answers():
    # the widely published first outputs of the Microsoft C runtime's rand() for seed 1
    r = MsvcRand()
    assert [r.next() for _ in range(10)] == [41, 18467, 6334, 26500, 19169, 15724, 11478, 29358, 26962, 24464]


def test_single_particle_known_answer(built):
    """snowSingleParticle.cpp: the first candidate (41, 18467, 6334) lies outside the ball, the
    second (26500, 19169, 15724) inside, so the particle is corner + those samples."""
    sc = built["snowSingleParticle"]
    s = lambda r: F(r) / F(32727.0) * F(1.0)
    d = [s(41) - F(0.5), s(18467) - F(0.5), s(6334) - F(0.5)]
    assert not (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < F(0.25))
    want = np.array([[F(5.0) + s(26500), F(1.7125) + s(19169), F(5.0) + s(15724)]], F)
    assert np.array_equal(sc.particles.x, want)
    assert sc.particles.mass[0] == F(1.25e-3) and not sc.particles.v.any()


@pytest.mark.parametrize("name", ["snowBallDrop", "snowBallSmash", "snowBallsCrash"])
def test_ball_scenes_match_independent_sampler(built, name):
    rng = MsvcRand()
    if name == "snowBallDrop":  # snowBallDrop.cpp:14-35
        parts = [sample_ball(rng, 32768, (5.0, 1.7125, 5.0), 1.0, (0, 0, 0), (True, 1.0e-4, 5e-3))]
    elif name == "snowBallSmash":  # snowBallSmash.cpp:6-21
        parts = [sample_ball(rng, 32768, (2.0, 1.525, 5.0), 1.0, (-1000000000, 0, 0), (False, 6.25e-3))]
    else:  # snowBallsCrash.cpp:6-49, one generator for both balls
        rule = (True, 8.0e-4, 10e-3)
        parts = [sample_ball(rng, 16384, (1.0, 1.7125, 5.0), 0.8, (1000000000, 0, 100000000), rule),
                 sample_ball(rng, 16384, (4.5, 1.7125, 5.5), 0.8, (-1200000000, 100000000, -500000000), rule)]
    x = np.concatenate([p[0] for p in parts])
    v = np.concatenate([p[1] for p in parts])
    m = np.concatenate([p[2] for p in parts])
    ps = built[name].particles
    assert np.array_equal(ps.x, x)
    assert np.array_equal(ps.v, v)
    assert np.array_equal(ps.mass, m)
    eye = np.eye(3, dtype=F).reshape(-1)
    assert (ps.fe == eye).all() and (ps.fp == eye).all()


def test_box_and_lattice_scenes(built):
    # snowPlowTrack.cpp:11-25: uniform in [0.8125, 2.8125) x [y0, y0 + 0.5) x [2.4, 3.2), at rest
    rng = MsvcRand()
    y0 = F(0.5125 + 4 * 0.05)
    want = np.array([[F(0.8125) + rng.sample(2.0), y0 + rng.sample(0.5), F(2.4) + rng.sample(0.8)]
                     for _ in range(32 * 32 * 64)], F)
    ps = built["snowPlowTrack"].particles
    assert np.array_equal(ps.x, want)
    assert (ps.mass == F(1.25e-3)).all() and not ps.v.any()
    # snowPlowField.cpp:11-24: two particles in a box of extents (0.2, 0.1, 0.2) x grid size
    rng = MsvcRand()
    ext = (F(201 * 0.05 * 0.2), F(201 * 0.05 * 0.1), F(201 * 0.05 * 0.2))
    want = np.array([[F(0.8125) + rng.sample(ext[0]), y0 + rng.sample(ext[1]), F(2.4) + rng.sample(ext[2])]
                     for _ in range(2)], F)
    assert np.array_equal(built["snowPlowField"].particles.x, want)
    # snowSquare.cpp:5-25: 32^3 lattice of pitch 0.025, y slowest, x fastest
    y, z, x = np.meshgrid(np.arange(32), np.arange(32), np.arange(32), indexing="ij")
    want = np.stack([F(0.9625) + x.astype(F) * F(0.025), F(0.72) + y.astype(F) * F(0.025),
                     F(0.9625) + z.astype(F) * F(0.025)], -1).reshape(-1, 3)
    assert np.array_equal(built["snowSquare"].particles.x, want)


def test_launch_snow_defaults(built):
    """Grid, material and dt of defines.hpp; the six boundary half-planes of explicit.cpp:9-67 after
    the scene's own sphere colliders, in order."""
    lo, hi = F(0.5125 + 4 * 0.05), (F(0.5125 + 197 * 0.05), F(0.5125 + 197 * 0.05), F(1.5125 + 197 * 0.05))
    planes = [((lo, lo, F(1.5125 + 4 * 0.05)), n) for n in ((1, 0, 0), (0, 1, 0), (0, 0, 1))] + \
             [(hi, n) for n in ((-1, 0, 0), (0, -1, 0), (0, 0, -1))]
    spheres = {"snowPlowTrack": [((0.2125, 1.1125, 2.8), (5.0, 0.0, 0.0), 0.3)],
               "snowPlowField": [((0.4125, 1.0125, 2.8), (10.0, 0.0, 0.0), 0.2),
                                 ((4.0125, 1.0125, 1.2), (-10.0, 0.0, 0.0), 0.2),
                                 ((1.2, 1.0125, 0.4125), (0.0, 0.0, 10.0), 0.2)]}
    for name, sc in built.items():
        g, mt = sc.grid, sc.material
        assert g.dims == (201, 201, 201) and g.h == F(0.05)
        assert (g.x_off, g.y_off, g.z_off) == (F(0.5125), F(0.5125), F(1.5125))
        assert (mt.young, mt.poisson, mt.hardening) == (F(6e6), F(0.2), F(30.0))
        assert (mt.crit_comp, mt.crit_stretch, mt.flip_alpha) == (F(1e-2), F(1e-3), F(0.95))
        assert mt.gravity == (0.0, F(-9.81), 0.0) and mt.grad_scale == 1.0
        assert sc.dt == F(1e-3)
        cols = sc.colliders.colliders
        own = spheres.get(name, [])
        assert len(cols) == len(own) + 6
        for c, (pos, vel, r) in zip(cols, own):
            assert c.type == SPHERE and c.position == tuple(map(F, pos)) and c.velocity == tuple(map(F, vel))
            assert c.radius == F(r) and c.friction == F(0.2)
        for c, (pos, n) in zip(cols[len(own):], planes):
            assert c.type == HALF_PLANE and c.position == tuple(map(F, pos)) and c.normal == tuple(map(F, n))
            assert c.friction == F(0.2) and c.velocity == (0.0, 0.0, 0.0)
        assert BOX not in [c.type for c in cols]


def test_bunny_scene_reads_particle_file(tmp_path):
    """snowBunnyScene.cpp:5-14 through initParticlesFromFile (particlesystem.cpp:19-38): "x,y,z"
    lines placed at scale * p + (5, 0, 5), scale 12.5 * 1.259, mass 6.25e-3, at rest."""
    pts = np.array([[0.1, 0.2, 0.3], [-0.05, 0.5, 0.25], [0.0, 0.0, 0.0]], F)
    path = tmp_path / "bunny.voxel"
    path.write_text("".join(f"{a!r},{b!r},{c!r}\n" for a, b, c in pts.tolist()))
    sc = scenes.reference("snowBunnyScene", asset=str(path))
    scale = F(12.5) * F(1.259)
    want = pts * scale + np.array([5.0, 0.0, 5.0], F)
    assert np.array_equal(sc.particles.x, want)
    assert (sc.particles.mass == F(6.25e-3)).all() and not sc.particles.v.any()
    # the Python loader of the same file format agrees
    ps = ParticleSystem.from_file(str(path), pos=(5.0, 0.0, 5.0), scale=(scale,) * 3)
    assert np.array_equal(ps.x, want)


def test_scene_errors():
    with pytest.raises(ValueError, match="unknown scene"):
        scenes.reference("snowNoSuchScene")
    with pytest.raises(ValueError, match="particle file"):
        scenes.reference("snowBunnyScene")
    with pytest.raises(ValueError, match="cannot open"):
        scenes.reference("snowBunnyScene", asset="/nonexistent/bunny.voxel")


def test_scene_file_round_trip(built, tmp_path):
    sc = built["snowBallsCrash"]
    sc.save(str(tmp_path / "s.scn"))
    back = PhysicalScene.load(str(tmp_path / "s.scn"))
    for k in ("x", "v", "mass", "fe", "fp"):
        assert np.array_equal(getattr(back.particles, k), getattr(sc.particles, k))
    assert back.colliders == sc.colliders and back.grid == sc.grid and back.material == sc.material
    assert back.dt == sc.dt
