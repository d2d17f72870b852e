"""Seeded scene generators for the benchmark workloads.

Particle sets come from a counter-based SplitMix64 hash of (seed, scene, candidate index), so
the same set is produced on any machine and independent of evaluation order. Shapes are
filled either by stratified 2x2x2 sub-cell jitter (8 particles per cell) or by rejection
sampling uniform candidates j = 0, 1, ... in the shape's bounding box and keeping the
first N accepted. Mass m_p = rho0 * V_shape / N, F_E = F_P = I.

Workloads (unit cube domain, node i at i*h):
* ``snowball_drop``      64^3, h=1/64, sphere r=0.15 at (0.5,0.6,0.5), 8 ppc, dt=1e-4
* ``two_snowballs``      128^3, h=1/128, spheres r=0.12 at x=0.3/0.7 moving at +-3 m/s, N=2^20, dt=5e-5
* ``two_snowballs_8ppc`` the same geometry at 8 particles per cell (~2.4e5 particles)
* ``slab_impact``        256^3, h=1/256, slab [0.2,0.8]x[0.5,0.7]x[0.2,0.8], box collider, sphere impactor
* ``parameter_sweep``    64 scenes, log-uniform N in [4k, 256k], grids {32,48,64,96}^3, random materials
The reference's own scene programs (src/exe/*.cpp) are available unchanged through ``reference``:
it builds them with the C++ scene builders of include/snow_scenes.hpp (the same particle sets as
the reference's rand()-based samplers) by way of the driver snow_amd/snow_run.
"""
from __future__ import annotations

import math
import os
import subprocess
import tempfile

import numpy as np

from . import (BOX, HALF_PLANE, SPHERE, Collider, CollisionObjects, Grid, Material, ParticleSystem,
               PhysicalScene)

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniforms(seed: int, scene: int, idx: np.ndarray, comp: int, ncomp: int = 4) -> np.ndarray:
    """U[0,1) doubles for candidate indices ``idx``, component ``comp``."""
    base = splitmix64(np.uint64(seed) * np.uint64(1000003) + np.uint64(scene))
    with np.errstate(over="ignore"):
        ctr = base + np.asarray(idx, np.uint64) * np.uint64(ncomp) + np.uint64(comp)
    return (splitmix64(ctr) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def snow_material(young=1.4e5, poisson=0.2, hardening=10.0, crit_comp=2.5e-2, crit_stretch=7.5e-3,
                  grad_scale=1.0) -> Material:
    return Material(young=young, poisson=poisson, hardening=hardening, crit_comp=crit_comp,
                    crit_stretch=crit_stretch, flip_alpha=0.95, gravity=(0.0, -9.81, 0.0), grad_scale=grad_scale)


def ground(y=0.05, friction=0.2) -> Collider:
    return Collider((0.0, y, 0.0), HALF_PLANE, normal=(0.0, 1.0, 0.0), friction=friction)


def stratified_sphere(seed, scene, center, radius, h, ppc_axis=2):
    """Stratified jitter: each cell of the sphere's bounding box is split into ppc_axis^3 sub-cells with
    one uniformly jittered point each; points inside the sphere are kept."""
    c = np.asarray(center, np.float64)
    lo = np.floor((c - radius) / h).astype(int)
    hi = np.ceil((c + radius) / h).astype(int)
    sub = ppc_axis
    nx, ny, nz = (hi - lo) * sub
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    idx = (i + nx * (j + ny * k)).reshape(-1).astype(np.uint64)
    pts = np.stack([i, j, k], -1).reshape(-1, 3).astype(np.float64)
    jit = np.stack([uniforms(seed, scene, idx, a) for a in range(3)], -1)
    pos = (lo * sub + pts + jit) * (h / sub)
    keep = np.sum((pos - c) ** 2, -1) < radius * radius
    return pos[keep]


def rejection_sample(seed, scene, n, lo, hi, inside, batch=1 << 22):
    """First n accepted uniform candidates j = 0, 1, ... in the box [lo, hi]."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    out, start = [], 0
    got = 0
    while got < n:
        idx = np.arange(start, start + batch, dtype=np.uint64)
        u = np.stack([uniforms(seed, scene, idx, a) for a in range(3)], -1)
        p = lo + u * (hi - lo)
        p = p[inside(p)]
        out.append(p)
        got += len(p)
        start += batch
    return np.concatenate(out)[:n]


def snowball_drop(seed=1001, grad_scale=1.0) -> PhysicalScene:
    """Plain MPM on a 64^3 grid over the unit cube: one sphere at 8 particles per cell."""
    n = 64
    h = 1.0 / n
    r, c = 0.15, (0.5, 0.6, 0.5)
    x = stratified_sphere(seed, 0, c, r, h)
    rho0 = 400.0
    m = rho0 * (4.0 / 3.0 * math.pi * r ** 3) / len(x)
    ps = ParticleSystem(x=x.astype(np.float32), mass=np.full(len(x), m, np.float32))
    return PhysicalScene(ps, Grid(n, n, n, h, 0.0, 0.0, 0.0), CollisionObjects([ground()]),
                         snow_material(grad_scale=grad_scale), dt=1e-4)


def two_snowballs(n_particles=1 << 20, seed=1002, grad_scale=1.0, dims=128) -> PhysicalScene:
    """Two spheres of radius 0.12 at x=0.3 and x=0.7 flying at each other at 3 m/s; 2^20 particles."""
    h = 1.0 / dims
    r = 0.12
    c1, c2 = np.array([0.3, 0.5, 0.5]), np.array([0.7, 0.5, 0.5])

    def inside(p):
        return (np.sum((p - c1) ** 2, -1) < r * r) | (np.sum((p - c2) ** 2, -1) < r * r)

    x = rejection_sample(seed, 0, n_particles, (0.3 - r, 0.5 - r, 0.5 - r), (0.7 + r, 0.5 + r, 0.5 + r), inside)
    v = np.zeros_like(x)
    v[:, 0] = np.where(x[:, 0] < 0.5, 3.0, -3.0)
    rho0 = 400.0
    m = rho0 * 2 * (4.0 / 3.0 * math.pi * r ** 3) / len(x)
    ps = ParticleSystem(x=x.astype(np.float32), v=v.astype(np.float32), mass=np.full(len(x), m, np.float32))
    return PhysicalScene(ps, Grid(dims, dims, dims, h, 0.0, 0.0, 0.0), CollisionObjects([ground()]),
                         snow_material(grad_scale=grad_scale), dt=5e-5)


def two_snowballs_8ppc(seed=1006, grad_scale=1.0, dims=128) -> PhysicalScene:
    """The two-snowball geometry of two_snowballs at 8 particles per cell (stratified 2x2x2 jitter in
    each cell of a sphere's bounding box; about 2.4e5 particles on 128^3)."""
    h = 1.0 / dims
    r = 0.12
    parts, vels = [], []
    for k, (c, vx) in enumerate((((0.3, 0.5, 0.5), 3.0), ((0.7, 0.5, 0.5), -3.0))):
        x = stratified_sphere(seed, k, c, r, h)
        parts.append(x)
        vels.append(np.tile([vx, 0.0, 0.0], (len(x), 1)))
    x, v = np.concatenate(parts), np.concatenate(vels)
    m = 400.0 * 2 * (4.0 / 3.0 * math.pi * r ** 3) / len(x)
    ps = ParticleSystem(x=x.astype(np.float32), v=v.astype(np.float32), mass=np.full(len(x), m, np.float32))
    return PhysicalScene(ps, Grid(dims, dims, dims, h, 0.0, 0.0, 0.0), CollisionObjects([ground()]),
                         snow_material(grad_scale=grad_scale), dt=5e-5)


def slab_impact(n_particles=1 << 24, seed=1003, grad_scale=1.0, dims=256, impactor_y=0.85) -> PhysicalScene:
    """Slab [0.2,0.8]x[0.5,0.7]x[0.2,0.8] dropped onto a box collider, hit by a sphere impactor of
    radius 0.1 (centre height impactor_y; 0.8 starts it in contact with the slab) at 10 m/s."""
    h = 1.0 / dims
    lo, hi = (0.2, 0.5, 0.2), (0.8, 0.7, 0.8)
    x = rejection_sample(seed, 0, n_particles, lo, hi, lambda p: np.ones(len(p), bool))
    rho0 = 400.0
    m = rho0 * (0.6 * 0.2 * 0.6) / len(x)
    cols = [ground(),
            Collider((0.5, 0.175, 0.5), BOX, normal=(0.2, 0.125, 0.2), friction=0.2),
            Collider((0.5, impactor_y, 0.5), SPHERE, velocity=(0.0, -10.0, 0.0), radius=0.1, friction=0.2)]
    ps = ParticleSystem(x=x.astype(np.float32), mass=np.full(len(x), m, np.float32))
    return PhysicalScene(ps, Grid(dims, dims, dims, h, 0.0, 0.0, 0.0), CollisionObjects(cols),
                         snow_material(grad_scale=grad_scale), dt=2.5e-5)


def parameter_sweep(num_scenes=64, seed=4000, nmin=4096, nmax=262144, dims_choices=(32, 48, 64, 96),
                    grad_scale=1.0) -> list:
    """Independent scenes with log-uniform particle counts, random shape, grid and material."""
    out = []
    for s in range(num_scenes):
        u = uniforms(seed, s, np.arange(16, dtype=np.uint64), 0, ncomp=1)
        n = int(round(math.exp(math.log(nmin) + u[0] * (math.log(nmax) - math.log(nmin)))))
        dims = dims_choices[min(len(dims_choices) - 1, int(u[1] * len(dims_choices)))]
        h = 1.0 / dims
        rad = 0.08 + 0.12 * u[2]
        c = np.array([0.25 + 0.5 * u[3], 0.25 + 0.5 * u[4], 0.25 + 0.5 * u[5]])
        vel = np.array([-2 + 4 * u[6], -2 + 4 * u[7], -2 + 4 * u[8]])
        sphere = u[9] < 0.5
        if sphere:
            inside = lambda p, c=c, rad=rad: np.sum((p - c) ** 2, -1) < rad * rad
            vol = 4.0 / 3.0 * math.pi * rad ** 3
        else:
            inside = lambda p: np.ones(len(p), bool)
            vol = (2 * rad) ** 3
        x = rejection_sample(seed + 1, s, n, c - rad, c + rad, inside, batch=1 << 18)
        mat = snow_material(young=4.8e4 + (1.4e5 - 4.8e4) * u[10], hardening=5 + 5 * u[11],
                            crit_comp=1.9e-2 + 0.6e-2 * u[12], crit_stretch=5e-3 + 2.5e-3 * u[13],
                            grad_scale=grad_scale)
        m = 400.0 * vol / n
        ps = ParticleSystem(x=x.astype(np.float32), v=np.tile(vel, (n, 1)).astype(np.float32),
                            mass=np.full(n, m, np.float32))
        scene = PhysicalScene(ps, Grid(dims, dims, dims, h, 0.0, 0.0, 0.0), CollisionObjects([ground()]), mat,
                              dt=1e-4 * 64 / dims)
        scene.shape = "sphere" if sphere else "box"
        out.append(scene)
    return out


def tiny_block(n_side=6, dims=16, seed=7, velocity=(0.5, -1.0, 0.25), grad_scale=1.0, jitter=0.5,
               center=(0.5, 0.5, 0.5), material=None) -> PhysicalScene:
    """A small jittered block of n_side^3 particles (test workload)."""
    h = 1.0 / dims
    idx = np.arange(n_side ** 3, dtype=np.uint64)
    k, j, i = np.meshgrid(np.arange(n_side), np.arange(n_side), np.arange(n_side), indexing="ij")
    base = np.stack([i, j, k], -1).reshape(-1, 3).astype(np.float64)
    jit = np.stack([uniforms(seed, 0, idx, a) for a in range(3)], -1)
    spacing = h * 0.5
    x = np.asarray(center) + (base - (n_side - 1) / 2.0 + jitter * (jit - 0.5)) * spacing
    v = np.tile(np.asarray(velocity, np.float64), (len(x), 1))
    m = 400.0 * spacing ** 3
    ps = ParticleSystem(x=x.astype(np.float32), v=v.astype(np.float32), mass=np.full(len(x), m, np.float32))
    return PhysicalScene(ps, Grid(dims, dims, dims, h, 0.0, 0.0, 0.0), CollisionObjects([ground()]),
                         material or snow_material(grad_scale=grad_scale), dt=1e-4)


def large_batch(num_small=512, seed=5000, nmin=1024, nmax=65536, dims_choices=(16, 24, 32, 48, 64),
                large_particles=1 << 25, large_dims=512, grad_scale=1.0) -> list:
    """Many tiny-to-medium scenes (log-uniform 1k-64k particles, grids 16^3-64^3) plus one large
    slab-impact scene (2^25 particles on 512^3, dt 1.25e-5), all in one batch."""
    small = parameter_sweep(num_scenes=num_small, seed=seed, nmin=nmin, nmax=nmax, dims_choices=dims_choices,
                            grad_scale=grad_scale)
    big = slab_impact(n_particles=large_particles, seed=seed + 7, grad_scale=grad_scale, dims=large_dims)
    big.dt = 1.25e-5
    return small + [big]



REFERENCE_SCENES = ("snowBallDrop", "snowBallSmash", "snowBallsCrash", "snowSingleParticle", "snowSquare",
                    "snowPlowTrack", "snowPlowField", "snowBunnyScene")


def reference(name: str, asset: str | None = None) -> PhysicalScene:
    """The scene of the reference's program src/exe/<name>.cpp with the grid, boundary half-planes,
    material and step size that launchSnow adds (explicit.cpp:2-67). Built by the C++ scene builders
    (include/snow_scenes.hpp) through ``snow_run --scene NAME --dump-scene FILE``. ``asset`` is the
    particle file of snowBunnyScene ("x,y,z" lines; the reference's model/bunny.voxel is not shipped)."""
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "snow_run")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built (python -m snow_amd.build)")
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, name + ".scn")
        cmd = [exe, "--scene", name] + (["--asset", asset] if asset else []) + ["--dump-scene", path]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise ValueError(res.stderr.strip() or f"snow_run failed ({res.returncode})")
        return PhysicalScene.load(path)
