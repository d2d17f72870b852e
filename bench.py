"""Benchmark of the MI355X snow MPM timestep.

Headline workload: two snowballs (2^20 particles, radius 0.12, +-3 m/s) on a 128^3 grid over the
unit cube, dt = 5e-5, ground plane. One "step" is one full explicit MPM step of all particles
(sort, P2G, grid update, G2P + advance) with the state resident in HBM. `value` is the plain
FP32 form's particle-steps/s. The same line also reports
  * forms.mixed    - FP32 with the FP64 SVD fallback, same workload,
  * errors         - max/RMS error of x, v, F_E, F_P and grid m, mv against the FP64 CPU oracle
                     after 100 steps of the same scene, for both forms,
  * forms.two_snowballs_8ppc - the same geometry at 8 particles per cell (~2.4e5 particles),
  * forms.batched  - the 64-scene parameter sweep as one batch (captured step graph, autotuned
                     launch shapes) against the same scenes run with the single-scene path, one
                     scene after another and all interleaved on their own streams,
  * forms.stress_slab - the 2^24-particle slab impact on 256^3 (plastic clamp active), plain and
                     mixed throughput, and both forms' errors against the FP64 oracle on a 2^20-
                     particle version on the same grid,
  * forms.semi_implicit - the semi-implicit step (beta = 1) on the same scene in contact at 4x the dt:
                     ms per step and per solver iteration, and its error against the FP64 oracle on the
                     64^3 snowball,
  * forms.large_batch - (--large) 512 small scenes plus one 2^25-particle 512^3 scene in one graph,
  * cpu_baseline   - the FP32 CPU restatement of the step on the host cores this process may use
                     (median of 3 runs), with the 64^3 scene and the sweep scene by scene beside it.
Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-errors] [--no-sweep]
                    [--no-second-form] [--no-stress] [--large] [--no-autotune]

Before its W warm-up steps the headline context runs snow_autotune (launch shapes that leave the
results unchanged, timed on a snapshot it restores); config.launch_shape reports its choice.

Kernel times come from HIP events around the reference's stages. The stage named k_p2g_chunk
("Transfer to Grid") spans k_p2g_chunk and k_fold_partials; `launches` lists what each measured
stage spans, and roofline.valu_floor counts the instructions of all of them.
roofline.traffic is the HBM traffic per step of the dominant stage's launches measured with
rocprofv3 FETCH_SIZE / WRITE_SIZE counters (calibrated against a plain copy), read from the
committed summary profiles/snow_step_two_snowballs_1M_128_hbm_traffic.md; null when the summary
has no entry for one of them.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s @1M particles: plain vs mixed/batched at equal error"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
TIMER_EVERY = 8        # stage timers on every 8th timed step

# Algorithmic bytes (FP32 SoA state): P2G reads x, v, F_E, F_P, m, V0 (104 B/particle) and writes
# m, mv, f (28 B/active node); the grid update reads 28 B and writes v* (12 B) per active node;
# G2P + advance reads x, v, F_E, F_P (96 B) and writes them (96 B) per particle, and reads v*, v_old
# (24 B) per active node. Sort/permutation and dense sweeps are not algorithmic.
BYTES = {
    "#Sim:Transfer to Grid": (104, 28),
    "#Sim:Compute Derivatives": (0, 40),
    "#Sim:Transfer to Particles": (192, 24),
}
KERNEL_NAMES = {
    "#Sim:Transfer to Grid": "k_p2g_chunk",
    "#Sim:Compute Derivatives": "k_grid",
    "#Sim:Transfer to Particles": "k_g2p_advance",
}
# The launches each stage timer spans: the P2G stage ("Transfer to Grid") also adds the
# continuation partials of tiles split between chunks (k_fold_partials, a few us).
STAGE_LAUNCHES = {
    "k_p2g_chunk": ("k_p2g_chunk", "k_fold_partials"),
    "k_grid": ("k_grid",),
    "k_g2p_advance": ("k_g2p_advance",),
}


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "snow_step_two_snowballs_1M_128_hbm_traffic.md")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload from the committed rocprofv3 counter summary
    (lines "traffic_bytes_per_launch[<kernel>]: <bytes>"), or None."""
    import re
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                m = re.match(r"\s*traffic_bytes_per_launch\[(\w+)\]:\s*(\d+)", line)
                if m and m.group(1) == kernel:
                    return int(m.group(2))
    except OSError:
        pass
    return None


def stage_traffic(kernel: str):
    """HBM bytes per step of the launches the stage timer of `kernel` spans (STAGE_LAUNCHES), or None
    when the summary lacks one of them."""
    parts = [measured_traffic(k) for k in STAGE_LAUNCHES.get(kernel, (kernel,))]
    return None if any(p is None for p in parts) else sum(parts)


INSTR_SUMMARY = os.path.join(ROOT, "profiles", "snow_step_two_snowballs_1M_128_instruction_counters.md")
PLAIN_STEP_KERNEL = {"k_g2p_advance": "snow::k_g2p_advance<false, false>", "k_p2g_chunk": "snow::k_p2g_chunk",
                     "k_fold_partials": "snow::k_fold_partials",
                     "k_grid": "snow::k_grid<false, true>"}  # the instantiations a plain FP32 step launches
VALU_CYCLES = 4          # cycles per wave64 FP32 VALU instruction (16 lanes per SIMD)
SIMDS = 256 * 4          # MI355X: 256 CUs of 4 SIMDs
CLOCK_HZ = 2.4e9         # peak engine clock


def valu_floor(kernel: str, measured_us: float):
    """The vector-ALU floor of the stage measured as `kernel` on this workload: the SQ_INSTS_VALU per
    step of the launches it spans (STAGE_LAUNCHES) from the committed rocprofv3 counter summary, at
    VALU_CYCLES per instruction on every SIMD at CLOCK_HZ, against the measured time. None when the
    summary has no row for one of them."""
    rows = {}
    try:
        with open(INSTR_SUMMARY) as f:
            col = None  # the INSTS_VALU column while inside the table that has one
            for line in f:
                if not line.startswith("|"):
                    col = None
                    continue
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if "INSTS_VALU" in cells:
                    col = cells.index("INSTS_VALU")
                elif col is not None and len(cells) > col and not cells[0].startswith("-"):
                    rows[cells[0]] = cells[col]
        names = [PLAIN_STEP_KERNEL.get(k) for k in STAGE_LAUNCHES.get(kernel, (kernel,))]
        if any(nm not in rows for nm in names):
            return None
        valu = sum(int(rows[nm]) for nm in names)
    except (OSError, ValueError):
        return None
    floor_us = valu * VALU_CYCLES / (SIMDS * CLOCK_HZ) * 1e6
    return {"instructions": valu, "floor_us": floor_us, "measured_us": measured_us, "frac": floor_us / measured_us}


def host_cpu() -> dict:
    """The host's CPU model (/proc/cpuinfo), its logical CPU count (nproc) and the CPUs this process
    may run on (its affinity set, which a container or job share narrows)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def time_cpu_steps(scene, threads: int, budget_s: float, repeats: int = 3, precision: int = 32) -> dict:
    """Median particle-steps/s of `repeats` timed runs of the CPU restatement of the step on one
    scene; each run takes about budget_s / repeats (the step count is sized from one warm-up step)."""
    import statistics
    from oracle import pyoracle
    o = pyoracle.Oracle(scene, precision=precision, nthreads=threads)
    o.init()
    t0 = time.perf_counter()
    o.step(scene.dt, 1)
    t_step = time.perf_counter() - t0
    steps = max(1, int(budget_s / repeats / max(t_step, 1e-6)))
    n = len(scene.particles)
    rates = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        o.step(scene.dt, steps)
        rates.append(n * steps / (time.perf_counter() - t0))
    return {"value": statistics.median(rates), "rates": rates, "steps_per_repeat": steps, "particles": n,
            "threads": threads}


def cpu_baseline(scene, budget_s: float = 12.0) -> dict:
    """The step restated on the CPU (oracle/mpm_oracle.hpp compiled with Real=float, std::thread
    workers, deterministic z-slab P2G) on the host cores this process may use: the headline scene
    (the value), the 64^3 snowball drop, and the 64-scene sweep one scene after another; plus the
    FP64 oracle on one thread. Each figure is the median of 3 repeats."""
    import statistics
    from oracle import pyoracle
    from snow_amd import scenes
    cpu = host_cpu()
    threads = cpu["usable_cpus"]
    head = time_cpu_steps(scene, threads, budget_s)
    legs = {"C1_snowball_drop_64^3": time_cpu_steps(scenes.snowball_drop(), threads, budget_s / 4)}
    # the 64 sweep scenes one after another, one step each per repeat
    sweep = scenes.parameter_sweep()
    ors = []
    for sc in sweep:
        o = pyoracle.Oracle(sc, precision=32, nthreads=threads)
        o.init()
        o.step(sc.dt, 1)
        ors.append((o, sc))
    total = sum(len(sc.particles) for sc in sweep)
    rates = []
    for _ in range(3):
        t0 = time.perf_counter()
        for o, sc in ors:
            o.step(sc.dt, 1)
        rates.append(total / (time.perf_counter() - t0))
    legs["C4_sweep_scene_by_scene"] = {"value": statistics.median(rates), "rates": rates, "steps_per_repeat": 1,
                                       "particles": total, "scenes": len(sweep), "threads": threads}
    del ors
    legs["C1_fp64_oracle_1_thread"] = time_cpu_steps(scenes.snowball_drop(), 1, 3.0, precision=64)
    n = len(scene.particles)
    return {"value": head["value"], "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"two snowballs ({n} particles, 128^3), FP32 restatement of the reference step "
                      f"(oracle/mpm_oracle.hpp, Real=float), {threads} std::threads, median of 3 runs of "
                      f"{head['steps_per_repeat']} steps",
            "rates": head["rates"], **cpu, "legs": legs}


def timed_steps(ctx, steps: int, dt=None, graph: bool = False) -> float:
    ctx.sync()
    t0 = time.perf_counter()
    if graph:
        ctx.run_graph(steps)
    else:
        for _ in range(steps):
            ctx.step(dt)
    ctx.sync()
    return time.perf_counter() - t0


def error_reports(scene, forms: dict, steps: int) -> dict:
    """Max / RMS error of every quantity against the FP64 oracle after `steps` steps, for each
    {name: flags} form; also the fraction of particles whose F_E ends at the singular-value clamp
    (oracle) and, for forms with FLAG_MIXED, the fraction that took the FP64 path in the last step."""
    import numpy as np
    from oracle import pyoracle
    from snow_amd import FLAG_KEEP_GRID, FLAG_MIXED, SnowContext
    o = pyoracle.Oracle(scene, nthreads=min(16, host_cpu()["usable_cpus"]))
    o.init()
    o.step(scene.dt, steps)
    mat = scene.material
    # a clamped F_E is W S' V^T with V rounded to 1e-5: its singular values lie within a few 1e-5 of
    # the bound
    sv = np.linalg.svd(o.fe.reshape(-1, 3, 3), compute_uv=False)
    at_clamp = (np.abs(sv - (1 - mat.crit_comp)) < 5e-5) | (np.abs(sv - (1 + mat.crit_stretch)) < 5e-5)
    out = {"steps": steps, "particles": len(scene.particles), "reference": "FP64 CPU oracle",
           "clamped_fraction": float(at_clamp.any(-1).mean())}
    for form, flags in forms.items():
        with SnowContext(scene, flags | FLAG_KEEP_GRID) as ctx:
            ctx.init()
            for _ in range(steps):
                ctx.step(scene.dt)
            p = ctx.read_particles()
            g = ctx.read_grid(fields=("m", "mv"))
            fp64 = ctx.fp64_particles() if flags & FLAG_MIXED else None
        rep = {}
        for name, a, b in (("x", p["x"], o.x), ("v", p["v"], o.v), ("F_E", p["fe"], o.fe), ("F_P", p["fp"], o.fp),
                           ("grid_m", g["m"], o.gm), ("grid_mv", g["mv"], o.gmv)):
            d = np.asarray(a, np.float64) - b
            if d.ndim > 1:
                d = np.sqrt((d ** 2).sum(-1))  # vector norm, Frobenius norm for F_E / F_P
            rep[name] = {"max": float(np.abs(d).max()), "rms": float(np.sqrt((d ** 2).mean()))}
        if fp64 is not None:
            rep["fp64_fraction_last_step"] = fp64 / len(scene.particles)
        out[form] = rep
    return out


def batched_sweep(steps: int) -> dict:
    """64 scenes in one batch (graph + autotune) vs the same scenes run with the single-scene path:
    one context per scene, (a) one scene after another for all steps and (b) all contexts stepped
    together, each on its own stream."""
    from snow_amd import SnowContext, scenes
    sc = scenes.parameter_sweep()
    n = sum(len(s.particles) for s in sc)
    with SnowContext(sc) as ctx:
        ctx.init()
        shape = ctx.autotune(steps=3)
        ctx.run_graph(5)
        el = timed_steps(ctx, steps, graph=True)
    batched = n * steps / el
    ctxs = [SnowContext(s) for s in sc]
    for c in ctxs:
        c.init()
        c.step(c.scenes[0].dt)
    for c in ctxs:
        c.sync()
    # (a) scene by scene
    t0 = time.perf_counter()
    for c in ctxs:
        dt = c.scenes[0].dt
        for _ in range(steps):
            c.step(dt)
        c.sync()
    el_seq = time.perf_counter() - t0
    # (b) all scenes' contexts interleaved (concurrent streams)
    t0 = time.perf_counter()
    for _ in range(steps):
        for c in ctxs:
            c.step(c.scenes[0].dt)
    for c in ctxs:
        c.sync()
    el_con = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    single = n * steps / el_seq
    concurrent = n * steps / el_con
    return {"value": batched, "scene_by_scene": single, "speedup": batched / single,
            "scene_by_scene_concurrent_streams": concurrent, "speedup_vs_concurrent_streams": batched / concurrent,
            "scenes": len(sc), "particles": n, "steps": steps, "launch_shape": shape,
            "grids": sorted({s.grid.dimx for s in sc}), "unit": "particle-steps/s"}


def stress_slab(steps: int, error_steps: int) -> dict:
    """The mixed-precision stress scene: a 2^24-particle slab on 256^3 dropped on a box, hit by a
    sphere at 10 m/s, physical weight gradient (1/h) so that particles reach the plastic clamp.
    The impactor starts in contact with the slab. Throughput of the plain and mixed forms at full
    size; errors of both against the FP64 oracle on the same grid and geometry with 2^20 particles
    (a size the oracle steps in about a second)."""
    from snow_amd import FLAG_MIXED, SnowContext, scenes
    sc = scenes.slab_impact(grad_scale=0.0, impactor_y=0.8)
    n = len(sc.particles)
    out = {"workload": "slab_impact_256^3", "particles": n, "steps": steps, "unit": "particle-steps/s",
           "note": "impactor starts in contact; timed after 40 steps (1 ms) of impact"}
    for name, flags in (("plain", 0), ("mixed", FLAG_MIXED)):
        with SnowContext(sc, flags) as ctx:
            ctx.init()
            for _ in range(40):
                ctx.step(sc.dt)
            el = timed_steps(ctx, steps, sc.dt)
            out[name] = n * steps / el
            if flags:
                out["fp64_fraction_last_step"] = ctx.fp64_particles() / n
    small = scenes.slab_impact(n_particles=1 << 20, grad_scale=0.0, impactor_y=0.8)
    out["errors"] = error_reports(small, {"plain": 0, "mixed": FLAG_MIXED}, error_steps)
    out["errors"]["grid"] = 256
    return out


def semi_implicit(steps: int) -> dict:
    """The semi-implicit step (snow_step_semi_implicit, beta = 1) on the headline geometry with the
    physical weight gradient (1/h; with the reference's gradient the stiffness term is h^2 smaller and
    the solve stops at its first residual), once the balls are in contact (600 explicit steps of 5e-5 s
    first), at four times the explicit dt: ms per step, solver iterations, particle-steps/s; and its
    error against the FP64 oracle on the 64^3 snowball with a deformed F_E and the same gradient after
    10 steps at five times that scene's dt."""
    import numpy as np
    from oracle import pyoracle
    from snow_amd import SnowContext, scenes
    sc = scenes.two_snowballs(grad_scale=0.0)
    n = len(sc.particles)
    dt, tol, max_iter = 4 * sc.dt, 1e-5, 100
    out = {"workload": "two_snowballs_128^3 in contact, gradient 1/h", "particles": n, "dt": dt, "beta": 1.0,
           "tol": tol,
           "steps": steps, "unit": "particle-steps/s"}
    with SnowContext(sc) as ctx:
        ctx.init()
        for _ in range(600):
            ctx.step(sc.dt)
        ctx.step_semi_implicit(dt, 1.0, tol, max_iter)
        ctx.sync()
        its = []
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.step_semi_implicit(dt, 1.0, tol, max_iter)
            its.append(ctx.implicit_stats()[0])
        ctx.sync()
        el = time.perf_counter() - t0
        ctx.check_finite()
        t0 = time.perf_counter()
        for _ in range(4 * steps):
            ctx.step(sc.dt)
        ctx.sync()
        el_ex = time.perf_counter() - t0
    out.update({"value": n * steps / el, "ms_per_step": el / steps * 1e3, "iterations": its,
                "ms_per_iteration": (el - steps * el_ex / (4 * steps)) / max(1, sum(its)) * 1e3,
                "explicit_ms_per_step": el_ex / (4 * steps) * 1e3})
    # error against the oracle (deformed F_E and the physical gradient, so that the solve iterates)
    small = scenes.snowball_drop(grad_scale=0.0)
    rng = np.random.default_rng(5)
    m = len(small.particles)
    fe = np.eye(3)[None] + 0.02 * rng.uniform(-1, 1, (m, 3, 3))
    small.particles.fe = np.ascontiguousarray(fe.transpose(0, 2, 1).reshape(m, 9), np.float32)
    sdt, ssteps = 5 * small.dt, 10
    o = pyoracle.Oracle(small, nthreads=min(16, host_cpu()["usable_cpus"]))
    o.init()
    o.rho0[:] = o.rho0.astype(np.float32)
    with SnowContext(small) as ctx:
        ctx.init()
        ctx.write_particles(rho0=o.rho0.astype(np.float32))
        for _ in range(ssteps):
            ctx.step_semi_implicit(sdt, 1.0, 1e-6, 200)
        g_it = ctx.implicit_stats()[0]
        p = ctx.read_particles()
    for _ in range(ssteps):
        o_it, _ = o.step_semi_implicit(sdt, 1.0, 1e-6, 200)
    err = {"workload": "snowball_drop_64^3, deformed F_E, gradient 1/h", "particles": m, "dt": sdt, "steps": ssteps,
           "iterations_last_step": {"gpu": g_it, "oracle": o_it}}
    for name, a, b in (("x", p["x"], o.x), ("v", p["v"], o.v), ("F_E", p["fe"], o.fe), ("F_P", p["fp"], o.fp)):
        d = np.sqrt(((np.asarray(a, np.float64) - b) ** 2).sum(-1))
        err[name] = {"max": float(d.max()), "rms": float(np.sqrt((d ** 2).mean()))}
    out["errors"] = err
    return out


def large_batch(steps: int) -> dict:
    """512 small scenes (1k-64k particles, grids 16^3-64^3) plus one 2^25-particle 512^3 scene in one
    step graph; throughput of the whole batch and of its two parts run alone."""
    from snow_amd import SnowContext, scenes
    sc = scenes.large_batch()
    parts = {"all": sc, "small_scenes": sc[:-1], "large_scene": sc[-1:]}
    out = {"scenes": len(sc), "unit": "particle-steps/s", "steps": steps}
    for name, part in parts.items():
        n = sum(len(s.particles) for s in part)
        with SnowContext(part) as ctx:
            ctx.init()
            ctx.run_graph(3)
            el = timed_steps(ctx, steps, graph=True)
        out[name] = {"particles": n, "value": n * steps / el}
        print(f"[bench] large batch: {name} done", file=sys.stderr, flush=True)
    return out


def reference_scenes(steps: int) -> dict:
    """The reference's own scene programs (src/exe/*.cpp, include/snow_scenes.hpp) on the reference
    grid (201^3, h = 0.05, dt = 1e-3), each stepped alone: particle-steps/s and ms per step. The bunny
    scene is left out (its particle file is not in the reference repository)."""
    from snow_amd import SnowContext, scenes
    out = {"unit": "particle-steps/s", "steps": steps, "grid": [201, 201, 201]}
    for name in scenes.REFERENCE_SCENES:
        if name == "snowBunnyScene":
            continue
        sc = scenes.reference(name)
        n = len(sc.particles)
        with SnowContext(sc) as ctx:
            ctx.init()
            for _ in range(5):
                ctx.step(sc.dt)
            el = timed_steps(ctx, steps, dt=sc.dt)
        out[name] = {"particles": n, "value": n * steps / el, "ms_per_step": el / steps * 1e3}
        print(f"[bench] reference scene {name} done", file=sys.stderr, flush=True)
    return out


def copy_calibration(mib: int) -> None:
    """One device-to-device hipMemcpy of `mib` MiB (HIP's copy kernel): the reference point for the
    FETCH_SIZE / WRITE_SIZE counters of the traffic measurement."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--form", choices=["plain", "mixed"], default="plain")
    ap.add_argument("--particles", type=int, default=1 << 20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of snow_autotune's choice")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-second-form", action="store_true", help="skip the mixed (or plain) second form")
    ap.add_argument("--no-stress", action="store_true", help="skip the 2^24-particle slab stress scene")
    ap.add_argument("--no-semi-implicit", action="store_true", help="skip the semi-implicit step")
    ap.add_argument("--stress-steps", type=int, default=10)
    ap.add_argument("--large", action="store_true", help="also run the 512+1-scene batch (2^25 + ~8M particles)")
    ap.add_argument("--large-steps", type=int, default=10)
    ap.add_argument("--reference-scenes", action="store_true",
                    help="also time the reference's own scene programs (src/exe/*.cpp) on the 201^3 grid")
    ap.add_argument("--only-sweep", action="store_true", help="run only the batched 64-scene sweep")
    ap.add_argument("--only-semi-implicit", action="store_true", help="run only the semi-implicit step")
    ap.add_argument("--error-steps", type=int, default=100)
    ap.add_argument("--stress-error-steps", type=int, default=20)
    ap.add_argument("--sweep-steps", type=int, default=100)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.only_sweep:
        print(json.dumps({"forms": {"batched": batched_sweep(args.sweep_steps)}}), flush=True)
        return
    if args.only_semi_implicit:
        print(json.dumps({"forms": {"semi_implicit": semi_implicit(5)}}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (multi-GPU sharding of the "
              f"snow step is a design only)", file=sys.stderr)
        sys.exit(2)

    from snow_amd import FLAG_MIXED, SnowContext, scenes

    print(f"[bench] building scene: two snowballs, {args.particles} particles, 128^3", file=sys.stderr, flush=True)
    scene = scenes.two_snowballs(n_particles=args.particles)
    n = len(scene.particles)
    flags = FLAG_MIXED if args.form == "mixed" else 0
    ctx = SnowContext(scene, flags)
    ctx.init()
    dt = scene.dt
    # the launch shapes that leave the results unchanged (G2P block, sort digit), timed by
    # snow_autotune on a snapshot of the state that it restores afterwards
    shape = None if args.no_autotune else ctx.autotune(steps=5)
    for _ in range(args.warmup):
        ctx.step(dt)
    ctx.sync()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    # Per-kernel times come from HIP events on the context stream around every TIMER_EVERY-th step
    # of the timed region (an event pair between kernels adds a few microseconds of idle time, so
    # timing every step would slow the measured loop itself).
    ctx.reset_stage_times()
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        timed = i % TIMER_EVERY == 0
        if timed:
            ctx.enable_stage_timers(True)
        ctx.step(dt)
        if timed:
            ctx.enable_stage_timers(False)
    ctx.sync()
    el = time.perf_counter() - t0
    ctx.check_finite()
    times = ctx.stage_times()
    active = ctx.active_nodes(0)
    print(f"[bench] timed {args.steps} steps in {el:.3f} s", file=sys.stderr, flush=True)

    kernels = {}
    for st, (bp, bn) in BYTES.items():
        ms, calls = times.get(st, (0.0, 0))
        if calls == 0:
            continue
        avg_s = ms / calls / 1e3
        alg = bp * n + bn * active
        kernels[KERNEL_NAMES[st]] = {"avg_us": avg_s * 1e6, "alg_bytes": alg, "GBps": alg / avg_s / 1e9,
                                     "launches": list(STAGE_LAUNCHES[KERNEL_NAMES[st]])}
        vf = valu_floor(KERNEL_NAMES[st], avg_s * 1e6)
        if vf:
            kernels[KERNEL_NAMES[st]]["valu_floor"] = vf
    sort_ms, sort_calls = times.get("#Sim:Sort Particles", (0.0, 0))
    if sort_calls:
        kernels["sort+cell_bounds"] = {"avg_us": sort_ms / sort_calls * 1e3}
    dom = max((k for k in kernels if "GBps" in k or "alg_bytes" in kernels[k]), key=lambda k: kernels[k]["avg_us"])
    achieved = kernels[dom]["GBps"]
    value = n * args.steps / el
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.form == "plain" else "fp32+fp64-svd",
        "data": "synthetic (seeded SplitMix64 particle sampling)",
        "config": {"workload": "two_snowballs_128^3", "particles": n, "grid": [128, 128, 128], "dt": dt,
                   "form": args.form, "active_nodes": active, "launch_shape": shape},
        "roofline": {"bound": "hbm", "kernel": dom, "launches": kernels[dom]["launches"], "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": stage_traffic(dom), "valu_floor": kernels[dom].get("valu_floor")},
        "step_roofline_frac": (296 * n + 92 * active) * args.steps / el / (HBM_PEAK_GBS * 1e9),
        "kernels": kernels,
    }
    ctx.close()

    out["forms"] = {args.form: {"value": value, "unit": "particle-steps/s"}}
    if not args.no_second_form:
        other = FLAG_MIXED if args.form == "plain" else 0
        with SnowContext(scene, other) as ctx2:
            ctx2.init()
            for _ in range(args.warmup):
                ctx2.step(dt)
            el2 = timed_steps(ctx2, args.steps, dt)
            fp64 = ctx2.fp64_particles() if other else 0
        out["forms"]["mixed" if other else "plain"] = {"value": n * args.steps / el2, "unit": "particle-steps/s",
                                                       "fp64_particles_last_step": fp64}
        print("[bench] second form timed", file=sys.stderr, flush=True)
        sc8 = scenes.two_snowballs_8ppc()
        with SnowContext(sc8) as ctx8:
            ctx8.init()
            for _ in range(args.warmup):
                ctx8.step(dt)
            el8 = timed_steps(ctx8, args.steps, dt)
        out["forms"]["two_snowballs_8ppc"] = {"particles": len(sc8.particles), "value": len(sc8.particles) *
                                              args.steps / el8, "ms_per_step": el8 / args.steps * 1e3,
                                              "unit": "particle-steps/s"}
    if not args.no_errors:
        out["errors"] = error_reports(scene, {"plain": 0, "mixed": FLAG_MIXED}, args.error_steps)
        print("[bench] error report done", file=sys.stderr, flush=True)
    if not args.no_sweep:
        out["forms"]["batched"] = batched_sweep(args.sweep_steps)
        print("[bench] batched sweep done", file=sys.stderr, flush=True)
    if not args.no_stress:
        out["forms"]["stress_slab"] = stress_slab(args.stress_steps, args.stress_error_steps)
        print("[bench] stress slab done", file=sys.stderr, flush=True)
    if not args.no_semi_implicit:
        out["forms"]["semi_implicit"] = semi_implicit(5)
        print("[bench] semi-implicit step done", file=sys.stderr, flush=True)
    if args.large:
        out["forms"]["large_batch"] = large_batch(args.large_steps)
    if args.reference_scenes:
        out["forms"]["reference_scenes"] = reference_scenes(100)
        print("[bench] large batch done", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
