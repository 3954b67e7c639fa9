"""Benchmark of the MLS-MPM substep engine on one MI355X.

Default workload (the headline): particle-substeps/s of the plain implicit
substep at 1M particles.  A Herschel-Bulkley paste column of 100^3 particles
(50^3 cells at 8 per cell, 125 mm cube, centred on the floor) in a 96^3 grid
with dx = 2.5 mm; rho 1200, bulk modulus 2e5 Pa, shear modulus 5e4 Pa, shear
yield stress 50 Pa, K_s 5 Pa s^n, n 0.5; sticky floor, free-slip walls;
adaptive dt at CFL 0.4; jitter +-0.25 dx from seed 0.  A step is one accepted
adaptive substep, run on the device loop without a host callback.  The CPU
baseline is the f64 oracle on a bounded sample of the same workload.

Other workloads (--workload), one JSON line each:
  liquid     6,291,456-particle dam break, Tait gamma 7, 128^3 grid, dx 1 mm, seed 1
  pellets    1,474,560-particle paste bed + 512 rigid pellets, 96^3 grid, dx 1.5 mm, seed 2
  surrogate  the same bed without pellets, exact return vs the learned return
  sweep      256 paste slumps of 16^3..64^3 grids stepped as one batch, seed 4
  copy       a plain 256 MiB device copy (calibration of the HBM byte counters)
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-substeps/s at 1M particles; surrogate vs implicit HB at ≤1e-3 stress err"
# Algorithmic HBM bytes per particle-substep of the substep's two particle
# kernels together, the f32 state read once and written once: paste reads x3 v3
# C9 F9 tau6 eps1 m1 V0 1 = 33 words and writes x3 v3 C9 F9 tau6 eps1 = 31 words
# (256 B); liquid reads x3 v3 C9 J1 tau6 m1 V0 1 = 24 words and writes 22 (184 B).
PASTE_BYTES_PER_PARTICLE = 4 * (33 + 31)
LIQUID_BYTES_PER_PARTICLE = 4 * (24 + 22)
# Per kernel: the deposit reads x3 v3 C9 tau6 m1 V0 1 = 23 words; G2P reads
# x3 F9 eps1 m1 V0 1 = 15 words (liquid: x3 J1 m1 V0 1 = 6) and writes the state.
P2G_BYTES_PER_PARTICLE = 4 * 23
G2P_PASTE_BYTES_PER_PARTICLE = 4 * (15 + 31)
G2P_LIQUID_BYTES_PER_PARTICLE = 4 * (6 + 22)
PEAK_HBM = 8.0e12
# The reference's largest published CPU rate of this path: its slump refinement
# at dx 2.5 mm, 4096 particles x 10000 substeps in 7.9454 s on an Apple M4, f64
# (docs/formal/data/refinement-report.json).
REFERENCE_PUBLISHED = 5.16e6
REFERENCE_BASIS = ("reference's largest published CPU rate: slump refinement dx 2.5 mm, 4096 particles x 10000 "
                   "substeps, Apple M4, f64 (5.16e6 particle-substeps/s)")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def engine():
    from __graft_entry__ import _import_engine
    return _import_engine()


def paste_material(E, shear_yield=50.0):
    bulk, shear = 2e5, 5e4
    young = 9.0 * bulk * shear / (3.0 * bulk + shear)
    poisson = (3.0 * bulk - 2.0 * shear) / (2.0 * (3.0 * bulk + shear))
    return E.paste(1200.0, young, poisson, shear_yield, 5.0, 0.5, shear_units=True)


FACES = [0.0, 0.0, 0.0, 0.0, math.inf, 0.0]  # free-slip walls, sticky floor


def workload():
    E = engine()
    mat = paste_material(E)
    dx = 0.0025
    domain = [96 * dx] * 3
    size = [0.125, 0.125, 0.125]
    cfg = E.config(wall_friction=0.0, cfl=0.4, max_dt=1e-4, min_dt=1e-12, face_friction=FACES)
    return E, mat, cfg, dx, domain, size


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """All cores this process may use: its CPU affinity, within the OpenMP
    thread count the host sets for it (its CPU share), if any."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(avail, int(omp)) if omp.isdigit() and int(omp) > 0 else avail


def _oracle_steps(O, sim, threads, seconds, max_steps):
    O.set_threads(threads)
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        dt = min(sim.stability_limit(), 1e-4)
        rc = sim.try_step(dt)
        if rc != 0:
            raise RuntimeError(f"oracle step rejected ({rc})")
        steps += 1
        el = time.perf_counter() - t0
        log(f"cpu baseline ({threads} threads): step {steps} at {el:.1f} s")
        if el >= seconds:
            break
    return steps, time.perf_counter() - t0


def cpu_baseline(size, dx, domain, seconds=10.0, max_steps=20):
    """The f64 oracle on the same workload, timed on this host's cores: all the
    cores the process may use, then one thread, each on a bounded sample of
    steps from the same placed state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    bulk, shear = 2e5, 5e4
    young = 9.0 * bulk * shear / (3.0 * bulk + shear)
    poisson = (3.0 * bulk - 2.0 * shear) / (2.0 * (3.0 * bulk + shear))
    m = O.paste(1200.0, young, poisson, 50.0, 5.0, 0.5, shear_units=True)
    cfg = O.config(wall_friction=0.0, cfl=0.4, max_dt=1e-4, min_dt=1e-12, face_friction=FACES)
    spec = O.fixture("slump", m, dx, domain, size, cfg, seed=0, jitter=0.5)
    base = O.Sim.from_fixture(spec)
    n = len(base)
    threads = cpu_threads()
    steps, el = _oracle_steps(O, base.clone(), threads, seconds, max_steps)
    steps1, el1 = _oracle_steps(O, base.clone(), 1, seconds, max_steps)
    return {"value": n * steps / el, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{n} particles x {steps} substeps of the same workload from the same placed state, f64 "
                      f"oracle, {threads} threads (serial P2G and traction, threaded grid update and G2P), {el:.1f} s",
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "single_thread": {"value": n * steps1 / el1, "unit": "particle-substeps/s", "cores": 1,
                              "sample": f"{n} particles x {steps1} substeps, 1 thread, {el1:.1f} s"}}


def read_traffic():
    path = os.path.join(ROOT, "profiles", "substep_paste_1M_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def timed_profiled(sim, steps):
    """Time `steps` accepted steps, with the G2P and deposit kernels of the trial
    steps of this timed region timed on the device (each kernel stamps its first
    workgroup start and last workgroup end; no idle time is added between
    kernels).  The stamps cost about 4 us per timed step at 1M particles, so in
    a long region every k-th step is timed, k = steps // 128 (at least 128
    timed launches); up to 255 steps, every step.  Timing needs stream
    launches, so the timed region runs in stream mode."""
    sim.set_launch(use_graph=0)
    sim.enable_profiling(True, every=max(1, steps // 128))
    t0 = time.perf_counter()
    accepted = sim.advance_steps(steps)
    elapsed = time.perf_counter() - t0
    g2p_ms, p2g_ms, launches = sim.kernel_profile()
    sim.enable_profiling(False)
    log(f"timed {accepted} steps in {elapsed:.3f} s; g2p avg {g2p_ms:.4f} ms, p2g avg {p2g_ms:.4f} ms over "
        f"{launches} timed steps")
    return accepted, elapsed, (g2p_ms, p2g_ms), launches


def roofline(bytes_per_particle, g2p_bytes, n, ms, launches, traffic=None):
    """HBM roofline of the substep's particle kernels (deposit + G2P): the
    algorithmic bytes per particle-substep over their summed average time, and
    each kernel on its own bytes."""
    g2p_ms, p2g_ms = ms
    both = g2p_ms + p2g_ms
    per_launch = bytes_per_particle * n
    achieved = per_launch / (both * 1e-3) if both > 0 else 0.0

    def one(b, t):
        return {"ms": t, "algorithmic_bytes_per_launch": b * n,
                "achieved_gbs": b * n / (t * 1e-3) / 1e9 if t > 0 else 0.0}

    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": achieved / PEAK_HBM, "traffic": traffic, "kernel": "p2g_kernel + g2p_kernel",
            "kernel_ms": both, "kernel_launches_timed": launches, "algorithmic_bytes_per_launch": per_launch,
            "kernels": {"g2p_kernel": one(g2p_bytes, g2p_ms), "p2g_kernel": one(P2G_BYTES_PER_PARTICLE, p2g_ms)}}


def line(metric, value, unit, steps, warmup, elapsed, dtype, config, rl=None, cpu=None, vs=None):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": 1e3 * elapsed / max(steps, 1), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": vs, "dtype": dtype, "data": "synthetic", "config": config, "roofline": rl,
           "cpu_baseline": cpu}
    if vs is not None:
        out["vs_baseline_basis"] = REFERENCE_BASIS
    return out


# ---------------------------------------------------------------- headline

def run_headline(args):
    E, mat, cfg, dx, domain, size = workload()
    t0 = time.perf_counter()
    st = E.place_particles("slump", mat, dx, domain, size, seed=0, jitter=0.5)
    n = len(st)
    log(f"placed {n} particles in {time.perf_counter() - t0:.1f} s")
    sim = E.Simulation(mat, cfg, domain, dx, st)
    log("engine created")
    sim.advance_steps(args.warmup)
    log(f"warmup {args.warmup} steps done")
    if not args.no_autotune:
        chosen = sim.autotune(steps=16)
        log(f"autotuned launch shape: {chosen.as_dict()}")
    accepted, elapsed, ms, launches = timed_profiled(sim, args.steps)
    h = sim.history()
    cpu = None if args.no_cpu_baseline else cpu_baseline(size, dx, domain, seconds=args.cpu_seconds)
    value = n * accepted / elapsed
    return line(METRIC, value, "particle-substeps/s", accepted, args.warmup, elapsed, "f32",
                {"workload": "HB paste column 100^3 particles (1,000,000) on a 96^3 grid, dx 2.5 mm, plain "
                             "implicit HB, CFL 0.4, sticky floor, free-slip walls, seed 0",
                 "particles": n, "grid_cells": [96, 96, 96], "rejected_steps": h.rejected_steps,
                 "dt_last_s": h.max_dt_used, "launch": sim.launch().as_dict()},
                roofline(PASTE_BYTES_PER_PARTICLE, G2P_PASTE_BYTES_PER_PARTICLE, n, ms, launches, read_traffic()),
                cpu, vs=value / REFERENCE_PUBLISHED)


# ---------------------------------------------------------------- liquid (dam break)

def run_liquid(args):
    E = engine()
    dx = 0.001
    cells = [128, 128, 128]
    domain = [c * dx for c in cells]
    block = [64 * dx, 128 * dx, 96 * dx]
    mat = E.liquid(1000.0, 3.6e5, 0.2, 1e-3, tait_gamma=7.0)
    cfg = E.config(wall_friction=0.0, cfl=0.4, max_dt=1e-4, min_dt=1e-12, face_friction=[0.0] * 6)
    st = E.place_particles("dam_break", mat, dx, domain, block, seed=1, jitter=0.5)
    n = len(st)
    log(f"placed {n} particles")
    sim = E.Simulation(mat, cfg, domain, dx, st)
    sim.advance_steps(args.warmup)
    if not args.no_autotune:
        log(f"autotuned launch shape: {sim.autotune(steps=8).as_dict()}")
    accepted, elapsed, ms, launches = timed_profiled(sim, args.steps)
    h = sim.history()
    return line("particle-substeps/s, dam break 6.3M particles, Tait gamma 7", n * accepted / elapsed,
                "particle-substeps/s", accepted, args.warmup, elapsed, "f32",
                {"workload": "water dam break 64x128x96 cells x 8 particles (6,291,456) on a 128^3 grid, dx 1 mm, "
                             "Tait gamma 7, K 2e5 Pa, viscosity 1e-3 Pa s, free-slip walls, CFL 0.4, seed 1",
                 "particles": n, "grid_cells": cells, "rejected_steps": h.rejected_steps,
                 "dt_last_s": h.max_dt_used, "launch": sim.launch().as_dict()},
                roofline(LIQUID_BYTES_PER_PARTICLE, G2P_LIQUID_BYTES_PER_PARTICLE, n, ms, launches))


# ---------------------------------------------------------------- paste bed (+ pellets)

def bed(E, with_pellets, seed=2):
    """96^3 grid, dx 1.5 mm, paste bed 96 x 96 x 20 cells; yield stress U[20, 200] Pa
    (shear) from `seed`; 512 pellets r 1.5 mm, L 8 mm, rho 650 at random positions
    in a band from 1.5 mm below to 3 mm above the bed surface (so that they touch the
    bed within the first steps) and random orientations from `seed`."""
    rng = np.random.default_rng(seed)
    dx = 0.0015
    cells = [96, 96, 96]
    domain = [c * dx for c in cells]
    mat = paste_material(E, shear_yield=float(rng.uniform(20.0, 200.0)))
    cfg = E.config(wall_friction=0.4, cfl=0.4, max_dt=1e-4, min_dt=1e-12, face_friction=FACES)
    st = E.place_particles("slump", mat, dx, domain, [96 * dx, 96 * dx, 20 * dx], seed=seed, jitter=0.5)
    pellets = []
    if with_pellets:
        top = 20 * dx
        for _ in range(512):
            pos = [rng.uniform(0.006, domain[0] - 0.006), rng.uniform(0.006, domain[1] - 0.006),
                   rng.uniform(top - 0.0015, top + 0.003)]
            q = rng.normal(size=4)
            q /= np.linalg.norm(q)
            pellets.append(E.new_pellet(0.0015, 0.008, 650.0, pos, orientation=list(q)))
    return mat, cfg, dx, domain, st, pellets, cells


def run_pellets(args):
    E = engine()
    mat, cfg, dx, domain, st, pellets, cells = bed(E, True)
    n = len(st)
    log(f"placed {n} particles and {len(pellets)} pellets")
    sim = E.Simulation(mat, cfg, domain, dx, st, pellets=pellets)
    sim.advance_steps(args.warmup)
    if not args.no_autotune:
        log(f"autotuned launch shape: {sim.autotune(steps=8).as_dict()}")
    accepted, elapsed, ms, launches = timed_profiled(sim, args.steps)
    h = sim.history()
    led = sim.ledger()
    return line("particle-substeps/s, paste bed 1.5M particles + 512 coupled pellets", n * accepted / elapsed,
                "particle-substeps/s", accepted, args.warmup, elapsed, "f32",
                {"workload": "HB paste bed 96x96x20 cells x 8 particles (1,474,560) on a 96^3 grid, dx 1.5 mm, "
                             "yield stress U[20,200] Pa from seed 2, 512 pellets r 1.5 mm L 8 mm rho 650 at random "
                             "positions in a band from 1.5 mm below to 3 mm above the bed surface and random "
                             "orientations (seed 2), two-way coupled, CFL 0.4",
                 "particles": n, "pellets": len(pellets), "grid_cells": cells, "rejected_steps": h.rejected_steps,
                 "dt_last_s": h.max_dt_used, "coupling_impulse": list(led.coupling_impulse),
                 "launch": sim.launch().as_dict()},
                roofline(PASTE_BYTES_PER_PARTICLE, G2P_PASTE_BYTES_PER_PARTICLE, n, ms, launches))


# ---------------------------------------------------------------- surrogate (cfg 2 bed, no pellets)

def run_surrogate(args):
    E = engine()
    import litter_physics_amd.surrogate as S
    mat, cfg, dx, domain, st, _, cells = bed(E, False)
    n = len(st)
    net = S.load_network()
    plain = E.Simulation(mat, cfg, domain, dx, st)
    learned = E.Simulation(mat, cfg, domain, dx, st)
    learned.set_surrogate(net, tolerance=1e-4, polish=2)
    # Stress against the f64 oracle after a bounded run of fixed steps from the same start.
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.set_threads(min(os.cpu_count() or 1, 16))
    dt = plain.stability_limit()
    ocfg = O.config(wall_friction=0.4, cfl=0.4, max_dt=1e-4, min_dt=1e-12, face_friction=FACES)
    om = O.Material()
    for f, _ in om._fields_:
        setattr(om, f, getattr(mat, f))
    # The oracle's own placement of the same fixture (bit-identical to the engine's).
    osim = O.Sim.from_fixture(O.fixture("slump", om, dx, domain, [96 * dx, 96 * dx, 20 * dx], ocfg, seed=2,
                                        jitter=0.5))
    assert len(osim) == n
    t0 = time.perf_counter()
    for k in range(args.oracle_steps):
        assert osim.try_step(dt) == 0
        assert plain.try_step(dt).accepted == 1 and learned.try_step(dt).accepted == 1
    log(f"{args.oracle_steps} fixed steps on oracle and both engines in {time.perf_counter() - t0:.1f} s")
    ref = osim.state()
    o = np.argsort(ref.id)

    def stress_err(sim):
        g = sim.state().by_id()
        return float(np.abs(g.kirchhoff - ref.kirchhoff[o]).max() / np.abs(ref.kirchhoff).max())

    err_plain, err_learned = stress_err(plain), stress_err(learned)
    considered0, accepted0 = learned.surrogate_stats()
    out = {}
    for name, sim in (("plain", plain), ("surrogate", learned)):
        sim.advance_steps(args.warmup)
        if name == "surrogate":
            learned.set_surrogate(net, tolerance=1e-4, polish=2)  # count the timed run only
        accepted, elapsed, ms, launches = timed_profiled(sim, args.steps)
        out[name] = {"value": n * accepted / elapsed, "ms_per_step": 1e3 * elapsed / accepted, "g2p_ms": ms[0],
                     "p2g_ms": ms[1]}
    considered, accepted = learned.surrogate_stats()
    fallback = 1.0 - accepted / max(considered, 1)
    return line("particle-substeps/s, 1.5M paste bed with the learned HB return (exact fallback)",
                out["surrogate"]["value"], "particle-substeps/s", args.steps, args.warmup,
                args.steps * out["surrogate"]["ms_per_step"] / 1e3, "f32",
                {"workload": "HB paste bed 96x96x20 cells x 8 particles (1,474,560) on a 96^3 grid, dx 1.5 mm, no "
                             "pellets, MLP 12-64-64-64-7 bf16 MFMA return + 2 Newton corrections, residual gate 1e-4",
                 "particles": n, "plain": out["plain"], "surrogate": out["surrogate"],
                 "speedup_vs_plain": out["surrogate"]["value"] / out["plain"]["value"],
                 "fallback_rate": fallback, "yielded_returns": considered,
                 "stress_err_vs_f64_oracle": {"plain": err_plain, "surrogate": err_learned,
                                              "steps": args.oracle_steps,
                                              "yielded_returns": considered0,
                                              "fallback_rate": 1.0 - accepted0 / max(considered0, 1)}})


# ---------------------------------------------------------------- batched sweep

def sweep_specs(E, count=256, seed=4):
    """256 paste slumps: grids log-uniform 16^3..64^3 cells over a 100 mm cube (dx =
    100 mm / cells), 40 x 40 mm columns 20-80 mm tall, shear yield stress U[20, 300] Pa,
    K_s U[0.5, 20], n U[0.3, 1.0], all from `seed`."""
    rng = np.random.default_rng(seed)
    bulk, shear = 2e5, 5e4
    young = 9.0 * bulk * shear / (3.0 * bulk + shear)
    poisson = (3.0 * bulk - 2.0 * shear) / (2.0 * (3.0 * bulk + shear))
    specs = []
    for i in range(count):
        cells = int(round(math.exp(rng.uniform(math.log(16), math.log(64)))))
        dx = 0.1 / cells
        domain = [cells * dx] * 3
        height = rng.uniform(0.02, 0.08)
        mat = E.paste(1200.0, young, poisson, rng.uniform(20.0, 300.0), rng.uniform(0.5, 20.0),
                      rng.uniform(0.3, 1.0), shear_units=True)
        cfg = E.config(wall_friction=0.0, cfl=0.4, max_dt=1e-4, min_dt=1e-12, face_friction=FACES)
        st = E.place_particles("slump", mat, dx, domain, [0.04, 0.04, height], seed=int(rng.integers(1 << 31)),
                               jitter=0.5)
        specs.append({"material": mat, "cfg": cfg, "domain": domain, "spacing": dx, "state": st, "cells": cells})
    return specs


def timed_batch(E, specs, args, net=None):
    """Warm up, autotune and time `args.steps` trial steps of every instance in one batch."""
    batch = E.Batch(specs)
    if net is not None:
        batch.set_surrogate(net, tolerance=1e-4, polish=2)
    batch.advance(1e9, max_steps=args.warmup)
    chosen = batch.autotune(steps=8).as_dict() if not args.no_autotune else batch.launch().as_dict()
    log(f"autotuned launch shape: {chosen}")
    if net is not None:
        batch.set_surrogate(net, tolerance=1e-4, polish=2)  # count the timed run only
    t0 = time.perf_counter()
    status, done_ms = batch.advance(1e9, max_steps=args.steps)
    elapsed = time.perf_counter() - t0
    stopped = int(np.sum(status == 1))  # instances that took all their steps
    log(f"{stopped} of {len(specs)} instances took all {args.steps} steps in {elapsed:.3f} s")
    return batch, elapsed, stopped, done_ms, chosen


def sequential_sample(E, specs, args, count=16):
    """The same instances one at a time (a sample of `count`, each its own engine,
    plain implicit return), as the reference's sweep runs its jobs in sequence."""
    rng = np.random.default_rng(5)
    pick = rng.choice(len(specs), count, replace=False)
    seq_particles, seq_time = 0, 0.0
    for i in pick:
        s = specs[i]
        sim = E.Simulation(s["material"], s["cfg"], s["domain"], s["spacing"], s["state"])
        sim.advance_steps(args.warmup)
        t1 = time.perf_counter()
        k = sim.advance_steps(args.steps)
        seq_time += time.perf_counter() - t1
        seq_particles += len(s["state"]) * k
        del sim
    return seq_particles / seq_time


SWEEP_WORKLOAD = ("256 paste slumps, grids log-uniform 16^3..64^3 over a 100 mm cube, 40x40 mm columns 20-80 mm "
                  "tall, yield U[20,300] Pa, K U[0.5,20], n U[0.3,1.0], seed 4, {steps} substeps each with its own "
                  "adaptive dt, one batch")


def run_sweep(args):
    E = engine()
    specs = sweep_specs(E)
    counts = np.array([len(s["state"]) for s in specs])
    total = int(counts.sum())
    log(f"{len(specs)} instances, {total} particles ({counts.min()}..{counts.max()} each)")
    batch, elapsed, stopped, done_ms, chosen = timed_batch(E, specs, args)
    del batch
    value = total * args.steps / elapsed
    seq_value = sequential_sample(E, specs, args)
    return line("aggregate particle-substeps/s, 256-instance batched paste sweep", value, "particle-substeps/s",
                args.steps, args.warmup, elapsed, "f32",
                {"workload": SWEEP_WORKLOAD.format(steps=args.steps),
                 "instances": len(specs), "instances_completed": stopped, "particles": total, "launch": chosen,
                 "instance_wall_ms_p50": float(np.percentile(done_ms, 50)),
                 "instance_wall_ms_p99": float(np.percentile(done_ms, 99)),
                 "sequential_sample_value": seq_value, "speedup_vs_sequential": value / seq_value})


def run_surrogate_sweep(args):
    """Surrogate plus batched mode against the plain GPU mode: the 256-instance
    sweep stepped in one batch with the learned return (exact fallback), beside
    the same batch with the exact return and the plain engine run one instance
    at a time.  The stress of every instance is compared with the exact batch
    after the same number of steps."""
    E = engine()
    import litter_physics_amd.surrogate as S
    net = S.load_network()
    specs = sweep_specs(E)
    counts = np.array([len(s["state"]) for s in specs])
    total = int(counts.sum())
    log(f"{len(specs)} instances, {total} particles ({counts.min()}..{counts.max()} each)")
    plain, el_plain, _, _, _ = timed_batch(E, specs, args)
    learned, elapsed, stopped, done_ms, chosen = timed_batch(E, specs, args, net)
    considered, accepted = learned.surrogate_stats()
    errs = []
    for i in range(len(specs)):
        a, b = plain.state(i).by_id(), learned.state(i).by_id()
        errs.append(float(np.abs(a.kirchhoff - b.kirchhoff).max() / max(np.abs(a.kirchhoff).max(), 1e-30)))
    del plain, learned
    value = total * args.steps / elapsed
    plain_value = total * args.steps / el_plain
    seq_value = sequential_sample(E, specs, args)
    return line("aggregate particle-substeps/s, 256-instance sweep, learned HB return in batched mode",
                value, "particle-substeps/s", args.steps, args.warmup, elapsed, "f32",
                {"workload": SWEEP_WORKLOAD.format(steps=args.steps) + ", MLP 12-64-64-64-7 bf16 MFMA return + 2 "
                             "Newton corrections, residual gate 1e-4",
                 "instances": len(specs), "instances_completed": stopped, "particles": total, "launch": chosen,
                 "instance_wall_ms_p50": float(np.percentile(done_ms, 50)),
                 "instance_wall_ms_p99": float(np.percentile(done_ms, 99)),
                 "fallback_rate": 1.0 - accepted / max(considered, 1), "yielded_returns": considered,
                 "stress_rel_diff_vs_exact_batch": {"max": max(errs), "median": float(np.median(errs))},
                 "plain_batched_value": plain_value, "plain_sequential_sample_value": seq_value,
                 "speedup_vs_plain_sequential": value / seq_value, "speedup_vs_plain_batched": value / plain_value})


# ---------------------------------------------------------------- copy calibration

def run_copy(args):
    """A plain device copy of 256 MiB, 10 times: the reference read and write of
    known size against which the HBM counters of the kernels are scaled."""
    import torch
    n = 64 << 20  # floats
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nbytes = 4 * n
    return line("device copy bandwidth (calibration)", 2 * nbytes * 10 / el, "B/s", 10, 0, el, "f32",
                {"workload": f"torch copy of {nbytes} bytes (read {nbytes} B, write {nbytes} B per launch), 10 launches",
                 "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes})


# Timed steps per workload: the substep counts of the configurations they follow.
DEFAULT_STEPS = {"headline": 500, "liquid": 1000, "pellets": 1000, "surrogate": 500, "sweep": 300,
                 "surrogate_sweep": 300, "copy": 10}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 500 for the headline and surrogate, 1000 for liquid and "
                         "pellets, 300 for the sweeps)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "liquid", "pellets", "surrogate", "sweep", "surrogate_sweep", "copy"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--oracle-steps", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported by this build "
              "(the MLS-MPM substep is measured on a single MI355X).", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = DEFAULT_STEPS[args.workload]
    run = {"headline": run_headline, "liquid": run_liquid, "pellets": run_pellets, "surrogate": run_surrogate,
           "sweep": run_sweep, "surrogate_sweep": run_surrogate_sweep, "copy": run_copy}[args.workload]
    print(json.dumps(run(args)), flush=True)


if __name__ == "__main__":
    main()
