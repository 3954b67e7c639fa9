"""Benchmark: particle-substeps/s of the plain implicit MLS-MPM substep at 1M particles.

Workload: a Herschel-Bulkley paste column of 100^3 particles (50^3 cells at 8 per
cell, 125 mm cube, centred on the floor) in a 96^3 grid with dx = 2.5 mm; rho 1200,
bulk modulus 2e5 Pa, shear modulus 5e4 Pa, shear yield stress 50 Pa, K_s 5 Pa s^n,
n 0.5; sticky floor, free-slip walls; adaptive dt at CFL 0.4; jitter +-0.25 dx from
seed 0.  A step is one accepted adaptive substep of the engine (advance_while).
The CPU baseline is the f64 oracle on a bounded sample of the same workload.
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
# Algorithmic HBM bytes per particle of the fused grid-to-particle-to-grid kernel
# (paste): reads x 3, F-I 9, eps_p 1, m 1, V0 1 words; writes x 3, v 3, C 9, F-I 9,
# tau 6, eps_p 1 words; f32.
PASTE_BYTES_PER_PARTICLE = 4 * (15 + 31)
PEAK_HBM = 8.0e12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload():
    from __graft_entry__ import _import_engine
    E = _import_engine()
    bulk, shear = 2e5, 5e4
    young = 9.0 * bulk * shear / (3.0 * bulk + shear)
    poisson = (3.0 * bulk - 2.0 * shear) / (2.0 * (3.0 * bulk + shear))
    mat = E.paste(1200.0, young, poisson, 50.0, 5.0, 0.5, shear_units=True)
    dx = 0.0025
    domain = [96 * dx] * 3
    size = [0.125, 0.125, 0.125]
    face = [0.0, 0.0, 0.0, 0.0, math.inf, 0.0]  # free-slip walls, sticky floor
    cfg = E.config(wall_friction=0.0, cfl=0.4, max_dt=1e-4, min_dt=1e-12, face_friction=face)
    return E, mat, cfg, dx, domain, size


def cpu_baseline(size, dx, domain, seconds=10.0, max_steps=20):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    bulk, shear = 2e5, 5e4
    young = 9.0 * bulk * shear / (3.0 * bulk + shear)
    poisson = (3.0 * bulk - 2.0 * shear) / (2.0 * (3.0 * bulk + shear))
    m = O.paste(1200.0, young, poisson, 50.0, 5.0, 0.5, shear_units=True)
    cfg = O.config(wall_friction=0.0, cfl=0.4, max_dt=1e-4, min_dt=1e-12,
                   face_friction=[0.0, 0.0, 0.0, 0.0, math.inf, 0.0])
    spec = O.fixture("slump", m, dx, domain, size, cfg, seed=0, jitter=0.5)
    sim = O.Sim.from_fixture(spec)
    n = len(sim)
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        dt = min(sim.stability_limit(), 1e-4)
        rc = sim.try_step(dt)
        if rc != 0:
            raise RuntimeError(f"oracle step rejected ({rc})")
        steps += 1
        el = time.perf_counter() - t0
        log(f"cpu baseline: step {steps} at {el:.1f} s")
        if el >= seconds:
            break
    el = time.perf_counter() - t0
    return {"value": n * steps / el, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{n} particles x {steps} substeps of the same workload, f64 oracle, {threads} threads "
                      f"(serial P2G and traction, threaded grid update and G2P), {el:.1f} s"}


def read_traffic():
    path = os.path.join(ROOT, "profiles", "g2p2g_paste_1M_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=50)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported by this build "
              "(the MLS-MPM substep is measured on a single MI355X).", file=sys.stderr)
        sys.exit(2)

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
    # Kernel timing pass: HIP events around every fused G2P2G launch (stream mode).
    shape = sim.launch()
    sim.set_launch(use_graph=0)
    sim.enable_profiling(True)
    sim.advance_steps(args.profile_steps)
    avg_ms, launches = sim.g2p2g_profile()
    sim.enable_profiling(False)
    sim.set_launch(use_graph=shape.use_graph)
    log(f"g2p2g avg {avg_ms:.4f} ms over {launches} launches")
    t0 = time.perf_counter()
    accepted = sim.advance_steps(args.steps)
    elapsed = time.perf_counter() - t0
    log(f"timed {accepted} steps in {elapsed:.3f} s")
    h = sim.history()
    value = n * accepted / elapsed
    bytes_per_launch = PASTE_BYTES_PER_PARTICLE * n
    achieved = bytes_per_launch / (avg_ms * 1e-3) if avg_ms > 0 else 0.0
    traffic = read_traffic()
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(size, dx, domain, seconds=args.cpu_seconds)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": accepted,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / max(accepted, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "HB paste column 100^3 particles (1,000,000) on a 96^3 grid, dx 2.5 mm, plain "
                               "implicit HB, CFL 0.4, sticky floor, free-slip walls, seed 0",
                   "particles": n, "grid_cells": [96, 96, 96], "rejected_steps": h.rejected_steps,
                   "dt_last_s": h.max_dt_used, "launch": sim.launch().as_dict()},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "kernel": "g2p2g_kernel", "kernel_ms": avg_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
