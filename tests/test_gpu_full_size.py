"""The 1M-particle paste column at full size, through size-independent checks.

The oracle cannot step a million particles in a test's time, so this case is
checked through what holds at any size: the particle set is preserved, the
mass ledger closes, the momentum ledger matches the state to f32 round-off,
no step is rejected across two locality sorts, and the state stays inside the
domain and finite.  The parity against the oracle itself runs on the smaller
cases of test_gpu_parity.py.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def column(E):
    # 100^3 particles (2 per cell and axis) of a Herschel-Bulkley paste on a
    # 96^3 grid, dx 2.5 mm: free-slip walls, sticky floor, CFL 0.4.
    bulk, shear = 2e5, 5e4
    young = 9.0 * bulk * shear / (3.0 * bulk + shear)
    poisson = (3.0 * bulk - 2.0 * shear) / (2.0 * (3.0 * bulk + shear))
    mat = E.paste(1200.0, young, poisson, 50.0, 5.0, 0.5, shear_units=True)
    dx = 0.0025
    domain = [96 * dx] * 3
    cfg = E.config(wall_friction=0.0, cfl=0.4, max_dt=1e-4, min_dt=1e-12,
                   face_friction=[0.0, 0.0, 0.0, 0.0, math.inf, 0.0])
    st = E.place_particles("slump", mat, dx, domain, [0.125] * 3, seed=0, jitter=0.5)
    return mat, cfg, dx, domain, st


def test_million_particle_column_invariants(engine):
    mat, cfg, dx, domain, st = column(engine)
    n = len(st)
    assert n == 1_000_000
    sim = engine.Simulation(mat, cfg, domain, dx, st)
    print("column placed and uploaded", flush=True)
    steps = 70  # crosses the locality sorts after steps 32 and 64
    assert sim.advance_steps(steps) == steps
    h, led = sim.history(), sim.ledger()
    assert h.step_count == steps and h.rejected_steps == 0
    out = sim.state().by_id()
    assert np.array_equal(out.id, np.sort(np.asarray(st.id)))
    # Mass closes in f64 (the exact particle masses are kept by id; two f64 sums
    # of a million masses in different orders differ by at most n eps); momentum
    # matches the state to f32 round-off of the gravity impulse.
    assert abs(sim.mass_residual()) <= 1e-10 * led.initial_mass
    g = np.linalg.norm(led.gravity_impulse)
    assert g > 0.0
    assert np.linalg.norm(sim.momentum_residual()) < 1e-5 * g
    # Every particle is still inside the padded grid (a stencil base in range:
    # -dx/2 <= x < L + dx/2; the walls act on the grid, so a particle may sit
    # slightly past a face), every field is finite, and the plastic strain never
    # goes below zero.
    assert np.isfinite(out.pos).all() and np.isfinite(out.vel).all() and np.isfinite(out.kirchhoff).all()
    assert (out.pos >= -0.5 * dx).all() and (out.pos < domain[0] + 0.5 * dx).all()
    print("lowest particle", out.pos.min(axis=0), flush=True)
    assert (out.plastic >= 0.0).all()
    # Gravity has started the slump: the column moves down, not up.
    assert float(np.mean(out.vel[:, 2])) < 0.0
    print(f"{steps} steps, t = {h.time:.3e} s, max speed {np.abs(out.vel).max():.3e} m/s", flush=True)


def check_invariants(sim, st, steps, dx, domain):
    h, led = sim.history(), sim.ledger()
    assert h.step_count == steps and h.rejected_steps == 0
    out = sim.state().by_id()
    assert np.array_equal(out.id, np.sort(np.asarray(st.id)))
    assert abs(sim.mass_residual()) <= 1e-10 * led.initial_mass
    g = np.linalg.norm(led.gravity_impulse)
    assert g > 0.0
    assert np.linalg.norm(sim.momentum_residual()) < 1e-5 * g
    assert np.isfinite(out.pos).all() and np.isfinite(out.vel).all() and np.isfinite(out.kirchhoff).all()
    for a in range(3):
        assert (out.pos[:, a] >= -0.5 * dx).all() and (out.pos[:, a] < domain[a] + 0.5 * dx).all()
    return out


def test_six_million_particle_dam_break_invariants(engine):
    # Water dam break, 64 x 128 x 96 cells x 8 particles (6,291,456) on a 128^3
    # grid, dx 1 mm, Tait gamma 7, free-slip walls: 40 substeps.
    E = engine
    dx = 0.001
    domain = [128 * dx] * 3
    mat = E.liquid(1000.0, 3.6e5, 0.2, 1e-3, tait_gamma=7.0)
    cfg = E.config(wall_friction=0.0, cfl=0.4, max_dt=1e-4, min_dt=1e-12, face_friction=[0.0] * 6)
    st = E.place_particles("dam_break", mat, dx, domain, [64 * dx, 128 * dx, 96 * dx], seed=1, jitter=0.5)
    assert len(st) == 6_291_456
    sim = E.Simulation(mat, cfg, domain, dx, st)
    print("dam break placed and uploaded", flush=True)
    assert sim.advance_steps(40) == 40
    out = check_invariants(sim, st, 40, dx, domain)
    # The volume ratio J of every particle stays finite and positive.
    assert np.isfinite(out.volume_ratio).all() and (out.volume_ratio > 0.0).all()
    print(f"dam break: t = {sim.history().time:.3e} s", flush=True)


def test_pellet_bed_full_size_invariants(engine):
    # 1.5M-particle paste bed (96 x 96 x 20 cells x 8) on a 96^3 grid, dx 1.5 mm,
    # with 512 rigid pellets in a band from 1.5 mm below to 3 mm above its
    # surface, so that they are coupled from the first step: 40 substeps.
    E = engine
    rng = np.random.default_rng(2)
    dx = 0.0015
    domain = [96 * dx] * 3
    bulk, shear = 2e5, 5e4
    young = 9.0 * bulk * shear / (3.0 * bulk + shear)
    poisson = (3.0 * bulk - 2.0 * shear) / (2.0 * (3.0 * bulk + shear))
    mat = E.paste(1200.0, young, poisson, float(rng.uniform(20.0, 200.0)), 5.0, 0.5, shear_units=True)
    cfg = E.config(wall_friction=0.4, cfl=0.4, max_dt=1e-4, min_dt=1e-12,
                   face_friction=[0.0, 0.0, 0.0, 0.0, math.inf, 0.0])
    st = E.place_particles("slump", mat, dx, domain, [96 * dx, 96 * dx, 20 * dx], seed=2, jitter=0.5)
    pellets = []
    for _ in range(512):
        pos = [rng.uniform(0.006, domain[0] - 0.006), rng.uniform(0.006, domain[1] - 0.006),
               rng.uniform(20 * dx - 0.0015, 20 * dx + 0.003)]
        q = rng.normal(size=4)
        pellets.append(E.new_pellet(0.0015, 0.008, 650.0, pos, orientation=list(q / np.linalg.norm(q))))
    assert len(st) == 1_474_560
    sim = E.Simulation(mat, cfg, domain, dx, st, pellets=pellets)
    print("pellet bed placed and uploaded", flush=True)
    assert sim.advance_steps(40) == 40
    out = check_invariants(sim, st, 40, dx, domain)
    assert (out.plastic >= 0.0).all()
    impulse = np.array(sim.ledger().coupling_impulse[:])
    assert np.isfinite(impulse).all() and np.abs(impulse).max() > 0.0, impulse
    print(f"pellet bed: coupling impulse {impulse}", flush=True)
    print(f"pellet bed: t = {sim.history().time:.3e} s", flush=True)


def test_full_size_256_instance_sweep_invariants(engine):
    # The batched calibration sweep at full size: 256 paste slumps on grids of
    # 16^3..64^3 cells over a 100 mm cube, 40 x 40 mm columns 20-80 mm tall,
    # yield stress U[20, 300] Pa, K U[0.5, 20], n U[0.3, 1] from seed 4, about
    # 11.3M particles, stepped as one batch for 40 substeps each (across a
    # locality sort).  Every instance keeps its particle set, closes its mass
    # ledger, matches its momentum ledger to f32 round-off, takes no rejected
    # step and stays finite inside its grid.
    import bench
    E = engine
    specs = bench.sweep_specs(E)
    total = sum(len(s["state"]) for s in specs)
    print(f"sweep placed: {len(specs)} instances, {total} particles", flush=True)
    assert len(specs) == 256 and total > 10_000_000
    batch = E.Batch(specs)
    status, done_ms = batch.advance(1e9, max_steps=40)
    assert list(status) == [1] * 256
    print(f"40 steps each; done at {np.min(done_ms):.1f}..{np.max(done_ms):.1f} ms", flush=True)
    for i, s in enumerate(specs):
        h, led = batch.history(i), batch.ledger(i)
        assert h.step_count == 40 and h.rejected_steps == 0, i
        out = batch.state(i).by_id()
        assert np.array_equal(out.id, np.sort(np.asarray(s["state"].id))), i
        mass = np.asarray(s["state"].mass, dtype=np.float64)
        assert abs(out.mass.sum() + led.outflow_mass - led.initial_mass) <= 1e-10 * led.initial_mass, i
        mom = (out.vel * out.mass[:, None]).sum(0)
        resid = (mom - np.array(led.initial_momentum[:]) - np.array(led.gravity_impulse[:]) -
                 np.array(led.wall_impulse[:]) - np.array(led.pellet_wall_impulse[:]) +
                 np.array(led.outflow_momentum[:]))
        assert np.linalg.norm(resid) < 1e-5 * np.linalg.norm(led.gravity_impulse), i
        assert np.isfinite(out.pos).all() and np.isfinite(out.vel).all() and np.isfinite(out.kirchhoff).all(), i
        dx, dom = s["spacing"], s["domain"]
        for a in range(3):
            assert (out.pos[:, a] >= -0.5 * dx).all() and (out.pos[:, a] < dom[a] + 0.5 * dx).all(), i
        assert (out.plastic >= 0.0).all(), i
        del out, mass
