"""Fixture specifications shared by the oracle and engine tests.

Each builder restates a setup that the reference's own tests or published
runs use, with the file:line it comes from, so that golden numbers quoted
from the reference apply to it.
"""
import numpy as np

import oracle as O

GRAVITY = 9.81


def water():
    # core/tests/hydrostatic_balance.rs:24 Material::liquid(1000, 1e5, 0.2, 1e-3)
    return O.liquid(1000.0, 1e5, 0.2, 1e-3)


def water_column(domain, height, spacing, max_dt, friction=0.4):
    """core/tests/hydrostatic_balance.rs:21-48."""
    cfg = O.config(wall_friction=friction, cfl=0.3, max_dt=max_dt, min_dt=1e-9)
    return O.fixture("hydrostatic", water(), spacing, domain, [domain[0], domain[1], height], cfg, seed=7)


def hydrostatic_errors(material, pos, volume_ratio, h0, depth=None):
    """Normalised pressure errors, fixtures.rs:529-561 (liquid branch)."""
    scale = material.density * GRAVITY * h0
    if depth is None:
        depth = np.maximum(h0 - pos[:, 2], 0.0)
    k = material.bulk_modulus
    ref = k * (np.exp(material.density * GRAVITY * depth / k) - 1.0)
    meas = k * (1.0 / volume_ratio - 1.0)
    return (meas - ref) / scale


def research_slump():
    """examples/research_slump.yaml through research.rs:171-232 (shear units)."""
    m = O.paste(1000.0, 1000.0, 0.2, 30.0, 10.0, 0.6, shear_units=True)
    cfg = O.config(wall_friction=0.4, cfl=0.3, max_dt=1e-4, min_dt=1e-12)
    return O.fixture("slump", m, 0.005, [0.06] * 3, [0.02] * 3, cfg, seed=7, max_particles=100_000,
                     max_nodes=1_000_000), 0.5, 0.01


def research_coupled_patch():
    """examples/research_coupled_patch.yaml through research.rs:171-232."""
    m = O.paste(1000.0, 1000.0, 0.2, 30.0, 10.0, 0.6, shear_units=True)
    cfg = O.config(wall_friction=0.4, cfl=0.3, max_dt=5e-5, min_dt=1e-12)
    return O.fixture("coupled_patch", m, 0.002, [0.04] * 3, [0.02] * 3, cfg, seed=7,
                     pellet=(0.003, 0.012, 1100.0), max_particles=100_000, max_nodes=1_000_000), 0.05, 0.005


def record_loop(sim, duration, interval, steps_cap=None):
    """research.rs:86-101: advance to successive record boundaries."""
    k = 1
    while sim.history().time < duration:
        target = min(float(k) * interval, duration)
        sim.advance_to(target)
        k += 1


def prestressed_block(friction, velocity):
    """core/tests/wall_friction.rs:12-52: uniaxially prestressed paste block on the floor."""
    m = O.paste(1000.0, 1e5, 0.0, 1e6, 1.0, 1.0)
    cfg = O.config(wall_friction=friction, cfl=0.3, max_dt=5e-5, min_dt=1e-9, contact_friction=friction)
    f = O.fixture("slump", m, 0.002, [0.04, 0.04, 0.03], [0.02, 0.02, 0.01], cfg, initial_velocity=velocity,
                  seed=7)
    sim = O.Sim.from_fixture(f)
    st = sim.state()
    depth = 0.01 - st.pos[:, 2]
    pressure = m.density * 9.81 * depth
    st.deformation[:] = 0.0
    st.deformation[:, 0, 0] = 1.0
    st.deformation[:, 1, 1] = 1.0
    st.deformation[:, 2, 2] = np.exp(-pressure / 1e5)
    st.kirchhoff[:] = 0.0
    st.kirchhoff[:, 2, 2] = -pressure
    sim.set_state(st)
    led = sim.ledger()
    led.initial_mechanical_energy = sim.mechanical_energy()
    sim.set_ledger(led)
    return f, sim


def paste_column_32(particles_scale=1.0, seed=0, jitter=0.5, cfl=0.4, max_dt=1e-4, friction=0.4):
    """Herschel-Bulkley paste column 40x40x60 mm on a 32^3 grid, dx = 2.5 mm (49,152 particles).

    rho 1200, K 2e5, G 5e4 (E = 9KG/(3K+G), nu = (3K-2G)/(2(3K+G))), shear
    yield 50 Pa, K_s 5 Pa s^n, n 0.5; jitter +-0.25 dx = 0.5 of the sub-cell.
    """
    bulk, shear = 2e5, 5e4
    young = 9.0 * bulk * shear / (3.0 * bulk + shear)
    poisson = (3.0 * bulk - 2.0 * shear) / (2.0 * (3.0 * bulk + shear))
    m = O.paste(1200.0, young, poisson, 50.0, 5.0, 0.5, shear_units=True)
    cfg = O.config(wall_friction=friction, cfl=cfl, max_dt=max_dt, min_dt=1e-12)
    dx = 0.0025
    size = [0.04 * particles_scale, 0.04 * particles_scale, 0.06 * particles_scale]
    size = [max(round(s / (dx / 2)) * dx / 2, dx) for s in size]
    return O.fixture("slump", m, dx, [32 * dx] * 3, size, cfg, seed=seed, jitter=jitter)


def dam_break_block(cells=(128, 128, 128), block=(64, 128, 96), dx=0.001, seed=1, gamma=1.0, jitter=0.5,
                   cfl=0.4, max_dt=1e-4, friction=0.0):
    """Water dam break: a block in the x=0, y=0 corner of the grid, Tait EOS (any size)."""
    # K = 2e5 from the wire pair E = 3.6e5, nu = 0.2 (constitutive.rs:211-223).
    m = O.liquid(1000.0, 3.6e5, 0.2, 1e-3, tait_gamma=gamma)
    cfg = O.config(wall_friction=friction, cfl=cfl, max_dt=max_dt, min_dt=1e-12)
    domain = [c * dx for c in cells]
    size = [b * dx for b in block]
    return O.fixture("dam_break", m, dx, domain, size, cfg, seed=seed, jitter=jitter)
