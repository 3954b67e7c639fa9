"""Pins of the f64 oracle against the reference's own golden numbers and closed forms.

Every expected value below is quoted from the reference (file:line given) or
follows from mathematics that a plausible restatement mistake would break.
"""
import math

import numpy as np
import pytest

import oracle as O
from cases import (GRAVITY, hydrostatic_errors, prestressed_block, record_loop, research_coupled_patch,
                   research_slump, water, water_column)


# ---------------------------------------------------------------- SplitMix64

def test_splitmix_published_vectors():
    # Steele-Lea-Flood 2014 reference outputs for seed 0 (rng.rs:7), plus the
    # seeds the build uses, computed by independent integer arithmetic.
    assert O.splitmix(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert O.splitmix(1, 2) == [0x910A2DEC89025CC1, 0xBEEB8DA1658EEC67]
    assert O.splitmix(2, 1) == [0x975835DE1C9756CE]
    assert O.splitmix(4, 1) == [0x6E73E372E2338ACA]
    assert O.splitmix(7, 1) == [0x63CBE1E459320DD7]


# ------------------------------------------------------------ constitutive

def test_manuscript_hb_return_golden():
    # core/tests/formal_examples.rs:12-26
    m = O.paste(1000.0, 1000.0, 0.2, 30.0, 10.0, 0.6, shear_units=True)
    trial_stress = 80.0
    strain = trial_stress / (2.0 * math.sqrt(3.0) * m.shear_modulus)
    trial = np.diag([math.exp(strain), math.exp(-strain), 1.0])
    rc, f, tau, inc, dd, eq = O.update_paste(trial, m, 0.0014)
    assert rc == 0
    assert abs(eq - 78.0059568837) < 1e-9
    assert abs(inc - 0.001595234493) < 1e-12
    assert abs(np.linalg.det(f) - 1.0) < 1e-12
    elastic_drop = (trial_stress ** 2 - eq ** 2) / (6.0 * m.shear_modulus)
    gap = 1.5 * m.shear_modulus * inc ** 2
    assert abs(elastic_drop - dd - gap) < 1e-12


def test_return_residual_and_bounds():
    # constitutive.rs:616-633
    for n in (0.3, 0.6, 1.0, 1.5):
        trial, ys, k, mu, dt = 500.0, 20.0, 5.0, 1.0e4, 1e-3
        rc, sigma, inc, _ = O.hb_return(trial, ys, k, n, mu, dt)
        assert rc == 0
        rate = ((sigma - ys) / k) ** (1.0 / n)
        assert abs(sigma - trial + 3.0 * mu * dt * rate) < 1e-8 * trial
        assert ys < sigma < trial
        assert abs(inc - rate * dt) < 1e-6 * rate * dt


def test_return_special_cases():
    # constitutive.rs:610-645
    assert O.hb_return(10.0, 20.0, 5.0, 0.6, 1e4, 1e-3)[2] == 0.0
    assert O.hb_return(500.0, 20.0, 0.0, 1.0, 1e4, 1e-3)[1] == 20.0
    assert O.hb_return(float("nan"), 20.0, 5.0, 1.0, 1e4, 1e-3)[0] == 5


def test_jacobi_reconstructs():
    # constitutive.rs:648-660
    m = np.array([[2.0, -1.0, 0.3], [-1.0, 3.0, 0.5], [0.3, 0.5, 1.5]])
    rc, vals, vecs = O.eigen3(m)
    assert rc == 0
    assert np.linalg.norm(vecs @ np.diag(vals) @ vecs.T - m) < 1e-12
    assert np.linalg.norm(vecs @ vecs.T - np.eye(3)) < 1e-12
    assert np.allclose(np.sort(vals), np.linalg.eigvalsh(m), atol=1e-13)
    near = np.eye(3) + np.array([[0, 1e-9, 0], [1e-9, 0, 0], [0, 0, 2e-9]])
    rc, vals, vecs = O.eigen3(near)
    assert np.linalg.norm(vecs @ np.diag(vals) @ vecs.T - near) < 1e-15
    assert O.eigen3(np.eye(3) * np.nan)[0] != 0


def test_inverted_and_hydrostatic_paste():
    m = O.paste(1000.0, 1.0e5, 0.3, 20.0, 5.0, 0.6)
    assert O.update_paste(np.diag([-1.0, 1.0, 1.0]), m, 1e-3)[0] == 1  # constitutive.rs:663-670
    s = 0.99  # constitutive.rs:672-682
    rc, f, tau, inc, _, _ = O.update_paste(np.eye(3) * s, m, 1e-3)
    assert rc == 0 and inc == 0.0
    expected = (2.0 * m.shear_modulus + 3.0 * m.lame_lambda) * math.log(s)
    assert abs(tau[0, 0] - expected) < 1e-9 * abs(expected)
    assert np.linalg.norm(f - np.eye(3) * s) < 1e-12


def test_steady_simple_shear_reproduces_hb_law():
    # constitutive.rs:688-723: 20,000 implicit steps reach tau = tau_y + K_s rate^n.
    m = O.paste(1000.0, 1.0e6, 0.3, 20.0, 5.0, 0.6, shear_units=True)
    rate, dt = 2.0, 1e-4
    f = np.eye(3)
    grad = np.zeros((3, 3))
    grad[0, 2] = rate
    step = np.eye(3) + grad * dt
    for _ in range(20_000):
        rc, f, tau, inc, _, eq = O.update_paste(step @ f, m, dt)
        assert rc == 0
    expected_shear = 20.0 + 5.0 * rate ** 0.6
    assert abs(eq - math.sqrt(3.0) * expected_shear) / (math.sqrt(3.0) * expected_shear) < 1e-6
    assert abs(tau[0, 2] - expected_shear) / expected_shear < 1e-6
    assert abs(inc / dt - rate / math.sqrt(3.0)) / (rate / math.sqrt(3.0)) < 1e-6


def test_liquid_laws_and_water_golden():
    w = water()
    # formal_examples.rs:29-44
    ratio = math.exp(-w.density * 9.81 * 0.01 / w.bulk_modulus)
    assert abs(O.liquid_pressure(w, ratio) - 98.186663) < 5e-7
    rc, j, tau, p = O.update_liquid(1.0, -np.eye(3), w, 1e-3)
    assert rc == 0 and abs(j - 0.997) < 1e-14
    assert abs(np.linalg.det(np.eye(3) - np.eye(3) * 1e-3) - j - 2.999e-6) < 1e-14
    assert O.update_liquid(1.0, -np.eye(3), w, 0.4)[0] == 1
    # constitutive.rs:739-749
    visc = O.liquid(1000.0, 1.0e5, 0.2, 0.5)
    g = np.zeros((3, 3))
    g[0, 2] = 4.0
    rc, j, tau, p = O.update_liquid(1.0, g, visc, 1e-4)
    assert abs(tau[0, 2] - 2.0) < 1e-12 and abs(p) < 1e-12


def test_tait_reduces_to_reference_and_small_strain_modulus():
    w1 = water()
    w7 = O.liquid(1000.0, 1e5, 0.2, 1e-3, tait_gamma=7.0)
    for j in (0.98, 0.999, 1.0, 1.003):
        assert O.liquid_pressure(w1, j) == w1.bulk_modulus * (1.0 / j - 1.0)
        # (K/g)(J^-g - 1) against its closed form
        assert abs(O.liquid_pressure(w7, j) - w7.bulk_modulus / 7.0 * (j ** -7.0 - 1.0)) < 1e-9
    h = 1e-7
    slope = (O.liquid_pressure(w7, 1.0 + h) - O.liquid_pressure(w7, 1.0 - h)) / (2 * h)
    assert abs(slope + w7.bulk_modulus) < 1e-4 * w7.bulk_modulus


# --------------------------------------------------------------- transfers

def test_single_particle_deposit_moments():
    # grid.rs:324-358 partition of unity and centroid, through the P2G.
    m = O.paste(1000.0, 1e4, 0.3, 1e6, 0.0, 1.0)
    cfg = O.config(gravity=(0, 0, 0))
    for x in ([0.0, 0.0, 0.0], [0.0031, 0.0177, 0.02], [0.0195, 0.0004, 0.011]):
        sim = O.Sim.create(m, cfg, [0.02] * 3, 0.002, [x], [[0.3, -0.2, 0.1]], [1.0], [1e-9])
        mass, mom = sim.probe_grid(1e-4, with_traction=False)
        n = 13
        idx = np.arange(n ** 3)
        coords = np.stack([idx // (n * n), (idx // n) % n, idx % n], 1)
        nodes = (coords - 1.0) * 0.002
        assert abs(mass.sum() - 1.0) < 1e-12
        assert np.linalg.norm((mass[:, None] * nodes).sum(0) - np.array(x)) < 1e-12
        assert np.linalg.norm(mom.sum(0) - np.array([0.3, -0.2, 0.1])) < 1e-12


def test_outside_particle_has_no_stencil():
    m = O.paste(1000.0, 1e4, 0.3, 1e6, 0.0, 1.0)
    sim = O.Sim.create(m, O.config(), [0.02] * 3, 0.002, [[-0.003, 0.0, 0.0]], [[0, 0, 0]], [1.0], [1e-9])
    mass, _ = sim.probe_grid(1e-4, with_traction=False)
    assert mass.sum() == 0.0


# ------------------------------------------------------- end-to-end pins

def test_hydrostatic_footprint_golden():
    # docs/hydrostatic-balance.md:26-31 (4 x 4 mm footprint, 160 particles, 0.002 s)
    spec = water_column([0.004, 0.004, 0.02], 0.01, 0.002, 5e-5)
    sim = O.Sim.from_fixture(spec)
    sim.advance_to(0.002)
    st = sim.state()
    assert len(sim) == 160 and sim.history().step_count == 40
    e = hydrostatic_errors(water(), st.pos, st.volume_ratio, 0.01)
    assert abs(np.abs(e).max() - 0.004780974064310695) < 1e-15
    assert abs(math.sqrt(np.mean(e * e)) - 0.00106261024765249) < 1e-15


def test_hydrostatic_original_column_golden():
    # docs/hydrostatic-balance.md:26-31; docs/validation.md:150-158, 232-236
    spec = water_column([0.02, 0.02, 0.02], 0.01, 0.002, 5e-5)
    sim = O.Sim.from_fixture(spec)
    sim.advance_to(0.01)
    h = sim.history()
    assert len(sim) == 4000 and h.step_count == 200 and h.rejected_steps == 0 and h.limited_steps == 0
    st = sim.state()
    e = hydrostatic_errors(water(), st.pos, st.volume_ratio, 0.01)
    assert abs(np.abs(e).max() - 0.011513856585915002) < 1e-15
    assert abs(math.sqrt(np.mean(e * e)) - 0.0034430221485745437) < 1e-15
    assert sim.mass_residual() == -8.673617379884035e-19
    assert abs(np.linalg.norm(sim.momentum_residual()) - 4.770490026058286e-18) < 1e-21
    assert abs(sim.energy_residual() - 1.4254529794e-6) < 1e-15
    assert abs(sim.ledger().wall_normal_traction_energy - 1.0219846729e-8) < 1e-18
    # Floor layer stays within 1e-4 (hydrostatic_balance.rs:155-164).
    assert np.abs(e[st.pos[:, 2] < 0.002]).max() < 1e-4


def test_rotated_gravity_golden():
    # docs/hydrostatic-balance.md:228; hydrostatic_balance.rs:339-370
    spec = water_column([0.02, 0.02, 0.02], 0.01, 0.002, 5e-5)
    sim = O.Sim.from_fixture(spec)
    cfg = sim.config()
    cfg.gravity[:] = [GRAVITY, 0.0, 0.0]
    sim.set_config(cfg)
    st = sim.state()
    st.pos = np.stack([0.02 - st.pos[:, 2], st.pos[:, 1], st.pos[:, 0]], 1)
    sim.set_state(st)
    led = sim.ledger()
    led.initial_mechanical_energy = sim.mechanical_energy()
    sim.set_ledger(led)
    sim.advance_to(0.002)
    st = sim.state()
    e = hydrostatic_errors(water(), st.pos, st.volume_ratio, 0.01, depth=np.maximum(st.pos[:, 0] - 0.01, 0.0))
    assert abs(np.abs(e).max() - 0.006727074734587234) < 1e-12
    led = sim.ledger()
    assert led.wall_impulse[0] < 0.0 and abs(led.wall_impulse[2]) < 1e-15
    assert np.linalg.norm(sim.momentum_residual()) < 1e-12


def test_free_fall_block_kinematics():
    # solver.rs:1212-1250
    m = O.paste(1000.0, 1.0e4, 0.3, 1.0e6, 0.0, 1.0)
    sub = 0.001
    g = np.arange(8) * sub + 0.5 * sub
    pos = np.stack(np.meshgrid(g + 0.006, g + 0.006, g + 0.028, indexing="ij"), -1).reshape(-1, 3)
    n = len(pos)
    sim = O.Sim.create(m, O.config(max_dt=1e-3, min_dt=1e-9), [0.02, 0.02, 0.04], 0.002, pos, np.zeros((n, 3)),
                       np.full(n, 1000.0 * sub ** 3), np.full(n, sub ** 3))
    z0 = pos[:, 2].mean()
    sim.advance_to(0.03)
    st = sim.state()
    drop = z0 - st.pos[:, 2].mean()
    expected = 0.5 * 9.81 * 0.03 ** 2
    assert abs(drop - expected) < 0.03 * expected
    assert np.linalg.norm(sim.momentum_residual()) < 1e-12
    assert abs(sim.mass_residual()) < 1e-18
    assert sim.history().rejected_steps == 0
    assert np.abs(st.kirchhoff).max() < 1e-6


def test_traction_energy_equals_full_scan_jump():
    # solver.rs:1429-1476 on the resting paste column (Hydrostatic fixture, paste)
    m = O.paste(1000.0, 1e5, 0.3, 1e4, 1.0, 1.0)
    cfg = O.config(wall_friction=0.4, cfl=0.3, max_dt=4e-5, min_dt=1e-9)
    spec = O.fixture("hydrostatic", m, 0.002, [0.02] * 3, [0.02, 0.02, 0.01], cfg, seed=1)
    sim = O.Sim.from_fixture(spec)
    dt = 4e-5
    total = 0.0
    for _ in range(6):
        _, energy, jump = sim.probe_transfer(dt)
        assert abs(energy - jump) <= 1e-12 * max(abs(jump), 1e-30)
        assert energy > 0.0
        total += energy
        assert sim.try_step(dt) == 0
    assert sim.ledger().wall_normal_traction_energy == total


def test_sliding_block_obeys_coulomb_total_reaction():
    # core/tests/wall_friction.rs:55-70; docs/hydrostatic-balance.md (I = (-7.850e-7, 0, 1.9624e-6))
    _, sim = prestressed_block(0.4, (1.0, 0.0, 0.0))
    sim.advance_to(5e-5)
    imp = np.array(sim.ledger().wall_impulse[:])
    assert imp[2] > 0.0
    assert abs(np.linalg.norm(imp[:2]) / (0.4 * imp[2]) - 1.0) < 1e-12
    assert abs(imp[0] - -7.850e-7) < 5e-10 and abs(imp[2] - 1.9624e-6) < 5e-10


def test_rejected_step_preserves_state():
    # solver.rs:1713-1740
    m = O.paste(1000.0, 1.0e5, 0.3, 1.0e6, 0.0, 1.0)
    sub = 0.001
    gx = np.arange(8) * sub + 0.5 * sub
    gz = np.arange(6) * sub + 0.5 * sub
    pos = np.stack(np.meshgrid(gx + 0.006, gx + 0.006, gz, indexing="ij"), -1).reshape(-1, 3)
    n = len(pos)
    vel = np.tile([1.0, 0.0, -1.0], (n, 1))
    sim = O.Sim.create(m, O.config(max_dt=1e-3, min_dt=1e-9), [0.02] * 3, 0.002, pos, vel,
                       np.full(n, 1000.0 * sub ** 3), np.full(n, sub ** 3))
    sim.advance_to(2e-4)
    before = sim.state()
    led = bytes(sim.ledger())
    rc = sim.try_step(5e-3)
    assert rc in (1, 2, 3, 4, 5, 7)
    after = sim.state()
    assert bytes(sim.ledger()) == led
    assert np.array_equal(after.pos, before.pos) and np.array_equal(after.kirchhoff, before.kirchhoff)


def test_pellet_geometry_and_signed_distance():
    # rigid.rs:452-479
    p = O.new_pellet(0.003, 0.012, 1100.0, [0.0, 0.0, 0.0])
    assert abs(p.mass - 1100.0 * math.pi * 0.003 ** 2 * 0.012) < 1e-12
    offs = list(p.sphere_offsets[:p.sphere_count])
    assert len(offs) >= 3 and all(b - a <= 0.003 + 1e-12 for a, b in zip(offs, offs[1:]))
    phi, normal = np.zeros(1), np.zeros(3)
    O.lib().oracle_pellet_signed_distance(p, O.ptr(np.array([0.0, 0.0, 0.004])), O.ptr(phi), O.ptr(normal))
    assert abs(phi[0] - 0.001) < 1e-12 and np.linalg.norm(normal - [0, 0, 1]) < 1e-12
    # The pellets of the 1.5M-particle bed: r = 1.5 mm, L = 8 mm gives five spheres.
    q = O.new_pellet(0.0015, 0.008, 650.0, [0.0, 0.0, 0.0])
    assert q.sphere_count == 5


@pytest.mark.slow
def test_research_slump_ledgers_golden():
    # docs/validation.md:232-236 (slump, 500 ms): traction -9.7731734973e-6 J, residual 7.3769595864e-5 J
    spec, duration, interval = research_slump()
    sim = O.Sim.from_fixture(spec)
    record_loop(sim, duration, interval)
    assert sim.history().step_count == 5000
    assert abs(sim.ledger().wall_normal_traction_energy - -9.7731734973e-6) < 1e-15
    assert abs(sim.energy_residual() - 7.3769595864e-5) < 1e-14


@pytest.mark.slow
def test_research_coupled_patch_ledgers_golden():
    # docs/validation.md:232-248 (coupled pellet, 50 ms): traction -3.4963973432e-6 J,
    # residual -7.1239144793e-6 J; 43 accepted steps book -4.5405977455e-9 J.
    spec, duration, interval = research_coupled_patch()
    sim = O.Sim.from_fixture(spec)
    sim.advance_to(43 * 5e-5)
    assert sim.history().step_count == 43
    assert abs(sim.ledger().wall_normal_traction_energy - -4.5405977455e-9) < 1e-18
    sim = O.Sim.from_fixture(spec)
    record_loop(sim, duration, interval)
    assert abs(sim.ledger().wall_normal_traction_energy - -3.4963973432e-6) < 1e-15
    assert abs(sim.energy_residual() - -7.1239144793e-6) < 1e-15
