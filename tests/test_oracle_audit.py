"""The energy audit of the oracle (core/src/mpm/audit.rs) against the analytic
and variance references of core/src/mpm/audit_tests.rs.

Every check here is fixed by mathematics: kinetic energies of chosen
velocities, exact APIC variance identities recomputed from the stencil, the
semi-implicit free-fall defect -m g^2 dt^2 / 2, and stage changes that must
equal the ledger increments of the same trial.
"""
import math

import numpy as np
import pytest

import oracle as O


def close(actual, expected, tol=2e-12):
    assert abs(actual - expected) < tol, f"actual={actual:e}, expected={expected:e}"


def isolated(material=None, particles=(((0.47, 0.52, 0.55), (0.0, 0.0, 0.0), 2.0, 0.002),), gravity=(0.0, 0.0, 0.0)):
    """audit_tests.rs:12-35: one interior particle with full quadratic support."""
    m = material if material is not None else O.paste(1000.0, 1000.0, 0.3, 1e6, 0.0, 1.0)
    cfg = O.config(gravity=gravity, wall_friction=0.4, cfl=0.3, max_dt=1e-3, min_dt=1e-9, stiffness=1000.0,
                   restitution=0.2, contact_friction=0.4)
    pos = np.array([p[0] for p in particles], dtype=np.float64)
    vel = np.array([p[1] for p in particles], dtype=np.float64)
    mass = np.array([p[2] for p in particles])
    vol = np.array([p[3] for p in particles])
    return O.Sim.create(m, cfg, [1.0, 1.0, 1.0], 0.1, pos, vel, mass, vol, max_nodes=10000)


def edit(sim, **fields):
    st = sim.state()
    for k, v in fields.items():
        getattr(st, k)[...] = v
    sim.set_state(st)
    return st


def snapshot(sim):
    st = sim.state()
    return (tuple(np.copy(getattr(st, f)) for f in ("id", "pos", "vel", "affine", "deformation", "volume_ratio",
                                                       "kirchhoff", "plastic", "mass", "volume0")),
            bytes(sim.ledger()), bytes(sim.history()), [bytes(p) for p in sim.pellets()])


def unchanged(sim, snap):
    now = snapshot(sim)
    return all(np.array_equal(a, b) for a, b in zip(now[0], snap[0])) and now[1:] == snap[1:]


def test_translation_and_affine_norm_are_analytic_and_nonmutating():
    sim = isolated()
    C = np.array([[0.2, 0.3, -0.1], [0.5, -0.4, 0.7], [0.1, 0.2, 0.4]])
    edit(sim, vel=[[1.0, -2.0, 0.5]], affine=[C])
    snap = snapshot(sim)
    a = sim.audit_step(1e-4)
    close(a["before"]["translation"], 5.25)
    expected_affine = 0.0025 * (0.04 + 0.09 + 0.01 + 0.25 + 0.16 + 0.49 + 0.01 + 0.04 + 0.16)
    close(a["before"]["affine"], expected_affine)
    close(a["stages"]["grid_transport"], 5.25 + expected_affine)
    close(a["after"]["translation"], 5.25)
    close(a["after"]["affine"], expected_affine)
    close(a["terms"]["p2g_transfer_loss"], 0.0)
    close(a["terms"]["g2p_transfer_loss"], 0.0)
    close(a["terms"]["telescoping_error"], 0.0)
    assert unchanged(sim, snap)
    close(sim.kinetic_energy(), a["before"]["translation"])


def test_colliding_velocities_have_exact_variance_loss():
    p = (0.47, 0.52, 0.55)
    sim = isolated(particles=((p, (1.0, 0.0, 0.0), 2.0, 0.002), (p, (-1.0, 0.0, 0.0), 2.0, 0.002)))
    a = sim.audit_step(1e-4)
    close(a["before"]["translation"], 2.0)
    close(a["stages"]["grid_transport"], 0.0)
    close(a["terms"]["p2g_transfer_loss"], 2.0)
    close(a["terms"]["g2p_transfer_loss"], 0.0)


def stencil(x, dx):
    """Quadratic B-spline weights and node offsets per axis (grid.rs:151-183), unpadded node index."""
    base, w, d = [], [], []
    for a in range(3):
        s = x[a] / dx
        b = math.floor(s - 0.5)
        f = s - b
        base.append(b)
        w.append([0.5 * (1.5 - f) ** 2, 0.75 - (f - 1.0) ** 2, 0.5 * (f - 0.5) ** 2])
        d.append([(o - f) * dx for o in range(3)])
    return base, w, d


def test_transfers_match_independent_stencil_variance_identities():
    C0 = np.eye(3) * 0.3
    C1 = np.array([[0.1, -0.2, 0.3], [0.4, 0.2, -0.1], [0.0, 0.3, -0.2]])
    sim = isolated(particles=(((0.47, 0.52, 0.55), (0.7, -0.2, 0.3), 2.0, 0.002),
                              ((0.51, 0.49, 0.58), (-0.4, 0.8, -0.1), 3.0, 0.003)))
    edit(sim, affine=[C0, C1])
    before = sim.state()
    a = sim.audit_step(1e-4)
    trial = sim.clone()
    assert trial.try_step(1e-4) == 0
    after = trial.state()
    dx = 0.1
    # No forces act, so the final grid velocity is the APIC transport velocity
    # sum w m (v + C d) / sum w m, rebuilt here node by node.
    nodes = {}
    for p in range(2):
        base, w, d = stencil(before.pos[p], dx)
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    key = (base[0] + i, base[1] + j, base[2] + k)
                    ww = w[0][i] * w[1][j] * w[2][k]
                    dd = np.array([d[0][i], d[1][j], d[2][k]])
                    m, mom = nodes.get(key, (0.0, np.zeros(3)))
                    nodes[key] = (m + ww * before.mass[p], mom + ww * before.mass[p] * (before.vel[p] + before.affine[p] @ dd))
    p2g = g2p = 0.0
    for p in range(2):
        base, w, d = stencil(before.pos[p], dx)
        q = int(np.where(after.id == before.id[p])[0][0])
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    key = (base[0] + i, base[1] + j, base[2] + k)
                    ww = w[0][i] * w[1][j] * w[2][k]
                    dd = np.array([d[0][i], d[1][j], d[2][k]])
                    m, mom = nodes[key]
                    vg = mom / m
                    old = before.vel[p] + before.affine[p] @ dd
                    rec = after.vel[q] + after.affine[q] @ dd
                    half = 0.5 * before.mass[p] * ww
                    p2g += half * np.sum((old - vg) ** 2)
                    g2p += half * np.sum((vg - rec) ** 2)
    assert p2g > 0.0 and g2p > 0.0
    close(a["terms"]["p2g_transfer_loss"], p2g)
    close(a["terms"]["g2p_transfer_loss"], g2p)


def test_freefall_has_semiimplicit_gravity_potential_defect():
    sim = isolated(gravity=(0.0, 0.0, -9.81))
    dt = 1e-3
    whole = sim.audit_step(dt)
    half = sim.audit_step(dt / 2.0)
    expected = -0.5 * 2.0 * 9.81 ** 2 * dt * dt
    close(whole["terms"]["uncoupled_gravity_potential_mismatch"], expected)
    close(half["terms"]["uncoupled_gravity_potential_mismatch"], expected / 4.0)
    assert whole["terms"]["coupled_grid_gravity_total_potential_change"] is None


def test_resting_and_prestressed_material_expose_storage_mismatch():
    rest = isolated().audit_step(1e-4)
    close(rest["terms"]["augmented_energy_change"], 0.0)
    liquid = O.liquid(1000.0, 1000.0, 0.3, 0.0)
    sim = isolated(material=liquid)
    edit(sim, volume_ratio=[0.95], kirchhoff=[np.eye(3) * (liquid.bulk_modulus * (0.95 - 1.0))])
    stressed = sim.audit_step(1e-4)
    assert stressed["before"]["elastic"] > 0.0
    assert stressed["stages"]["grid_after_stress"] > stressed["stages"]["grid_transport"]
    assert abs(stressed["terms"]["stress_storage_plastic_mismatch"]) > 1e-10
    close(stressed["terms"]["telescoping_error"], 0.0)


def test_signed_wall_stages_match_existing_histories():
    sim = isolated(gravity=(0.0, 0.0, -9.81))
    edit(sim, pos=[[0.47, 0.52, 0.025]], vel=[[0.1, 0.0, -0.5]], kirchhoff=[np.eye(3) * -1000.0])
    r = sim.audit_step(1e-4)
    inc, terms = r["increments"], r["terms"]
    assert inc["wall_normal_projection_energy"] < 0.0
    assert inc["wall_friction_dissipation"] > 0.0
    assert inc["wall_normal_traction_energy"] < 0.0
    close(terms["traction_change"], inc["wall_normal_traction_energy"])
    close(terms["walls_change"], inc["wall_normal_projection_energy"] - inc["wall_friction_dissipation"])
    close(terms["telescoping_error"], 0.0)
    edit(sim, vel=[[0.0, 0.0, 0.0]])
    resting = sim.audit_step(1e-4)
    assert resting["increments"]["wall_normal_traction_energy"] > 0.0
    close(resting["terms"]["traction_change"], resting["increments"]["wall_normal_traction_energy"])


@pytest.mark.parametrize("density", [1.0, 500.0])
def test_coupled_stages_are_combined_not_fake_gravity_work(density):
    sim = isolated(particles=(((0.5, 0.5, 0.5), (0.0, 0.0, 1.0), 2.0, 0.002),), gravity=(0.0, 0.0, -9.81))
    sim.add_pellet(O.new_pellet(0.15, 0.3, density, [0.5, 0.5, 0.65]))
    r = sim.audit_step(1e-4)
    terms, inc = r["terms"], r["increments"]
    assert terms["uncoupled_gravity_potential_mismatch"] is None
    assert terms["coupled_grid_gravity_total_potential_change"] is not None
    assert inc["coupling_grid_energy"] < 0.0
    assert inc["coupling_pellet_energy"] > 0.0
    close(terms["coupling_change"], inc["coupling_grid_energy"] + inc["coupling_pellet_energy"])
    close(terms["pellet_integrator_combined_change"], r["after"]["pellet"] - r["stages"]["pellet_after_coupling"])
    close(terms["telescoping_error"], 0.0)


def test_invalid_unstable_escaped_and_rejected_trials_do_not_mutate():
    sim = isolated()
    snap = snapshot(sim)
    for dt in (0.0, -1.0, float("nan"), float("inf"), 1e-12, 1.0):
        with pytest.raises(ValueError):
            sim.audit_step(dt)
        assert unchanged(sim, snap)
    limited = isolated()
    h = limited.history()
    h.dt_scale = 0.001
    limited.set_history(h)
    with pytest.raises(ValueError):
        limited.audit_step(1e-4)
    for fields in ({"affine": [np.eye(3) * 1e200]}, {"pos": [[-10.0, 0.52, 0.55]]},
                   {"kirchhoff": [np.eye(3) * 1e20]}):
        s = isolated()
        edit(s, **fields)
        snap = snapshot(s)
        with pytest.raises(ValueError):
            s.audit_step(1e-4)
        assert unchanged(s, snap)


def plastic_fixture(history):
    sim = isolated(material=O.paste(1000.0, 1000.0, 0.3, 0.1, 0.0, 1.0))
    C = np.zeros((3, 3))
    C[0, 1] = 50.0
    edit(sim, affine=[C])
    led = sim.ledger()
    led.plastic_dissipation = history
    sim.set_ledger(led)
    return sim


def test_plastic_increment_is_history_difference_not_total_or_invented_dissipation():
    sim = plastic_fixture(3.0)
    a = sim.audit_step(1e-3)
    sim.advance_to(1e-3)
    assert a["increments"]["plastic_dissipation"] > 0.0
    close(a["increments"]["plastic_dissipation"], sim.ledger().plastic_dissipation - 3.0)
    expected = (a["stages"]["grid_after_stress"] - a["stages"]["grid_transport"] + a["after"]["elastic"]
                - a["before"]["elastic"] + a["stages"]["plastic_dissipation"])
    close(a["terms"]["stress_storage_plastic_mismatch"], expected)
    close(a["terms"]["telescoping_error"], 0.0)


def test_local_plastic_work_survives_cumulative_history_rounding():
    sim = plastic_fixture(1e100)
    snap = snapshot(sim)
    a = sim.audit_step(1e-3)
    assert a["stages"]["plastic_dissipation"] > 0.0
    assert np.float64(a["increments"]["plastic_dissipation"]).view(np.uint64) == np.float64(0.0).view(np.uint64)
    close(a["terms"]["stress_storage_plastic_mismatch"],
          a["stages"]["grid_after_stress"] - a["stages"]["grid_transport"] + a["after"]["elastic"]
          - a["before"]["elastic"] + a["stages"]["plastic_dissipation"])
    assert unchanged(sim, snap)


def test_rigid_gravity_contact_rotation_stay_one_combined_term():
    sim = isolated(gravity=(0.0, 0.0, -9.81))
    # UnitQuaternion::from_euler_angles(0, 0.03, 0): a rotation of 0.03 rad about y.
    q = (0.0, math.sin(0.015), 0.0, math.cos(0.015))
    p = O.new_pellet(0.1, 0.3, 500.0, [0.5, 0.5, 0.095], q)
    p.velocity[:] = [0.1, 0.0, -0.2]
    p.angular_momentum[:] = [0.01, 0.02, 0.03]
    sim.add_pellet(p)
    snap = snapshot(sim)
    a = sim.audit_step(1e-4)
    assert abs(a["increments"]["pellet_wall_work"]) > 0.0
    assert a["terms"]["uncoupled_gravity_potential_mismatch"] is None
    assert abs(a["terms"]["pellet_integrator_combined_change"]) > 0.0
    assert abs(a["terms"]["pellet_integrator_combined_change"] - a["increments"]["pellet_wall_work"]) > 1e-9
    close(a["terms"]["telescoping_error"], 0.0)
    assert unchanged(sim, snap)


def test_audit_matches_actual_advance_endpoint_and_existing_histories():
    sim = isolated(gravity=(0.0, 0.0, -9.81))
    edit(sim, vel=[[0.1, -0.2, 0.3]])
    a = sim.audit_step(1e-4)
    sim.advance_to(1e-4)
    st = sim.state()
    close(a["after"]["translation"], float(np.sum(0.5 * st.mass * np.sum(st.vel ** 2, axis=1))))
    close(a["after"]["elastic"], sim.elastic_energy())
    close(a["after"]["potential"], sim.potential_energy())
    close(a["increments"]["plastic_dissipation"], sim.ledger().plastic_dissipation)
