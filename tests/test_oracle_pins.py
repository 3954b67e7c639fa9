"""Pins of the oracle's extensions beyond the reference: the per-face wall law
(free-slip and sticky faces) and several pellets coupled to one grid.

The per-face law reduces to the reference's single Coulomb coefficient
(solver.rs:822-892, 1145-1155); the reference's own wall tests
(core/tests/wall_friction.rs) are restated on the oracle first, then each
face's coefficient is shown to act on that face only, and the two limits of
the benchmark workloads (mu = 0, mu = inf) are checked against the closed
forms they reduce to at the grid.

Several pellets apply couple_grid (rigid.rs:385-438) once per pellet in index
order.  Their coupling is pinned by what the mathematics fixes: exact exchange
of linear and angular momentum between grid and pellets (the reference's
rigid.rs:482-557 for one pellet), pellets with disjoint boxes acting exactly
as each alone, the sequential composition of single-pellet couplings, and the
order dependence where two pellets constrain the same node.
"""
import math

import numpy as np
import pytest

import cases
import oracle as O

DT = 5e-5


def _block(friction, velocity, faces=None):
    spec, sim = cases.prestressed_block(friction, velocity)
    if faces is not None:
        cfg = sim.config()
        cfg.face_friction[:] = list(faces)
        sim.set_config(cfg)
    return spec, sim


# ---------------------------------------------------------------- wall_friction.rs restated

def test_zero_friction_preserves_even_a_tiny_tangential_velocity():
    # core/tests/wall_friction.rs:172-188
    speed = 1e-16
    _, sim = _block(0.0, (speed, 0.0, 0.0))
    sim.advance_to(DT)
    led = sim.ledger()
    assert math.copysign(1.0, led.wall_impulse[0]) > 0 and led.wall_impulse[0] == 0.0
    assert math.copysign(1.0, led.wall_friction_dissipation) > 0 and led.wall_friction_dissipation == 0.0
    assert np.all(np.abs(sim.state().vel[:, 0] - speed) < 1e-29)


def test_slow_block_sticks_within_the_coulomb_budget():
    # core/tests/wall_friction.rs:73-107
    speed = 1e-5
    spec, sim = _block(0.4, (speed, 0.0, 0.0))
    sim.advance_to(DT)
    imp = np.array(sim.ledger().wall_impulse[:])
    ratio = np.linalg.norm(imp[:2]) / (0.4 * imp[2])
    st = sim.state()
    floor = st.pos[:, 2] < 0.5 * spec.grid_spacing
    floor_speed = np.linalg.norm(st.vel[floor, :2], axis=1).max()
    assert imp[2] > 0.0
    assert imp[0] < 0.0 and ratio < 0.5, ratio
    assert floor_speed < 1e-15, floor_speed
    assert np.all(st.vel[:, 0] >= -1e-15)
    assert sim.ledger().wall_friction_dissipation > 0.0


def test_release_with_friction_receives_no_normal_or_tangential_impulse():
    # core/tests/wall_friction.rs:109-135
    _, sim = _block(0.4, (1.0, 0.0, 0.3))
    st = sim.state()
    st.kirchhoff[:] = 0.0
    st.deformation[:] = np.eye(3)
    sim.set_state(st)
    led = sim.ledger()
    led.initial_mechanical_energy = sim.mechanical_energy()
    sim.set_ledger(led)
    sim.advance_to(DT)
    expected = np.array([1.0, 0.0, 0.3 - 9.81 * DT])
    defect = np.linalg.norm(sim.state().vel - expected, axis=1).max()
    assert defect < 1e-12, defect
    assert list(sim.ledger().wall_impulse) == [0.0, 0.0, 0.0]
    assert sim.ledger().wall_friction_dissipation <= 0.0


def test_vanishing_contact_does_not_turn_negative_image_weight_into_attraction():
    # core/tests/wall_friction.rs:137-170
    inward_speed = 1e-7
    spec, sim = _block(0.4, (1.0, 0.0, 9.81 * DT - inward_speed))
    st = sim.state()
    st.kirchhoff[:] = 0.0
    st.deformation[:] = np.eye(3)
    low = st.pos[:, 2] < 0.5 * spec.grid_spacing
    st.pos[low, 2] = 0.05 * spec.grid_spacing
    sim.set_state(st)
    led = sim.ledger()
    led.initial_mechanical_energy = sim.mechanical_energy()
    sim.set_ledger(led)
    # The step's P2G grid: nodes on the wall side along z (grid.rs:135-143).
    mass, _, _, _ = sim.probe_update(DT)
    _, cells = sim.layout()
    nodes = [c + 3 for c in cells]
    k = np.arange(mass.size) % nodes[2]
    wall_side = (k <= 1) | (k > cells[2])
    constrained_mass = mass[wall_side].sum()
    sim.advance_to(DT)
    expected_normal = constrained_mass * inward_speed
    imp = sim.ledger().wall_impulse
    assert imp[2] >= 0.0
    assert abs(imp[2] - expected_normal) < 1e-16
    assert abs(imp[0] + 0.4 * expected_normal) < 1e-16
    assert np.linalg.norm(sim.momentum_residual()) < 1e-15


# ---------------------------------------------------------------- per-face coefficients

def _block_on_face(axis, high, faces, velocity=(1.0, 0.0, 0.0)):
    """The prestressed block of wall_friction.rs:12-52 turned so that it rests on
    face (axis, high) under gravity pointing into that face, sliding along the
    face's first tangential axis."""
    spec, floor = _block(0.4, velocity)
    others = [b for b in range(3) if b != axis]
    Q = np.zeros((3, 3))
    Q[others[0], 0] = 1.0
    Q[others[1], 1] = 1.0
    Q[axis, 2] = -1.0 if high else 1.0
    dom0 = list(spec.domain)
    domain = [0.0] * 3
    domain[others[0]], domain[others[1]], domain[axis] = dom0[0], dom0[1], dom0[2]
    offset = np.zeros(3)
    if high:
        offset[axis] = dom0[2]
    st = floor.state()
    cfg = O.config(gravity=tuple(Q @ np.array([0.0, 0.0, -9.81])), wall_friction=0.4, cfl=0.3, max_dt=DT,
                   min_dt=1e-9, contact_friction=0.4, face_friction=faces)
    sim = O.Sim.create(spec.material, cfg, domain, spec.grid_spacing, st.pos @ Q.T + offset, st.vel @ Q.T,
                       st.mass, st.volume0)
    turn = lambda m: np.einsum("ab,pbc,dc->pad", Q, m, Q)
    st.pos = st.pos @ Q.T + offset
    st.vel = st.vel @ Q.T
    st.affine = turn(st.affine)
    st.deformation = turn(st.deformation)
    st.kirchhoff = turn(st.kirchhoff)
    sim.set_state(st)
    return sim, Q


def _face_run(axis, high, faces):
    sim, Q = _block_on_face(axis, high, faces)
    assert sim.try_step(DT) == 0
    return sim, Q


@pytest.mark.parametrize("axis,high", [(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1)])
def test_face_coefficient_acts_on_its_own_face(axis, high):
    # face_friction index 2 axis + (low face 0, high face 1), (x-, x+, y-, y+, z-, z+):
    # a block resting on one face feels that face's coefficient and no other.
    # A wrong index would put 99 (or 0) on the contact face and change every bit.
    face = 2 * axis + high
    uniform, Q = _face_run(axis, high, [0.4] * 6)
    only = [99.0] * 6
    only[face] = 0.4
    own, _ = _face_run(axis, high, only)
    assert bytes(own.ledger()) == bytes(uniform.ledger())
    a, b = own.state(), uniform.state()
    assert np.array_equal(a.pos, b.pos) and np.array_equal(a.vel, b.vel) and np.array_equal(a.kirchhoff, b.kirchhoff)
    # The contact face's coefficient is the one that matters: free slip there
    # leaves no tangential impulse at all, and the uniform law is the floor
    # run turned (the reference's sliding impulse, to round-off).
    slip = [0.4] * 6
    slip[face] = 0.0
    free, _ = _face_run(axis, high, slip)
    imp = Q.T @ np.array(free.ledger().wall_impulse[:])
    assert imp[0] == 0.0 and imp[1] == 0.0 and free.ledger().wall_friction_dissipation == 0.0
    ref = Q.T @ np.array(uniform.ledger().wall_impulse[:])
    assert abs(ref[0] - -7.850e-7) < 5e-10 and abs(ref[2] - 1.9624e-6) < 5e-10


def test_sticky_floor_stops_every_loaded_floor_node():
    # mu = inf on the floor (z-): every floor node that approaches the wall, and
    # every node that receives a wall-traction reaction, keeps no tangential
    # velocity (solver.rs:1145-1155 with an infinite reduction); nothing else
    # changes.  The tangential wall impulse is then -sum m v_t over those nodes
    # and the friction dissipation sum 1/2 m |v_t|^2, v_t the velocity the walls
    # act on (solver.rs:833).
    faces = [0.4, 0.4, 0.4, 0.4, math.inf, 0.4]
    _, sim = _block(0.4, (1.0, 0.0, 0.0), faces)
    mass, before, after, budget = sim.probe_update(DT)
    _, cells = sim.layout()
    nodes = [c + 3 for c in cells]
    k = np.arange(mass.size) % nodes[2]
    live = mass > 0.0
    floor = live & (k <= 1) & (before[:, 2] < 0.0)
    loaded = live & (budget[:, 2] > 0.0)
    stopped = floor | loaded
    assert floor.sum() > 0 and loaded.sum() > 0 and not np.any(floor & loaded)
    assert np.all(after[stopped, :2] == 0.0)
    assert np.all(after[floor, 2] == 0.0)
    untouched = live & ~stopped & (k > 1)
    assert np.array_equal(after[untouched], before[untouched])
    assert np.array_equal(after[loaded, 2], before[loaded, 2])
    expect_impulse = -(mass[stopped, None] * before[stopped, :2]).sum(0)
    expect_friction = 0.5 * (mass[stopped] * (before[stopped, :2] ** 2).sum(1)).sum()
    assert sim.try_step(DT) == 0
    led = sim.ledger()
    assert abs(led.wall_impulse[0] - expect_impulse[0]) < 1e-12 * abs(expect_impulse[0])
    assert abs(led.wall_impulse[1] - expect_impulse[1]) <= 1e-12 * abs(expect_impulse[0])
    assert abs(led.wall_friction_dissipation - expect_friction) < 1e-12 * expect_friction
    # Harder than the Coulomb cone: the sticky floor takes more tangential impulse.
    _, coulomb = _block(0.4, (1.0, 0.0, 0.0))
    assert coulomb.try_step(DT) == 0
    assert led.wall_impulse[0] < coulomb.ledger().wall_impulse[0] < 0.0


def test_free_slip_floor_leaves_tangential_velocity_untouched():
    faces = [0.4, 0.4, 0.4, 0.4, 0.0, 0.4]
    _, sim = _block(0.4, (1.0, 0.0, 0.0), faces)
    mass, before, after, _ = sim.probe_update(DT)
    live = mass > 0.0
    assert np.array_equal(after[live, :2], before[live, :2])
    assert sim.try_step(DT) == 0
    led = sim.ledger()
    assert led.wall_impulse[0] == 0.0 and led.wall_impulse[1] == 0.0
    assert led.wall_friction_dissipation == 0.0
    assert led.wall_impulse[2] > 0.0


# ---------------------------------------------------------------- several pellets

DOMAIN = [0.02, 0.02, 0.02]
SPACING = 0.001
NODES = 23  # 20 cells + 3 padding nodes per axis


def _pellet(pos, axis_angle=None, radius=0.002, length=0.008, vel=(0.0, 0.0, 0.0), ang=(0.0, 0.0, 0.0)):
    q = (0.0, 0.0, 0.0, 1.0)
    if axis_angle is not None:
        ax, ang_rad = np.asarray(axis_angle[0], float), axis_angle[1]
        ax = ax / np.linalg.norm(ax)
        s = math.sin(0.5 * ang_rad)
        q = (ax[0] * s, ax[1] * s, ax[2] * s, math.cos(0.5 * ang_rad))
    p = O.new_pellet(radius, length, 1100.0, pos, orientation=q)
    p.velocity[:] = list(vel)
    p.angular_momentum[:] = list(ang)
    return p


def _phi(p, x):
    phi, normal = np.zeros(1), np.zeros(3)
    O.lib().oracle_pellet_signed_distance(p, O.ptr(np.ascontiguousarray(x, dtype=np.float64)), O.ptr(phi),
                                          O.ptr(normal))
    return phi[0], normal


def _rotation(q):
    i, j, k, w = q
    return np.array([[1 - 2 * (j * j + k * k), 2 * (i * j - w * k), 2 * (i * k + w * j)],
                     [2 * (i * j + w * k), 1 - 2 * (i * i + k * k), 2 * (j * k - w * i)],
                     [2 * (i * k - w * j), 2 * (j * k + w * i), 1 - 2 * (i * i + j * j)]])


def _point_velocity(p, x):
    r = _rotation(list(p.orientation))
    inv = np.diag(1.0 / np.array(p.inertia_body[:]))
    omega = r @ inv @ r.T @ np.array(p.angular_momentum[:])
    return np.array(p.velocity[:]) + np.cross(omega, np.asarray(x) - np.array(p.position[:]))


def _box(p):
    r = _rotation(list(p.orientation))
    half = max(abs(o) for o in p.sphere_offsets[:p.sphere_count])
    reach = np.abs(r[:, 0]) * half + p.radius
    c = np.array(p.position[:])
    return c - reach, c + reach


def _overlap(a, b):
    (la, ha), (lb, hb) = _box(a), _box(b)
    return bool(np.all(la <= hb) and np.all(lb <= ha))


def _nodes(pellets, seed=0):
    """Nodes within 1 mm of any pellet, carrying an off-centre upward flow with a
    random part (rigid.rs:482-503 generalised), in dense index order."""
    rng = np.random.default_rng(seed)
    idx, pos = [], []
    for i in range(NODES):
        for j in range(NODES):
            for k in range(NODES):
                x = np.array([(i - 1.0) * SPACING, (j - 1.0) * SPACING, (k - 1.0) * SPACING])
                if min(_phi(p, x)[0] for p in pellets) <= 0.001:
                    idx.append((i * NODES + j) * NODES + k)
                    pos.append(x)
    pos = np.array(pos)
    vel = np.stack([0.2 * (pos[:, 0] - 0.01) / 0.006, np.full(len(pos), 0.05), np.full(len(pos), 0.3)], 1)
    vel += 0.05 * rng.standard_normal(vel.shape)
    mass = 1e-6 * (1.0 + 0.5 * rng.random(len(pos)))
    return np.array(idx), pos, mass, vel


def _couple(pellets, idx, mass, vel):
    return O.couple_nodes(DOMAIN, SPACING, idx, mass, vel, pellets, 0.4)


def _three():
    # a and b move towards each other, so a node inside both is pushed apart by
    # the first and then constrained again by the second.
    towards = np.array([0.001, -0.001, 0.002]) / math.sqrt(6e-6)
    a = _pellet([0.010, 0.010, 0.010], vel=tuple(0.1 * towards), ang=(0.0, 1e-9, 0.0))
    b = _pellet([0.011, 0.009, 0.012], axis_angle=([0, 0, 1], 0.5 * math.pi), vel=tuple(-0.1 * towards),
                ang=(2e-10, 0.0, 1e-10))
    c = _pellet([0.004, 0.016, 0.004], axis_angle=([1, 1, 0], 0.7), vel=(0.0, -0.02, 0.1))
    assert _overlap(a, b) and not _overlap(a, c) and not _overlap(b, c)
    return [a, b, c]


def test_several_pellets_conserve_linear_and_angular_momentum():
    # rigid.rs:482-557 generalised to three pellets (two sharing nodes): the grid
    # loses exactly what the pellets gain, linear and angular about the origin.
    pellets = _three()
    idx, pos, mass, vel = _nodes(pellets)
    new, res = _couple(pellets, idx, mass, vel)
    assert np.all(res[:, 7] > 0)
    dv = (new - vel) * mass[:, None]
    p_before = (vel * mass[:, None]).sum(0) + sum(p.mass * np.array(p.velocity[:]) for p in pellets)
    l_before = np.cross(pos, vel * mass[:, None]).sum(0) + sum(
        np.array(p.angular_momentum[:]) + np.cross(p.position[:], p.mass * np.array(p.velocity[:])) for p in pellets)
    dp = dv.sum(0) + res[:, 0:3].sum(0)
    dl = np.cross(pos, dv).sum(0) + sum(res[k, 3:6] + np.cross(pellets[k].position[:], res[k, 0:3])
                                        for k in range(len(pellets)))
    assert np.linalg.norm(dp) < 1e-12 * np.linalg.norm(p_before), dp
    assert np.linalg.norm(dl) < 1e-12 * np.linalg.norm(l_before), dl
    assert np.linalg.norm(res[:, 0:3], axis=1).min() > 0.0


def test_disjoint_pellets_act_exactly_as_each_alone():
    a, _, c = _three()
    idx, pos, mass, vel = _nodes([a, c], seed=1)
    both, res = _couple([a, c], idx, mass, vel)
    only_a, ra = _couple([a], idx, mass, vel)
    only_c, rc = _couple([c], idx, mass, vel)
    assert np.array_equal(res[0], ra[0]) and np.array_equal(res[1], rc[0])
    moved_a = np.any(only_a != vel, axis=1)
    moved_c = np.any(only_c != vel, axis=1)
    assert moved_a.sum() > 0 and moved_c.sum() > 0 and not np.any(moved_a & moved_c)
    expect = vel.copy()
    expect[moved_a] = only_a[moved_a]
    expect[moved_c] = only_c[moved_c]
    assert np.array_equal(both, expect)


def test_pellets_couple_one_after_another_in_index_order():
    # N pellets = the reference's single-pellet couple_grid applied N times, each
    # on the grid the previous one left; where two pellets constrain one node the
    # order matters, and the node ends up not approaching the last of them.
    a, b, _ = _three()
    idx, pos, mass, vel = _nodes([a, b], seed=2)
    ab, rab = _couple([a, b], idx, mass, vel)
    after_a, ra = _couple([a], idx, mass, vel)
    then_b, rb = _couple([b], idx, mass, after_a)
    assert np.array_equal(ab, then_b)
    assert np.array_equal(rab[0], ra[0]) and np.array_equal(rab[1], rb[0])
    ba, rba = _couple([b, a], idx, mass, vel)
    in_a = np.array([_phi(a, x)[0] <= 0.0 for x in pos])
    in_b = np.array([_phi(b, x)[0] <= 0.0 for x in pos])
    shared = in_a & in_b
    assert shared.sum() > 0
    differ = np.any(ab != ba, axis=1)
    assert np.any(differ & shared) and not np.any(differ & ~shared)
    for x, v, inside in zip(pos, ab, in_b):
        if inside:
            _, n = _phi(b, x)
            assert np.dot(v - _point_velocity(b, x), n) >= -1e-12
    for x, v, inside in zip(pos, ba, in_a):
        if inside:
            _, n = _phi(a, x)
            assert np.dot(v - _point_velocity(a, x), n) >= -1e-12


def test_pellet_order_in_a_step_follows_the_coupling_composition():
    # One oracle step with two pellets pressed into a paste bed: the step's
    # coupling impulse is the sum of the per-pellet couplings of the step's grid,
    # composed in index order.
    spec, _, _ = cases.research_coupled_patch()
    sim = O.Sim.from_fixture(spec)
    p0 = sim.pellets()[0]
    p0.position[2] = 0.02 + p0.radius - 0.001
    p0.velocity[:] = [0.0, 0.0, -0.2]
    sim.set_pellet(0, p0)
    p1 = O.new_pellet(0.003, 0.012, 1100.0, [p0.position[0] + 0.004, p0.position[1] + 0.002, p0.position[2]],
                      orientation=(0.0, 0.0, math.sin(0.6), math.cos(0.6)))
    p1.velocity[:] = [0.0, 0.0, -0.15]
    sim.add_pellet(p1)
    assert _overlap(p0, p1)
    mass, _, after, _ = sim.probe_update(DT)
    live = np.nonzero(mass > 0.0)[0]
    spacing, cells = sim.layout()
    domain = [c * spacing for c in cells]
    _, res = O.couple_nodes(domain, spacing, live, mass[live], after[live], [p0, p1],
                            sim.config().wall_friction)
    before = np.array(sim.ledger().coupling_impulse[:])
    assert sim.try_step(DT) == 0
    got = np.array(sim.ledger().coupling_impulse[:]) - before
    assert np.all(res[:, 7] > 0)
    assert np.allclose(got, res[:, 0:3].sum(0), rtol=1e-12, atol=1e-18)
