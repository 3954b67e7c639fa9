"""Parity of the HIP engine against the f64 oracle, through the C ABI.

Tolerances (stated per test) follow the engine's contract: positions and
velocities within 1e-4 relative (normwise: largest error over the largest
magnitude) after 100 substeps, stress and plastic dissipation within 1e-3,
mass and momentum ledgers to f32 round-off.
"""
import math

import numpy as np
import pytest

import cases
import oracle as O

pytestmark = pytest.mark.gpu


def as_engine(E, obj, cls):
    out = cls()
    for name, _ in obj._fields_:
        val = getattr(obj, name)
        if hasattr(val, "__len__"):
            getattr(out, name)[:] = list(val)
        else:
            setattr(out, name, val)
    return out


def engine_from_oracle(E, osim, cfg=None, pellets=True):
    st = osim.state()
    est = E.State(st.id.copy(), st.pos, st.vel, st.affine, st.deformation, st.volume_ratio, st.kirchhoff,
                  st.plastic, st.mass, st.volume0)
    m = as_engine(E, osim_material(osim), E.Material)
    c = as_engine(E, cfg if cfg is not None else osim.config(), E.Config)
    spacing, cells = osim.layout()
    domain = [cells[a] * spacing for a in range(3)]
    pl = [as_engine(E, p, E.Pellet) for p in osim.pellets()] if pellets else []
    return E.Simulation(m, c, domain, spacing, est, pellets=pl)


_materials = {}


def osim_material(osim):
    return _materials[id(osim)]


def build_oracle(spec):
    s = O.Sim.from_fixture(spec)
    _materials[id(s)] = spec.material
    return s


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def compare_states(gpu, ref, pos_tol=1e-4, vel_tol=1e-4, stress_tol=1e-3):
    g = gpu.state().by_id()
    r = ref.state()
    o = np.argsort(r.id)
    assert np.array_equal(g.id, r.id[o])
    ep = rel(g.pos, r.pos[o])
    ev = rel(g.vel, r.vel[o])
    et = rel(g.kirchhoff, r.kirchhoff[o])
    print(f"parity: pos {ep:.2e} vel {ev:.2e} stress {et:.2e}", flush=True)
    assert ep < pos_tol, ep
    assert ev < vel_tol, ev
    assert et < stress_tol, et
    return g, r, o


# ---------------------------------------------------------------- kernels

def test_paste_batch_matches_oracle(engine):
    rng = np.random.default_rng(3)
    m = O.paste(1200.0, 138461.538, 0.384615, 50.0, 5.0, 0.5, shear_units=True)
    em = as_engine(engine, m, engine.Material)
    n = 4096
    scale = np.where(rng.random(n) < 0.5, 2e-4, 4e-3)[:, None, None]
    trial = np.eye(3)[None] + rng.standard_normal((n, 3, 3)) * scale
    trial[0] = np.eye(3) * 0.99  # hydrostatic compression
    dt = 6.7e-5
    f, tau, inc, dd, codes = engine.paste_update_batch(em, dt, trial)
    ref_f = np.zeros_like(f)
    ref_t = np.zeros_like(tau)
    ref_i = np.zeros(n)
    for p in range(n):
        rc, rf, rt, ri, rdd, req = O.update_paste(trial[p], m, dt)
        assert rc == codes[p] == 0
        ref_f[p], ref_t[p], ref_i[p] = rf, rt, ri
    yielded = ref_i > 0
    assert 0.2 < yielded.mean() < 0.8
    # Stress: each particle within 1e-3 of its own stress scale (or of the yield stress).
    denom = np.maximum(np.abs(ref_t).reshape(n, -1).max(1), m.yield_stress)
    err = np.abs(tau - ref_t).reshape(n, -1).max(1) / denom
    print("paste batch stress err", err.max())
    assert err.max() < 1e-3
    assert np.abs(f - ref_f).max() < 2e-6
    ierr = np.abs(inc - ref_i)[yielded] / ref_i[yielded]
    assert ierr.max() < 1e-3


def test_manuscript_return_on_device(engine):
    # core/tests/formal_examples.rs:12-26 through the f32 kernel: q = 78.0059568837,
    # plastic increment 0.001595234493, each within 1e-5 relative.
    m = engine.paste(1000.0, 1000.0, 0.2, 30.0, 10.0, 0.6, shear_units=True)
    strain = 80.0 / (2.0 * math.sqrt(3.0) * m.shear_modulus)
    trial = np.diag([math.exp(strain), math.exp(-strain), 1.0])[None]
    f, tau, inc, dd, codes = engine.paste_update_batch(m, 0.0014, trial)
    assert codes[0] == 0
    dev = tau[0] - np.eye(3) * np.trace(tau[0]) / 3.0
    q = math.sqrt(1.5) * np.linalg.norm(dev)
    assert abs(q - 78.0059568837) / 78.0059568837 < 1e-5
    assert abs(inc[0] - 0.001595234493) / 0.001595234493 < 1e-5
    assert abs(np.linalg.det(f[0]) - 1.0) < 1e-6


@pytest.mark.parametrize("gamma", [1.0, 7.0])
def test_liquid_batch_matches_oracle(engine, gamma):
    rng = np.random.default_rng(5)
    m = O.liquid(1000.0, 3.6e5, 0.2, 1e-3, tait_gamma=gamma)
    em = as_engine(engine, m, engine.Material)
    n = 2048
    j = 1.0 + rng.standard_normal(n) * 2e-3
    grad = rng.standard_normal((n, 3, 3)) * 5.0
    dt = 3e-5
    nj, tau, codes = engine.liquid_update_batch(em, dt, j, grad)
    for p in range(0, n, 7):
        rc, rj, rt, rp = O.update_liquid(j[p], grad[p], m, dt)
        assert rc == codes[p] == 0
        assert abs(nj[p] - rj) < 1e-7
        assert np.abs(tau[p] - rt).max() < 1e-3 * max(np.abs(rt).max(), 1.0)


# ---------------------------------------------------------------- substeps

def fixed_dt_run(E, spec, steps, dt=None):
    osim = build_oracle(spec)
    gsim = engine_from_oracle(E, osim)
    if dt is None:
        dt = min(osim.stability_limit(), spec.config.max_dt)
    for k in range(steps):
        rc = osim.try_step(dt)
        rep = gsim.try_step(dt)
        assert rc == 0 and rep.accepted == 1, (k, rc, rep.failure)
    return osim, gsim, dt


def test_paste_column_parity_100_substeps(engine):
    # 49,152-particle Herschel-Bulkley column (32^3 grid, dx 2.5 mm), 100 substeps.
    spec = cases.paste_column_32()
    osim, gsim, dt = fixed_dt_run(engine, spec, 100)
    g, r, o = compare_states(gsim, osim)
    lo, lg = osim.ledger(), gsim.ledger()
    assert lo.plastic_dissipation > 0.0
    assert abs(lg.plastic_dissipation - lo.plastic_dissipation) / lo.plastic_dissipation < 1e-3
    assert abs(gsim.mass_residual()) <= 1e-7 * lo.initial_mass
    scale = np.linalg.norm(lo.gravity_impulse)
    assert np.linalg.norm(gsim.momentum_residual()) < 1e-5 * scale
    for a in range(3):
        assert abs(lg.wall_impulse[a] - lo.wall_impulse[a]) < 1e-3 * scale
    # Plastic strain of yielded particles.
    eps_err = np.abs(g.plastic - r.plastic[o]).max() / max(r.plastic.max(), 1e-30)
    assert eps_err < 1e-3


def test_hydrostatic_column_matches_golden(engine):
    # docs/hydrostatic-balance.md:26-31: max 0.011513856585915002, rms 0.0034430221485745437
    spec = cases.water_column([0.02] * 3, 0.01, 0.002, 5e-5)
    osim = build_oracle(spec)
    gsim = engine_from_oracle(engine, osim)
    gsim.advance_to(0.01)
    h = gsim.history()
    assert h.step_count == 200 and h.rejected_steps == 0 and h.limited_steps == 0
    st = gsim.state()
    e = cases.hydrostatic_errors(cases.water(), st.pos, st.volume_ratio, 0.01)
    mx, rms = np.abs(e).max(), math.sqrt(np.mean(e * e))
    print(f"hydrostatic gpu max {mx!r} rms {rms!r}")
    assert abs(mx - 0.011513856585915002) < 2e-4
    assert abs(rms - 0.0034430221485745437) < 5e-5
    assert np.abs(e[st.pos[:, 2] < 0.002]).max() < 2e-4
    led = gsim.ledger()
    assert abs(gsim.mass_residual()) <= 1e-7 * led.initial_mass
    assert np.linalg.norm(gsim.momentum_residual()) < 1e-5 * np.linalg.norm(led.gravity_impulse)
    assert abs(led.wall_normal_traction_energy - 1.0219846729e-8) / 1.0219846729e-8 < 2e-2


def test_adaptive_advance_matches_oracle(engine):
    spec = cases.paste_column_32(max_dt=1e-3)
    osim = build_oracle(spec)
    gsim = engine_from_oracle(engine, osim)
    osim.advance_to(4e-3)
    gsim.advance_to(4e-3)
    ho, hg = osim.history(), gsim.history()
    assert hg.step_count == ho.step_count and hg.rejected_steps == ho.rejected_steps
    assert abs(hg.min_dt_used - ho.min_dt_used) / ho.min_dt_used < 1e-4
    compare_states(gsim, osim)


def test_rejected_step_leaves_state_untouched(engine):
    spec = cases.paste_column_32()
    osim = build_oracle(spec)
    gsim = engine_from_oracle(engine, osim)
    gsim.advance_to(1e-3)
    before = gsim.state()
    led = bytes(gsim.ledger())
    rep = gsim.try_step(0.05)
    assert rep.accepted == 0 and rep.failure in (1, 2, 3, 4, 5, 7)
    after = gsim.state()
    assert bytes(gsim.ledger()) == led
    assert np.array_equal(after.pos, before.pos) and np.array_equal(after.kirchhoff, before.kirchhoff)
    # The engine recovers: the next regular step is accepted.
    gsim.advance_to(1.5e-3)


def test_sliding_block_reaction_on_the_coulomb_cone(engine):
    # core/tests/wall_friction.rs:55-70: a uniaxially prestressed paste block sliding at
    # 1 m/s on a floor with mu = 0.4.  The total wall reaction lies on the Coulomb cone,
    # |I_t| = mu I_n (1e-12 in f64; the f32 node sums hold it to about 1e-6), with
    # I = (-7.850e-7, 0, 1.9624e-6) N s (docs/hydrostatic-balance.md).
    spec, osim = cases.prestressed_block(0.4, (1.0, 0.0, 0.0))
    _materials[id(osim)] = spec.material
    gsim = engine_from_oracle(engine, osim)
    gsim.advance_to(5e-5)
    osim.advance_to(5e-5)
    imp = np.array(gsim.ledger().wall_impulse[:])
    assert imp[2] > 0.0
    ratio = np.linalg.norm(imp[:2]) / (0.4 * imp[2])
    print(f"coulomb ratio {ratio!r}, impulse {imp}", flush=True)
    assert abs(ratio - 1.0) < 1e-5
    assert abs(imp[0] - -7.850e-7) < 1e-3 * 7.850e-7 and abs(imp[2] - 1.9624e-6) < 1e-3 * 1.9624e-6
    assert gsim.history().step_count == osim.history().step_count
    compare_states(gsim, osim)


@pytest.mark.parametrize("depth", [2, 3])
def test_thin_grid_wall_traction_on_both_faces(engine, depth):
    # A paste block filling a grid `depth` cells deep, compressed along z, so that
    # the wall-image deposits of the floor and of the ceiling both reach the grid
    # (solver.rs:725-819).  Two cells deep, both faces deposit on node 2.
    m = O.paste(1000.0, 1e5, 0.2, 1e6, 1.0, 1.0)
    cfg = O.config(wall_friction=0.4, cfl=0.3, max_dt=5e-5, min_dt=1e-9)
    dx = 0.002
    spec = O.fixture("slump", m, dx, [10 * dx, 10 * dx, depth * dx], [6 * dx, 6 * dx, depth * dx], cfg, seed=7)
    osim = build_oracle(spec)
    st = osim.state()
    st.kirchhoff[:] = 0.0
    st.kirchhoff[:, 2, 2] = -2000.0
    st.deformation[:, 2, 2] = np.exp(-2000.0 / 1e5)
    osim.set_state(st)
    gsim = engine_from_oracle(engine, osim)
    dt = min(osim.stability_limit(), cfg.max_dt)
    for k in range(30):
        assert osim.try_step(dt) == 0, k
        rep = gsim.try_step(dt)
        assert rep.accepted == 1, (k, rep.failure)
    # The block sits near equilibrium, so its velocities (about 1.5e-4 m/s) are the
    # small residual of cancelling stress, traction and gravity forces: they are
    # compared against the velocity scale of the prestress, sigma / (rho c) with
    # c = sqrt(E / rho) (0.2 m/s), not against their own size.
    g, r, o = compare_states(gsim, osim, vel_tol=1.0)
    vscale = 2000.0 / (1000.0 * np.sqrt(1e5 / 1000.0))
    assert np.abs(g.vel - r.vel[o]).max() < 1e-4 * vscale
    lo, lg = osim.ledger(), gsim.ledger()
    assert lo.wall_normal_traction_energy != 0.0
    assert abs(lg.wall_normal_traction_energy - lo.wall_normal_traction_energy) <= \
        1e-3 * abs(lo.wall_normal_traction_energy)
    scale = np.linalg.norm(lo.gravity_impulse)
    for a in range(3):
        assert abs(lg.wall_impulse[a] - lo.wall_impulse[a]) < 1e-3 * scale


@pytest.mark.parametrize("gamma", [1.0, 7.0])
def test_dam_break_liquid_parity(engine, gamma):
    spec = cases.dam_break_block(cells=(24, 16, 16), block=(12, 16, 12), dx=0.001, gamma=gamma)
    osim, gsim, dt = fixed_dt_run(engine, spec, 100)
    g, r, o = compare_states(gsim, osim)
    assert np.abs(g.volume_ratio - r.volume_ratio[o]).max() < 1e-5


def test_coupled_patch_pellet_parity(engine):
    # research_coupled_patch with the pellet already pressing into the paste
    # (1 mm below the block top, moving down at 0.2 m/s) so coupling acts every step.
    spec, _, _ = cases.research_coupled_patch()
    osim = build_oracle(spec)
    p = osim.pellets()[0]
    p.position[2] = 0.02 + p.radius - 0.001
    p.velocity[:] = [0.0, 0.0, -0.2]
    p.angular_momentum[:] = [0.0, 2e-9, 0.0]
    osim.set_pellet(0, p)
    led = osim.ledger()
    led.initial_momentum[2] += p.mass * -0.2
    osim.set_ledger(led)
    gsim = engine_from_oracle(engine, osim)
    for k in range(100):
        assert osim.try_step(5e-5) == 0
        assert gsim.try_step(5e-5).accepted == 1, k
    compare_states(gsim, osim)
    po, pg = osim.pellets()[0], gsim.pellets()[0]
    assert rel(np.array(pg.position[:]), np.array(po.position[:])) < 1e-5
    vo = np.array(po.velocity[:])
    print("pellet velocity", list(pg.velocity), list(po.velocity), flush=True)
    assert np.abs(np.array(pg.velocity[:]) - vo).max() < 1e-3 * np.abs(vo).max()
    lo, lg = osim.ledger(), gsim.ledger()
    ci = np.array(lo.coupling_impulse[:])
    assert np.abs(ci).max() > 0.0
    assert np.abs(np.array(lg.coupling_impulse[:]) - ci).max() < 2e-3 * np.abs(ci).max()
    assert np.linalg.norm(gsim.momentum_residual()) < 1e-5 * np.linalg.norm(lo.gravity_impulse)


def _pellet_box(p):
    i, j, k, w = list(p.orientation)
    axis = np.array([1 - 2 * (j * j + k * k), 2 * (i * j + w * k), 2 * (i * k - w * j)])
    half = max(abs(o) for o in p.sphere_offsets[:p.sphere_count])
    reach = np.abs(axis) * half + p.radius
    c = np.array(p.position[:])
    return c - reach, c + reach


def test_overlapping_pellets_parity(engine):
    # Four pellets (r = 1 dx, L = 8 mm, rho 650: the 512-pellet bed's geometry at
    # this grid) pressed into the top of the 49,152-particle paste column with
    # overlapping boxes, so grid nodes see two or three pellets one after another
    # in index order (couple_grid per pellet, rigid.rs:385-438).  100 substeps.
    spec = cases.paste_column_32()
    osim = build_oracle(spec)
    r = 0.0025
    top = 0.06 + r - 0.001
    layout = [((0.0385, 0.0400, top), None, (0.02, 0.0, -0.2)),
              ((0.0410, 0.0405, top + 0.0005), ((0, 0, 1), 0.5 * math.pi), (-0.02, 0.01, -0.15)),
              ((0.0400, 0.0370, top), ((0, 0, 1), 0.25 * math.pi), (0.0, 0.03, -0.2)),
              ((0.0405, 0.0430, top + 0.001), ((1, 0, 1), 0.6), (0.01, -0.03, -0.25))]
    mom = np.zeros(3)
    for pos, rot, vel in layout:
        q = (0.0, 0.0, 0.0, 1.0)
        if rot is not None:
            ax = np.asarray(rot[0], float) / np.linalg.norm(rot[0])
            q = tuple(list(ax * math.sin(0.5 * rot[1])) + [math.cos(0.5 * rot[1])])
        p = O.new_pellet(r, 0.008, 650.0, list(pos), orientation=q)
        p.velocity[:] = list(vel)
        osim.add_pellet(p)
        mom += p.mass * np.array(vel)
    pels = osim.pellets()
    boxes = [_pellet_box(p) for p in pels]
    overlaps = sum(bool(np.all(boxes[a][0] <= boxes[b][1]) and np.all(boxes[b][0] <= boxes[a][1]))
                   for a in range(4) for b in range(a + 1, 4))
    assert overlaps >= 3, overlaps
    led = osim.ledger()
    for a in range(3):
        led.initial_momentum[a] += mom[a]
    osim.set_ledger(led)
    gsim = engine_from_oracle(engine, osim)
    dt = min(osim.stability_limit(), spec.config.max_dt)
    for k in range(100):
        assert osim.try_step(dt) == 0, k
        rep = gsim.try_step(dt)
        assert rep.accepted == 1, (k, rep.failure)
    compare_states(gsim, osim)
    po, pg = osim.pellets(), gsim.pellets()
    for a, b in zip(pg, po):
        assert rel(np.array(a.position[:]), np.array(b.position[:])) < 1e-5
        vo = np.array(b.velocity[:])
        assert np.abs(np.array(a.velocity[:]) - vo).max() < 1e-3 * np.abs(vo).max()
        lo_ = np.array(b.angular_momentum[:])
        assert np.abs(np.array(a.angular_momentum[:]) - lo_).max() <= 1e-3 * max(np.abs(lo_).max(), 1e-12)
    lo, lg = osim.ledger(), gsim.ledger()
    for f in ("coupling_impulse", "coupling_angular_impulse"):
        ci = np.array(getattr(lo, f)[:])
        assert np.abs(ci).max() > 0.0
        assert np.abs(np.array(getattr(lg, f)[:]) - ci).max() < 2e-3 * np.abs(ci).max(), f
    assert abs(lg.coupling_grid_energy - lo.coupling_grid_energy) < 2e-3 * abs(lo.coupling_grid_energy)
    assert np.linalg.norm(gsim.momentum_residual()) < 1e-5 * np.linalg.norm(lo.gravity_impulse)


@pytest.mark.parametrize("depth", [2, 3])
def test_thin_grid_traction_out_of_equilibrium(engine, depth):
    # The thin block of test_thin_grid_wall_traction_on_both_faces out of
    # equilibrium: a lateral compressive stress that its deformation does not
    # hold makes it spread across both faces at a few cm/s, under the faces'
    # wall-image traction (solver.rs:725-819) and a low Coulomb coefficient whose
    # traction budget (solver.rs:866-886) acts on the sliding.  The velocities are
    # their own scale here, so the plain 1e-4 bar applies.
    m = O.paste(1000.0, 1e5, 0.2, 1e6, 1.0, 1.0)
    cfg = O.config(wall_friction=0.05, cfl=0.3, max_dt=5e-5, min_dt=1e-9)
    dx = 0.002
    spec = O.fixture("slump", m, dx, [10 * dx, 10 * dx, depth * dx], [6 * dx, 6 * dx, depth * dx], cfg, seed=7)
    osim = build_oracle(spec)
    st = osim.state()
    st.kirchhoff[:] = 0.0
    for a in range(3):
        st.kirchhoff[:, a, a] = -2000.0
    st.deformation[:, 2, 2] = np.exp(-2000.0 / 1e5)
    osim.set_state(st)
    gsim = engine_from_oracle(engine, osim)
    dt = min(osim.stability_limit(), cfg.max_dt)
    for k in range(30):
        assert osim.try_step(dt) == 0, k
        rep = gsim.try_step(dt)
        assert rep.accepted == 1, (k, rep.failure)
    g, r, o = compare_states(gsim, osim)
    assert np.abs(r.vel).max() > 1e-2
    lo, lg = osim.ledger(), gsim.ledger()
    for f in ("wall_normal_traction_energy", "wall_friction_dissipation"):
        ref = getattr(lo, f)
        assert ref != 0.0
        assert abs(getattr(lg, f) - ref) <= 1e-3 * abs(ref), f


def _f32_key_cells(pos, dx):
    """The cell of each coordinate as an f32 quotient by the f32 spacing (the
    rounding the f64 key avoids)."""
    return np.floor(pos.astype(np.float32) / np.float32(dx)).astype(np.int64)


def test_reference_order_is_the_reference_key_exactly(engine):
    # solver.rs:501-517: every 32 accepted steps the particles are ordered by
    # (clamped cell, id), the cell from the f64 position over the f64 spacing.
    # A paste column at rest without gravity keeps every position bit for bit
    # through a step, so the engine and the oracle sort the same positions; a
    # third of the coordinates sit on, or one f32 ulp either side of, a cell
    # face, where a quotient rounded to f32 can floor into the neighbouring cell.
    spec = cases.paste_column_32()
    spec.config.gravity[:] = [0.0, 0.0, 0.0]
    osim = build_oracle(spec)
    st = osim.state()
    n = len(st.id)
    dx = spec.grid_spacing
    rng = np.random.default_rng(11)
    pos = st.pos.astype(np.float32)
    face = (np.round(pos.astype(np.float64) / dx) * dx).astype(np.float32)
    way = rng.integers(0, 3, pos.shape)
    moved = np.where(way == 0, face, np.where(way == 1, np.nextafter(face, np.float32(np.inf)),
                                              np.nextafter(face, np.float32(-np.inf))))
    pos = np.where(rng.random(pos.shape) < 0.3, moved, pos).astype(np.float64)
    assert np.any(_f32_key_cells(pos, dx) != np.floor(pos / dx).astype(np.int64))
    perm = rng.permutation(n)
    for f in ("id", "pos", "vel", "affine", "deformation", "volume_ratio", "kirchhoff", "plastic", "mass", "volume0"):
        setattr(st, f, np.ascontiguousarray(getattr(st, f)[perm]))
    st.pos = np.ascontiguousarray(pos[perm])
    osim.set_state(st)
    h = osim.history()
    h.step_count = 31
    osim.set_history(h)
    gsim = engine_from_oracle(engine, osim)
    gh = engine.History()
    for name, _ in h._fields_:
        setattr(gh, name, getattr(h, name))
    gsim.set_state(gsim.state(), history=gh)
    osim.advance_to(1e-5)
    gsim.advance_to(1e-5)
    assert osim.history().step_count == 32 and gsim.history().step_count == 32
    g, r = gsim.state(), osim.state()
    o = np.argsort(st.id)
    assert np.array_equal(g.by_id().pos, st.pos[o]), "positions changed during the step at rest"
    assert np.array_equal(g.id, r.id)
