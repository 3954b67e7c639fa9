"""The GPU energy audit (mpm_audit_step) against the oracle's audit
(core/src/mpm/audit.rs), on the reference's audit fixtures and on states of
the research slump and coupled patch.

The engine's stage energies are f64 sums over f32 node values, so they agree
with the f64 oracle to f32 resolution of the energies involved: every audit
quantity within 2e-6 of the fixture's energy scale (the sum of the magnitudes
of its endpoint and stage energies).  The analytic identities themselves are
pinned on the oracle in tests/test_oracle_audit.py.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from cases import record_loop, research_coupled_patch, research_slump
from test_oracle_audit import edit, isolated, plastic_fixture

pytestmark = pytest.mark.gpu


def convert(src, cls):
    dst = cls()
    C.memmove(C.byref(dst), C.byref(src), C.sizeof(cls))
    return dst


def engine_twin(E, osim, max_nodes=1 << 40):
    """An engine simulation holding the oracle's state, ledger and history."""
    st = osim.state()
    est = E.State(st.id.copy(), st.pos, st.vel, st.affine, st.deformation, st.volume_ratio, st.kirchhoff, st.plastic,
                  st.mass, st.volume0)
    spacing, cells = osim.layout()
    cfg = convert(osim.config(), E.Config)
    mat = convert(osim_material(osim), E.Material)
    return E.Simulation(mat, cfg, [cells[a] * spacing for a in range(3)], spacing, est,
                        [convert(p, E.Pellet) for p in osim.pellets()], max_nodes=max_nodes,
                        ledger=convert(osim.ledger(), E.Ledger), history=convert(osim.history(), E.History))


def osim_material(osim):
    m = O.Material()
    O.lib().oracle_sim_get_material(osim.h, C.byref(m))
    return m


def flatten(a):
    out = []
    for group in ("before", "after", "stages", "increments"):
        out += [(f"{group}.{k}", v) for k, v in a[group].items()]
    out += [(f"terms.{k}", v) for k, v in a["terms"].items() if v is not None]
    return dict(out)


def compare(gpu, ref, rel=2e-6):
    scale = sum(abs(v) for g in ("before", "after", "stages") for v in ref[g].values()) + 1e-300
    fg, fr = flatten(gpu), flatten(ref)
    assert fg.keys() == fr.keys()
    worst = max(abs(fg[k] - fr[k]) for k in fr) / scale
    for k in fr:
        assert abs(fg[k] - fr[k]) <= rel * scale, f"{k}: gpu {fg[k]!r} oracle {fr[k]!r} (scale {scale:.3e})"
    assert (gpu["terms"]["uncoupled_gravity_potential_mismatch"] is None) == \
        (ref["terms"]["uncoupled_gravity_potential_mismatch"] is None)
    return worst


def state_bytes(sim):
    st = sim.state()
    return ([np.copy(getattr(st, f)) for f in ("id", "pos", "vel", "affine", "deformation", "kirchhoff", "plastic",
                                                "mass", "volume0")],
            bytes(sim.ledger()), bytes(sim.history()), [bytes(p) for p in sim.pellets()])


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[0], b[0])) and a[1:] == b[1:]


def analytic_fixtures():
    """The fixtures of audit_tests.rs, as (name, oracle sim, dt)."""
    out = []
    s = isolated()
    edit(s, vel=[[1.0, -2.0, 0.5]], affine=[np.array([[0.2, 0.3, -0.1], [0.5, -0.4, 0.7], [0.1, 0.2, 0.4]])])
    out.append(("translation_affine", s, 1e-4))
    p = (0.47, 0.52, 0.55)
    out.append(("colliding", isolated(particles=((p, (1.0, 0.0, 0.0), 2.0, 0.002), (p, (-1.0, 0.0, 0.0), 2.0, 0.002))),
                1e-4))
    s = isolated(particles=(((0.47, 0.52, 0.55), (0.7, -0.2, 0.3), 2.0, 0.002),
                            ((0.51, 0.49, 0.58), (-0.4, 0.8, -0.1), 3.0, 0.003)))
    edit(s, affine=[np.eye(3) * 0.3, np.array([[0.1, -0.2, 0.3], [0.4, 0.2, -0.1], [0.0, 0.3, -0.2]])])
    out.append(("variance", s, 1e-4))
    out.append(("freefall", isolated(gravity=(0.0, 0.0, -9.81)), 1e-3))
    liquid = O.liquid(1000.0, 1000.0, 0.3, 0.0)
    s = isolated(material=liquid)
    edit(s, volume_ratio=[0.95], kirchhoff=[np.eye(3) * (liquid.bulk_modulus * (0.95 - 1.0))])
    out.append(("prestressed_liquid", s, 1e-4))
    s = isolated(gravity=(0.0, 0.0, -9.81))
    edit(s, pos=[[0.47, 0.52, 0.025]], vel=[[0.1, 0.0, -0.5]], kirchhoff=[np.eye(3) * -1000.0])
    out.append(("walls", s, 1e-4))
    for density in (1.0, 500.0):
        s = isolated(particles=(((0.5, 0.5, 0.5), (0.0, 0.0, 1.0), 2.0, 0.002),), gravity=(0.0, 0.0, -9.81))
        s.add_pellet(O.new_pellet(0.15, 0.3, density, [0.5, 0.5, 0.65]))
        out.append((f"coupled_{density}", s, 1e-4))
    out.append(("plastic", plastic_fixture(3.0), 1e-3))
    return out


@pytest.mark.parametrize("index", range(9))
def test_audit_matches_oracle_on_reference_fixtures(engine, index):
    name, osim, dt = analytic_fixtures()[index]
    sim = engine_twin(engine, osim, max_nodes=10000)
    before = state_bytes(sim)
    gpu = sim.audit_step(dt)
    ref = osim.audit_step(dt)
    worst = compare(gpu, ref)
    print(f"{name}: worst {worst:.2e} of the energy scale, p2g {gpu['terms']['p2g_transfer_loss']:.6e} "
          f"(oracle {ref['terms']['p2g_transfer_loss']:.6e})", flush=True)
    assert same(state_bytes(sim), before)
    assert abs(gpu["terms"]["telescoping_error"]) <= 1e-9 * max(1.0, abs(gpu["before"]["potential"]))


def test_audit_rejections_do_not_mutate(engine):
    osim = isolated()
    sim = engine_twin(engine, osim, max_nodes=10000)
    before = state_bytes(sim)
    for dt in (0.0, -1.0, float("nan"), float("inf"), 1e-12, 1.0):
        with pytest.raises(engine.MpmError):
            sim.audit_step(dt)
        assert same(state_bytes(sim), before)
    for fields in ({"pos": [[-10.0, 0.52, 0.55]]}, {"kirchhoff": [np.eye(3) * 1e20]}):
        o = isolated()
        edit(o, **fields)
        s = engine_twin(engine, o, max_nodes=10000)
        snap = state_bytes(s)
        with pytest.raises(engine.MpmError):
            s.audit_step(1e-4)
        assert same(state_bytes(s), snap)
    o = isolated()
    h = o.history()
    h.dt_scale = 0.001
    o.set_history(h)
    with pytest.raises(engine.MpmError, match="outside finite accepted range"):
        engine_twin(engine, o, max_nodes=10000).audit_step(1e-4)


def test_audit_then_advance_continues_like_the_oracle(engine):
    # The audit leaves nothing behind: stepping afterwards follows the oracle.
    osim = isolated(gravity=(0.0, 0.0, -9.81))
    edit(osim, vel=[[0.1, -0.2, 0.3]])
    sim = engine_twin(engine, osim, max_nodes=10000)
    a = sim.audit_step(1e-4)
    sim.advance_to(1e-4)
    osim.advance_to(1e-4)
    st, ost = sim.state(), osim.state()
    assert sim.history().step_count == osim.history().step_count == 1
    assert np.abs(st.pos - ost.pos).max() < 1e-6
    assert abs(a["after"]["translation"] - float(np.sum(0.5 * ost.mass * np.sum(ost.vel ** 2, axis=1)))) < 1e-6


@pytest.mark.parametrize("case", ["slump", "coupled_patch"])
def test_audit_of_research_states_matches_oracle(engine, case):
    spec, _, interval = research_slump() if case == "slump" else research_coupled_patch()
    osim = O.Sim.from_fixture(spec)
    record_loop(osim, 5 * interval, interval)
    sim = engine_twin(engine, osim)
    dt = 0.5 * min(sim.stability_limit(), osim.config().max_dt)
    gpu = sim.audit_step(dt)
    ref = osim.audit_step(dt)
    worst = compare(gpu, ref, rel=1e-5)
    print(f"{case}: worst {worst:.2e} of the energy scale; transport {gpu['stages']['grid_transport']:.6e} "
          f"(oracle {ref['stages']['grid_transport']:.6e})", flush=True)
    assert gpu["stages"]["grid_transport"] > 0.0
