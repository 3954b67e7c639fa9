"""Batched instances, captured-graph stepping and the autotuner, against the f64 oracle.

A batch steps independent instances (different materials, grids, particle
counts, pellets and adaptive dt) in one launch sequence; each instance must
match its own oracle run to the single-instance tolerances: positions and
velocities within 1e-4 relative (normwise), stress within 1e-3, identical step
and rejection counts.
"""
import numpy as np
import pytest

import cases
import oracle as O
from test_gpu_parity import as_engine, build_oracle, compare_states, osim_material

pytestmark = pytest.mark.gpu


def spec_from_oracle(E, osim, cfg=None):
    st = osim.state()
    est = E.State(st.id.copy(), st.pos, st.vel, st.affine, st.deformation, st.volume_ratio, st.kirchhoff,
                  st.plastic, st.mass, st.volume0)
    spacing, cells = osim.layout()
    return {"material": as_engine(E, osim_material(osim), E.Material),
            "cfg": as_engine(E, cfg if cfg is not None else osim.config(), E.Config),
            "domain": [cells[a] * spacing for a in range(3)], "spacing": spacing, "state": est,
            "pellets": [as_engine(E, p, E.Pellet) for p in osim.pellets()]}


class View:
    """One instance of a batch seen through the single-instance comparison helpers."""

    def __init__(self, batch, i):
        self.batch, self.i = batch, i

    def state(self):
        return self.batch.state(self.i)


def small_slump():
    m = O.paste(1200.0, 138461.538, 0.384615, 50.0, 5.0, 0.5, shear_units=True)
    cfg = O.config(wall_friction=0.0, cfl=0.4, max_dt=1e-3, min_dt=1e-12,
                   face_friction=[0.0, 0.0, 0.0, 0.0, float("inf"), 0.0])
    dx = 0.0025
    return O.fixture("slump", m, dx, [16 * dx] * 3, [0.02, 0.02, 0.03], cfg, seed=7, jitter=0.5)


def pressed_patch():
    spec, _, _ = cases.research_coupled_patch()
    osim = build_oracle(spec)
    p = osim.pellets()[0]
    p.position[2] = 0.02 + p.radius - 0.001
    p.velocity[:] = [0.0, 0.0, -0.2]
    osim.set_pellet(0, p)
    led = osim.ledger()
    led.initial_momentum[2] += p.mass * -0.2
    osim.set_ledger(led)
    return osim


def test_batch_instances_match_their_oracles(engine):
    # Three unlike instances: a paste slump (16^3 grid), a gamma = 7 dam break
    # (24 x 16 x 16 grid) and a pellet pressed into a paste patch.
    oracles = [build_oracle(small_slump()),
               build_oracle(cases.dam_break_block(cells=(24, 16, 16), block=(12, 16, 12), dx=0.001, gamma=7.0)),
               pressed_patch()]
    specs = [spec_from_oracle(engine, o) for o in oracles]
    batch = engine.Batch(specs)
    targets = [3e-3, 1.5e-3, 2e-3]
    for o, t in zip(oracles, targets):
        o.advance_to(t)
    status, done_ms = batch.advance(targets)
    print("batch status", status, "done ms", done_ms, flush=True)
    assert list(status) == [0, 0, 0]
    for i, o in enumerate(oracles):
        ho, hg = o.history(), batch.history(i)
        print(f"instance {i}: steps {hg.step_count} vs {ho.step_count}", flush=True)
        assert hg.step_count == ho.step_count and hg.rejected_steps == ho.rejected_steps
        assert hg.time == ho.time
        compare_states(View(batch, i), o)
        lo, lg = o.ledger(), batch.ledger(i)
        scale = np.linalg.norm(lo.gravity_impulse)
        assert abs(lg.initial_mass - lo.initial_mass) <= 1e-7 * lo.initial_mass
        assert np.abs(np.array(lg.gravity_impulse[:]) - np.array(lo.gravity_impulse[:])).max() < 1e-6 * scale


def test_batch_of_one_equals_simulation(engine):
    osim = build_oracle(small_slump())
    spec = spec_from_oracle(engine, osim)
    sim = engine.Simulation(spec["material"], spec["cfg"], spec["domain"], spec["spacing"], spec["state"])
    batch = engine.Batch([spec])
    sim.advance_to(2e-3)
    batch.advance([2e-3])
    assert sim.history().step_count == batch.history(0).step_count
    a, b = sim.state().by_id(), batch.state(0).by_id()
    assert np.abs(a.pos - b.pos).max() <= 1e-6 * np.abs(a.pos).max()


@pytest.mark.parametrize("use_graph", [0, 1])
def test_launch_shapes_match_oracle(engine, use_graph):
    # Captured-graph and stream stepping, with one or two blocks per work item,
    # both follow the oracle.
    spec = cases.paste_column_32(max_dt=1e-3)
    osim = build_oracle(spec)
    s = spec_from_oracle(engine, osim)
    sim = engine.Simulation(s["material"], s["cfg"], s["domain"], s["spacing"], s["state"])
    sim.set_launch(use_graph=use_graph, steps_per_sync=8, merge_blocks=2 if use_graph else 1)
    osim.advance_to(3e-3)
    sim.advance_to(3e-3)
    assert sim.history().step_count == osim.history().step_count
    compare_states(sim, osim)


def test_autotune_leaves_state_unchanged(engine):
    osim = build_oracle(small_slump())
    s = spec_from_oracle(engine, osim)
    sim = engine.Simulation(s["material"], s["cfg"], s["domain"], s["spacing"], s["state"])
    osim.advance_to(1e-3)
    sim.advance_to(1e-3)
    before, led, hist = sim.state(), bytes(sim.ledger()), bytes(sim.history())
    chosen = sim.autotune(steps=4)
    print("autotune chose", chosen.as_dict(), flush=True)
    after = sim.state()
    assert bytes(sim.ledger()) == led and bytes(sim.history()) == hist
    assert np.array_equal(after.id, before.id)
    assert np.array_equal(after.pos, before.pos) and np.array_equal(after.kirchhoff, before.kirchhoff)
    assert sim.launch().as_dict() == chosen.as_dict()
    # Stepping goes on under the chosen shape and still follows the oracle.
    osim.advance_to(2e-3)
    sim.advance_to(2e-3)
    compare_states(sim, osim)


def test_step_budget_and_callback(engine):
    osim = build_oracle(small_slump())
    s = spec_from_oracle(engine, osim)
    a = engine.Simulation(s["material"], s["cfg"], s["domain"], s["spacing"], s["state"])
    b = engine.Simulation(s["material"], s["cfg"], s["domain"], s["spacing"], s["state"])
    assert a.advance_steps(10, target=1.0) == 10
    calls = [0]

    def keep_going():
        calls[0] += 1
        return calls[0] <= 10

    assert b.advance_while(1.0, keep_going) is False
    assert b.history().step_count == 10 == a.history().step_count
    # Same steps; the times agree to the round-off of the atomically summed grid.
    assert abs(a.history().time - b.history().time) <= 1e-9 * a.history().time
