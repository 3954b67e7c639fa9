"""Run-to-run and order independence of the engine's trajectories.

The reference's results do not depend on its thread count (solver.rs:5-9:
fixed-order reductions).  The engine gets the same property a different way:
every sum whose order of additions varies with the scheduling -- the deposit,
the wall reactions, the pellets' coupling sums, the plastic dissipation -- is
an exact integer, and the deposit's units are chosen from the state alone.  So
two runs of one state agree bit for bit, and so does a run of the same
particles uploaded in another order (another device order, other work items,
other locality sorts), compared particle by particle.
"""
import numpy as np
import pytest

import cases
import oracle as O
from test_gpu_parity import as_engine, build_oracle, engine_from_oracle

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "affine", "deformation", "volume_ratio", "kirchhoff", "plastic", "mass", "volume0")


def shuffled(E, st, seed):
    perm = np.random.default_rng(seed).permutation(len(st.id))
    return E.State(st.id[perm].copy(), *[np.ascontiguousarray(getattr(st, f)[perm]) for f in FIELDS])


def same_by_id(a, b):
    a, b = a.by_id(), b.by_id()
    assert np.array_equal(a.id, b.id)
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def paste_run(E, order_seed=None, steps_to=3e-3):
    spec = cases.paste_column_32(max_dt=1e-3)
    osim = build_oracle(spec)
    sim = engine_from_oracle(E, osim)
    if order_seed is not None:
        # The ledger as seeded from the unshuffled state (its f64 sums follow the given order).
        sim.set_state(shuffled(E, sim.state(), order_seed), ledger=sim.ledger(), history=sim.history())
    sim.advance_to(steps_to)
    return sim


def test_two_runs_are_identical(engine):
    a, b = paste_run(engine), paste_run(engine)
    assert a.history().step_count > 40  # past a locality sort and a reference-order sort
    same_by_id(a.state(), b.state())
    assert bytes(a.ledger()) == bytes(b.ledger()) and bytes(a.history()) == bytes(b.history())


def test_particle_order_does_not_change_the_trajectory(engine):
    # The same particles in two other upload orders: other device orders and work
    # items until the first locality sort, other tile and global deposit paths.
    a = paste_run(engine)
    for seed in (1, 2):
        b = paste_run(engine, order_seed=seed)
        same_by_id(a.state(), b.state())
        assert bytes(a.ledger()) == bytes(b.ledger()), seed
        assert bytes(a.history()) == bytes(b.history()), seed


def test_pellet_coupling_is_order_independent(engine):
    # Several pellets on shared nodes: the coupling sums are exact, so the
    # pellets' trajectories do not depend on the particle order either.
    spec, _, _ = cases.research_coupled_patch()
    osim = build_oracle(spec)
    p = osim.pellets()[0]
    p.position[2] = 0.02 + p.radius - 0.001
    p.velocity[:] = [0.0, 0.0, -0.2]
    osim.set_pellet(0, p)
    q = O.new_pellet(0.003, 0.012, 1100.0, [p.position[0] + 0.004, p.position[1] + 0.002, p.position[2]],
                     orientation=(0.0, 0.0, 0.5646424733950354, 0.8253356149096783))
    q.velocity[:] = [0.0, 0.0, -0.15]
    osim.add_pellet(q)
    runs = []
    for seed in (None, 3):
        sim = engine_from_oracle(engine, osim)
        if seed is not None:
            sim.set_state(shuffled(engine, sim.state(), seed), pellets=sim.pellets(), ledger=sim.ledger(),
                          history=sim.history())
        for _ in range(60):
            assert sim.try_step(5e-5).accepted == 1
        runs.append(sim)
    a, b = runs
    same_by_id(a.state(), b.state())
    assert bytes(a.ledger()) == bytes(b.ledger())
    for pa, pb in zip(a.pellets(), b.pellets()):
        assert bytes(pa) == bytes(pb)
