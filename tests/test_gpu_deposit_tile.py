"""The fixed-point LDS deposit tile of the fused G2P2G kernel against the oracle,
on states whose stored Kirchhoff stresses do not predict the next ones.

A particle's new stress comes from its deformation (paste) or volume ratio
(liquid) alone (constitutive.rs:456-589); the stored stress only enters the
next deposit (solver.rs:632-661).  The engine scales its deposit tile's stress
sums by the stored stresses of each workgroup, so states whose stored stresses
are zeroed or scaled down send the first step's stressed particles to the
global deposit, and sparse outliers scaled up coarsen the scale of their
workgroups (their neighbours fall below it and take the global path too).  The
oracle follows each edited state as it is; the engine must match it to the
parity bar (positions and velocities 1e-4, stress 1e-3, normwise) after 30
substeps.
"""
import ctypes as C

import pytest

import cases
import oracle as O
from test_gpu_parity import compare_states

pytestmark = pytest.mark.gpu


def convert(src, cls):
    dst = cls()
    C.memmove(C.byref(dst), C.byref(src), C.sizeof(cls))
    return dst


def engine_twin(E, osim):
    """An engine simulation holding the oracle's state, ledger and history."""
    st = osim.state()
    est = E.State(st.id.copy(), st.pos, st.vel, st.affine, st.deformation, st.volume_ratio, st.kirchhoff, st.plastic,
                  st.mass, st.volume0)
    spacing, cells = osim.layout()
    m = O.Material()
    O.lib().oracle_sim_get_material(osim.h, C.byref(m))
    return E.Simulation(convert(m, E.Material), convert(osim.config(), E.Config),
                        [cells[a] * spacing for a in range(3)], spacing, est,
                        [convert(p, E.Pellet) for p in osim.pellets()],
                        ledger=convert(osim.ledger(), E.Ledger), history=convert(osim.history(), E.History))


def spec_of(name):
    if name == "slump":
        return cases.paste_column_32()
    return cases.dam_break_block(cells=(24, 16, 16), block=(12, 16, 12), dx=0.001, gamma=7.0)


@pytest.mark.parametrize("edit", ["stored", "zero", "scaled_down", "outliers"])
@pytest.mark.parametrize("name", ["slump", "dam_break"])
def test_deposit_paths_follow_the_oracle(engine, name, edit):
    spec = spec_of(name)
    osim = O.Sim.from_fixture(spec)
    dt = min(osim.stability_limit(), spec.config.max_dt)
    for _ in range(20):
        assert osim.try_step(dt) == 0
    st = osim.state()
    assert abs(st.kirchhoff).max() > 0.0
    if edit == "zero":
        st.kirchhoff[:] = 0.0
    elif edit == "scaled_down":
        st.kirchhoff *= 1e-6
    elif edit == "outliers":
        st.kirchhoff[::61] *= 10.0
    osim.set_state(st)
    sim = engine_twin(engine, osim)
    for k in range(30):
        assert osim.try_step(dt) == 0, k
        rep = sim.try_step(dt)
        assert rep.accepted == 1, (k, rep.failure)
    print(f"{name}/{edit}", end=" ", flush=True)
    compare_states(sim, osim)
