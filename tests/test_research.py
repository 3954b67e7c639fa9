"""Research wire layer on the CPU: the checkpoint codec, request validation and
the observables (core/src/research.rs, core/src/research/tests.rs,
core/src/mpm/fixtures.rs:369-615).

No test here touches the GPU: the observables and the codec are host code that
takes a state; the states come from the f64 oracle, whose trajectories the
reference's documents pin (docs/validation.md, docs/hydrostatic-balance.md).
"""
import copy
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from cases import record_loop, research_coupled_patch, research_slump, water_column

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_request(name):
    with open(os.path.join(GOLDEN, f"request_{name}.json")) as f:
        return json.load(f)


def small_request():
    """The request of core/src/research/tests.rs:4-23."""
    return {
        "schema_version": 1, "mode": "research", "seed": 3,
        "duration_s": 0.002, "dt_s": 0.0005, "record_interval_s": 0.001,
        "max_wall_time_s": 60.0, "boxes": [], "events": [],
        "materials": {"friction": 0.3, "restitution": 0.1, "normal_stiffness_n_m": 1000.0,
                      "water_capacity_ratio": 3.0, "uptake_rate_s": 0.1, "breakdown_rate_s": 0.01,
                      "evaporation_rate_s": 0.0},
        "research": {"fixture": "slump", "grid_spacing_m": 0.005, "domain_m": [0.03, 0.03, 0.03],
                     "material": "paste", "density_kg_m3": 1000.0, "young_modulus_pa": 1000.0,
                     "poisson_ratio": 0.2, "yield_stress_pa": 30.0, "consistency_pa_s_n": 10.0,
                     "flow_index": 0.6, "initial_size_m": [0.01, 0.01, 0.01],
                     "initial_velocity_m_s": [0.0, 0.0, 0.0]},
    }


def convert(E, src, cls):
    dst = cls()
    C.memmove(C.byref(dst), C.byref(src), C.sizeof(cls))
    return dst


def engine_state(E, ost):
    return E.State(ost.id.copy(), ost.pos, ost.vel, ost.affine, ost.deformation, ost.volume_ratio, ost.kirchhoff,
                   ost.plastic, ost.mass, ost.volume0)


def observables_of(E, request, osim):
    return E.research_observables(json.dumps(request), engine_state(E, osim.state()),
                                  [convert(E, p, E.Pellet) for p in osim.pellets()],
                                  convert(E, osim.ledger(), E.Ledger), convert(E, osim.history(), E.History))


# ------------------------------------------------------------------ validation

def test_protocol_version(engine):
    assert engine.protocol_version() == 1


def test_valid_requests_pass(engine):
    engine.research_validate(json.dumps(small_request()))
    for name in ("research_slump", "research_coupled_patch", "research_hydrostatic_water"):
        engine.research_validate(json.dumps(golden_request(name)))


def test_unsupported_fines_and_bad_parameters_fail_closed(engine):
    # research/tests.rs:84-101
    for key, value in [("material", "fines"), ("material", "unknown"), ("poisson_ratio", 0.5), ("flow_index", 0),
                       ("grid_spacing_m", 1e-300), ("initial_size_m", [0.011, 0.01, 0.01])]:
        req = small_request()
        req["research"][key] = value
        with pytest.raises(ValueError):
            engine.research_validate(json.dumps(req))
    req = small_request()
    req["unknown"] = 1
    with pytest.raises(ValueError, match="unknown field"):
        engine.research_validate(json.dumps(req))


@pytest.mark.parametrize("edit, message", [
    (lambda r: r.update(dt_s=0.01), "dt exceeds duration"),
    (lambda r: r.update(max_wall_time_s=4000.0), "budget exceeds 3600"),
    (lambda r: r.update(record_interval_s=1e-8), "resource cap"),
    (lambda r: r.update(schema_version=2), "schema_version=1"),
    (lambda r: r.update(seed=-1), "expected u64"),
    (lambda r: r.update(seed=3.0), "expected u64"),
    (lambda r: r.pop("materials"), "missing field `materials`"),
    (lambda r: r["materials"].update(restitution=1.5), "restitution"),
    (lambda r: r["research"].update(fixture="avalanche"), "unsupported research fixture"),
    (lambda r: r["research"].update(domain_m=[0.03, 0.03, 0.0301]), "integer multiple"),
    (lambda r: r["research"].update(grid_spacing_m=0.0005, domain_m=[0.06, 0.06, 0.06]), "node"),
    (lambda r: r["research"].update(material="water"), "water requires zero yield stress"),
    (lambda r: r["research"].update(fixture="hydrostatic"), "hydrostatic"),
    (lambda r: r["research"].update(fixture="coupled_patch"), "pellet geometry"),
    (lambda r: r["research"].pop("flow_index"), "missing field `flow_index`"),
    (lambda r: r.update(research=None), "research parameters are required"),
    (lambda r: r.update(events=[{"time_s": 0.0, "kind": "dig", "box_id": "a", "position_m": [0, 0, 0],
                                 "direction": [1, 0, 0], "duration_s": 0.0, "amount_kg": 0.0,
                                 "water_fraction": 0.0, "radius_m": 0.01}]), "no household events"),
])
def test_request_rejections(engine, edit, message):
    req = small_request()
    edit(req)
    with pytest.raises(ValueError, match=message):
        engine.research_validate(json.dumps(req))


def test_particle_and_frame_caps(engine):
    req = small_request()
    # 0.06 m cube at 2.5 mm, 0.05 m block: 40^3 = 64000 particles fit the 100k cap ...
    req["research"].update(grid_spacing_m=0.0025, domain_m=[0.06, 0.06, 0.06], initial_size_m=[0.05, 0.05, 0.05])
    engine.research_validate(json.dumps(req))
    # ... 48^3 = 110592 do not.
    req["research"].update(initial_size_m=[0.06, 0.06, 0.06])
    with pytest.raises(ValueError, match="exceeding the cap of '100000'"):
        engine.research_validate(json.dumps(req))
    req = small_request()
    req["research"].update(grid_spacing_m=0.0025, domain_m=[0.06, 0.06, 0.06], initial_size_m=[0.05, 0.05, 0.05])
    req.update(duration_s=1.0, record_interval_s=0.01, dt_s=1e-4)  # 101 frames x 64064 points
    with pytest.raises(ValueError, match="frame-point resource cap"):
        engine.research_validate(json.dumps(req))


def test_coupled_patch_rules(engine):
    req = golden_request("research_coupled_patch")
    req["boxes"][0]["pellet_radius_m"] = 0.0025  # under 1.5 cells of 2 mm
    req["boxes"][0]["pellet_length_m"] = 0.01
    with pytest.raises(ValueError, match="under-resolved"):
        engine.research_validate(json.dumps(req))
    req = golden_request("research_coupled_patch")
    req["research"]["fixture"] = "slump"
    with pytest.raises(ValueError, match="only coupled_patch accepts"):
        engine.research_validate(json.dumps(req))


def test_run_json_rejects_before_touching_the_device(engine):
    with pytest.raises(ValueError, match="invalid request JSON"):
        engine.run_json("{not json")
    with pytest.raises(ValueError, match="mode must be household or research"):
        engine.run_json(json.dumps({"mode": "other"}))
    req = small_request()
    req["mode"] = "household"
    with pytest.raises(ValueError, match="household"):
        engine.run_json(json.dumps(req))
    req = small_request()
    req["research"]["material"] = "fines"
    with pytest.raises(ValueError, match="fines"):
        engine.run_json(json.dumps(req))
    with pytest.raises(ValueError, match="invalid checkpoint JSON"):
        engine.run_json(json.dumps(small_request()), "[1, 2")


# ------------------------------------------------------------------ checkpoint codec

def random_state(E, n, rng):
    """A state whose f64 fields are arbitrary finite bit patterns (all exponents, subnormals, signed zeros)."""
    def f64(*shape):
        bits = rng.integers(0, 2 ** 63, size=shape, dtype=np.int64).astype(np.uint64) * np.uint64(2) + \
            rng.integers(0, 2, size=shape, dtype=np.int64).astype(np.uint64)
        v = bits.view(np.float64)
        v[~np.isfinite(v)] = 1.5
        return v
    st = E.State(rng.permutation(n).astype(np.uint32), f64(n, 3), f64(n, 3), f64(n, 3, 3), f64(n, 3, 3), f64(n),
                 f64(n, 3, 3), f64(n), f64(n), f64(n))
    st.pos[0] = [0.0, -0.0, 5e-324]
    st.vel[0] = [1e-7, 1e16, 1.7976931348623157e308]
    return st


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_checkpoint_round_trip_is_bit_exact(engine):
    E = engine
    rng = np.random.default_rng(11)
    st = random_state(E, 257, rng)
    pel = E.new_pellet(0.003, 0.012, 1100.0, [0.02, 0.021, 0.0235])
    pel.orientation[:] = [0.1, -0.2, 0.3, math.sqrt(1 - 0.14)]
    pel.velocity[:] = [1e-300, -2.5, 3.0]
    led = E.Ledger()
    raw = (C.c_double * (C.sizeof(E.Ledger) // 8)).from_buffer(led)
    for k in range(len(raw)):
        raw[k] = float(rng.normal()) * 10.0 ** int(rng.integers(-300, 300))
    hist = E.History(0.0123, 4000, 17, 3, 0.5, 1e-9, 2e-4)
    req = json.dumps(small_request())
    text = E.checkpoint_encode(req, st, [pel], led, hist, 7)
    back, pels, led2, hist2, nri = E.checkpoint_decode(text)
    assert nri == 7 and len(pels) == 1
    assert np.array_equal(back.id, st.id)
    for f in ("pos", "vel", "affine", "deformation", "volume_ratio", "kirchhoff", "plastic", "mass", "volume0"):
        assert np.array_equal(bits(getattr(back, f)), bits(getattr(st, f))), f
    assert bytes(led2) == bytes(led)
    assert bytes(hist2) == bytes(hist)
    assert bytes(pels[0]) == bytes(pel)
    # An independent parser (Python's, correctly rounded) reads the same numbers.
    doc = json.loads(text)
    assert np.array_equal(bits(np.array(doc["state"]["particles"]["position"])), bits(st.pos))
    assert doc["time_s"] == 0.0123 and doc["state"]["min_dt_used"] == 1e-9
    # Encoding the decoded state gives the same text.
    assert E.checkpoint_encode(req, back, pels, led2, hist2, 7) == text


def test_checkpoint_layout_follows_serde(engine):
    E = engine
    st = E.State.empty(1)
    F = np.arange(1.0, 10.0).reshape(3, 3)
    st.deformation[0] = F
    st.mass[:] = 1.0
    st.volume0[:] = 1e-7
    st.pos[0] = [0.001, 1e16, 1e15]
    st.vel[0] = [-0.0, 123456.789, 1.5e300]
    hist = E.History(0.0, 0, 0, 0, 1.0, float("inf"), 0.0)
    text = E.checkpoint_encode(json.dumps(small_request()), st, [], E.Ledger(), hist, 1)
    pairs = []
    doc = json.loads(text, object_pairs_hook=lambda kv: pairs.append([k for k, _ in kv]) or dict(kv))
    # serde_json maps are sorted by key at every level.
    assert all(keys == sorted(keys) for keys in pairs)
    # nalgebra Matrix3: the nine entries column by column.
    assert doc["state"]["particles"]["deformation"][0] == list(F.T.flatten())
    assert doc["state"]["pellet"] is None and doc["state"]["min_dt_used"] is None
    assert doc["schema_version"] == 1 and doc["mode"] == "research"
    # f64 layout of the ryu crate used by serde_json.
    for literal in ('"position":[[0.001,1e16,1000000000000000.0]]', '"velocity":[[-0.0,123456.789,1.5e300]]',
                    '"volume0":[1e-7]', '"mass":[1.0]', '"time_s":0.0', '"step_count":0'):
        assert literal in text, literal


def test_checkpoint_decode_rejects_malformed(engine):
    E = engine
    st = E.State.empty(2)
    st.mass[:] = 1.0
    st.volume0[:] = 1.0
    st.id[:] = [0, 1]
    hist = E.History(0.0, 0, 0, 0, 1.0, float("inf"), 0.0)
    good = json.loads(E.checkpoint_encode(json.dumps(small_request()), st, [], E.Ledger(), hist, 1))
    cases = []
    c = copy.deepcopy(good)
    del c["state"]["ledger"]["wall_normal_traction_energy"]
    cases.append((c, "missing field `wall_normal_traction_energy`"))
    c = copy.deepcopy(good)
    c["state"]["ledger"]["wall_normal_traction_energy"] = None
    cases.append((c, "expected f64"))
    c = copy.deepcopy(good)
    c["state"]["extra"] = 1
    cases.append((c, "unknown field `extra`"))
    c = copy.deepcopy(good)
    c["state"]["particles"]["mass"] = [1.0]
    cases.append((c, "has length '1', expected '2'"))
    c = copy.deepcopy(good)
    c["state"]["particles"]["position"][0] = [1.0, 2.0]
    cases.append((c, "expected an array of length 3"))
    c = copy.deepcopy(good)
    c["state"]["next_record_index"] = -1
    cases.append((c, "expected u64"))
    for doc, message in cases:
        with pytest.raises(ValueError, match=message):
            E.checkpoint_decode(json.dumps(doc))


# ------------------------------------------------------------------ observables on oracle states

def test_initial_observables_match_oracle_energies(engine):
    req = golden_request("research_slump")
    spec, _, _ = research_slump()
    osim = O.Sim.from_fixture(spec)
    obs = observables_of(engine, req, osim)
    assert obs["kinetic_energy_j"] == osim.kinetic_energy()
    assert obs["potential_energy_j"] == osim.potential_energy()
    assert obs["elastic_energy_j"] == osim.elastic_energy()
    assert obs["mechanical_energy_j"] == osim.mechanical_energy()
    assert obs["particle_count"] == len(osim) == 512
    assert obs["step_count"] == 0 and obs["min_dt_s"] == 0.0
    assert obs["height_m"] == pytest.approx(0.02, abs=0.1 * 0.0025)
    # Pashias et al. 1996 estimate with the shear yield stress (fixtures.rs:510-527).
    tp = (30.0 * math.sqrt(3.0) / math.sqrt(3.0)) / (1000.0 * 9.81 * 0.02)
    assert obs["dimensionless_yield_stress"] == pytest.approx(tp, rel=1e-12)
    assert obs["pashias_slump_m"] == pytest.approx(0.02 * (1 - 2 * tp * (1 - math.log(2 * tp))), rel=1e-12)


@pytest.mark.slow
def test_slump_observables_reproduce_documented_energies(engine):
    # docs/validation.md:232-236 (slump, 500 ms): traction -9.7731734973e-6 J,
    # energy residual 7.3769595864e-5 J.  The residual is formed here from the
    # observables' own kinetic, potential and elastic energies.
    req = golden_request("research_slump")
    spec, duration, interval = research_slump()
    osim = O.Sim.from_fixture(spec)
    record_loop(osim, duration, interval)
    obs = observables_of(engine, req, osim)
    assert obs["step_count"] == 5000
    assert abs(obs["wall_normal_traction_energy_j"] - -9.7731734973e-6) < 1e-15
    assert abs(obs["energy_residual_j"] - 7.3769595864e-5) < 1e-14
    assert abs(obs["mass_residual_kg"]) < 1e-12
    assert obs["energy_residual_j"] == osim.energy_residual()


def test_hydrostatic_observables_reproduce_documented_errors(engine):
    # docs/hydrostatic-balance.md:26-31 and docs/validation.md:150-158, 232-236:
    # 4000-particle column, 10 ms: max 0.011513856585915002, rms 0.0034430221485745437,
    # energy residual 1.4254529794e-6 J, traction 1.0219846729e-8 J.
    req = golden_request("research_hydrostatic_water")
    osim = O.Sim.from_fixture(water_column([0.02, 0.02, 0.02], 0.01, 0.002, 5e-5))
    osim.advance_to(0.01)
    obs = observables_of(engine, req, osim)
    assert obs["particle_count"] == 4000 and obs["step_count"] == 200
    assert abs(obs["hydrostatic_max_relative_error"] - 0.011513856585915002) < 1e-15
    assert abs(obs["hydrostatic_rms_relative_error"] - 0.0034430221485745437) < 1e-15
    assert abs(obs["energy_residual_j"] - 1.4254529794e-6) < 1e-15
    assert abs(obs["wall_normal_traction_energy_j"] - 1.0219846729e-8) < 1e-18
    assert obs["mass_residual_kg"] == -8.673617379884035e-19
    assert abs(obs["momentum_residual_kg_m_s"] - 4.770490026058286e-18) < 1e-21
    # reference_bottom_stress_pa = K (exp(rho g h / K) - 1), fixtures.rs:354-366
    k = 1e5 / (3.0 * (1.0 - 2.0 * 0.2))
    assert obs["reference_bottom_stress_pa"] == pytest.approx(k * math.expm1(1000.0 * 9.81 * 0.01 / k), rel=1e-12)


@pytest.mark.slow
def test_coupled_patch_observables_reproduce_documented_energies(engine):
    # docs/validation.md:232-248 (coupled pellet, 50 ms): traction -3.4963973432e-6 J,
    # residual -7.1239144793e-6 J; the residual includes the pellet's kinetic and
    # potential energy as the observables compute them.
    req = golden_request("research_coupled_patch")
    spec, duration, interval = research_coupled_patch()
    osim = O.Sim.from_fixture(spec)
    record_loop(osim, duration, interval)
    obs = observables_of(engine, req, osim)
    assert abs(obs["wall_normal_traction_energy_j"] - -3.4963973432e-6) < 1e-15
    assert abs(obs["energy_residual_j"] - -7.1239144793e-6) < 1e-15
    p = osim.pellets()[0]
    assert obs["pellet_z_m"] == p.position[2]
    assert obs["pellet_penetration_m"] == max(0.02 + p.radius - p.position[2], 0.0)
    for key in ("coupling_impulse_z_kg_m_s", "pellet_wall_work_j", "coupling_grid_energy_j", "pellet_kinetic_energy_j",
                "pellet_axis_tilt_rad", "pellet_angular_speed_rad_s"):
        assert math.isfinite(obs[key]), key
