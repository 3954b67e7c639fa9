"""Sweep and refinement-study expansion (python/litter_physics/sweep.py,
verification.py), and the batched run's handling of invalid jobs, on the CPU."""
import importlib.util
import json
import os

import pytest

from test_research import small_request

HERE = os.path.dirname(os.path.abspath(__file__))


def sweep_module(engine):
    spec = importlib.util.spec_from_file_location(
        "litter_physics_amd.sweep", os.path.join(os.path.dirname(engine.__file__), "sweep.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_set_by_path_errors(engine):
    # tests/test_sweep_benchmark.py:24-36 of the reference
    S = sweep_module(engine)
    tree = {"a": {"b": [1, 2]}, "c": 1}
    S.set_by_path(tree, "a.b.1", 5)
    assert tree["a"]["b"] == [1, 5]
    with pytest.raises(KeyError):
        S.set_by_path(tree, "a.zzz", 1)
    with pytest.raises(TypeError):
        S.set_by_path(tree, "a.b.x", 1)
    with pytest.raises(TypeError):
        S.set_by_path(tree, "c.d", 1)
    with pytest.raises(IndexError):
        S.set_by_path(tree, "a.b.9", 1)


def test_expand_jobs_is_the_grid_product_times_seeds(engine):
    S = sweep_module(engine)
    base = small_request()
    axes = [("materials.friction", [0.1, 0.3]), ("research.yield_stress_pa", [10.0, 30.0])]
    jobs = S.expand_jobs(base, axes, seeds=[0, 1])
    assert len(jobs) == 8
    assert [j[1] for j in jobs] == [3, 4] * 4
    assert [tuple(j[0].values()) for j in jobs[::2]] == [(0.1, 10.0), (0.1, 30.0), (0.3, 10.0), (0.3, 30.0)]
    assert jobs[3][2]["research"]["yield_stress_pa"] == 30.0 and jobs[3][2]["seed"] == 4
    assert base["materials"]["friction"] == 0.3  # the base is not modified
    for _, _, req in jobs:
        engine.research_validate(json.dumps(req))


def test_refinement_cases_are_spatial_then_temporal(engine):
    S = sweep_module(engine)
    cases = S.refinement_cases(small_request(), spacings=[0.005, 0.0025], spatial_dt=5e-4, dts=[5e-4, 2.5e-4],
                               temporal_spacing=0.005)
    assert [(c[0], c[1]) for c in cases] == [("spatial", 0), ("spatial", 1), ("temporal", 0), ("temporal", 1)]
    assert cases[1][4]["research"]["grid_spacing_m"] == 0.0025 and cases[3][4]["dt_s"] == 2.5e-4
    with pytest.raises(ValueError):
        S.refinement_cases({"mode": "household"})


def test_batch_of_invalid_jobs_reports_each_without_a_device(engine):
    bad1 = small_request()
    bad1["research"]["material"] = "fines"
    bad2 = small_request()
    bad2["unknown"] = 1
    out = engine.run_json_batch([json.dumps(bad1), json.dumps(bad2)])
    assert [o["ok"] for o in out] == [False, False]
    assert "fines" in out[0]["error"] and "unknown field" in out[1]["error"]
    with pytest.raises(ValueError, match="invalid batch JSON"):
        engine.run_json_batch(["{"])
    with pytest.raises(ValueError, match="one checkpoint or null per request"):
        engine.run_json_batch([json.dumps(bad1)], [None, None])
