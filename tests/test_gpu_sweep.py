"""Batched research runs (mpm_run_json_batch) against run_json one job at a time.

Jobs in one batch share the engine's launches; each must end exactly as its
own run_json would (same status, frames, step/limiter/rejection counts), with
states equal to the f32 atomic round-off of the grid sums (1e-6 relative).
"""
import copy
import json

import numpy as np
import pytest

from test_research import golden_request, small_request
from test_sweep import sweep_module

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def same_run(a, b):
    assert a["status"] == b["status"]
    assert [f["time_s"] for f in a["frames"]] == pytest.approx([f["time_s"] for f in b["frames"]], rel=1e-12)
    for key in ("step_count", "limited_steps", "rejected_steps", "particle_count"):
        assert a["observables"][key] == b["observables"][key], key
    pa, pb = a["checkpoint"]["state"]["particles"], b["checkpoint"]["state"]["particles"]
    assert pa["id"] == pb["id"] and pa["mass"] == pb["mass"]
    for key in ("position", "velocity", "kirchhoff"):
        assert rel(pa[key], pb[key]) < 1e-6, key
    assert a["checkpoint"]["state"]["ledger"]["initial_mechanical_energy"] == \
        b["checkpoint"]["state"]["ledger"]["initial_mechanical_energy"]


def test_batched_sweep_matches_run_json_per_job(engine):
    E = engine
    S = sweep_module(E)
    jobs = S.expand_jobs(small_request(), [("research.yield_stress_pa", [10.0, 30.0, 60.0])], seeds=[0, 1])
    requests = [j[2] for j in jobs]
    bad = small_request()
    bad["research"]["poisson_ratio"] = 0.5
    starved = small_request()
    starved["max_wall_time_s"] = 1e-12
    batch = S.run_batched(E, requests + [bad, starved])
    print("batch:", [b["ok"] and b["output"]["status"] for b in batch], flush=True)
    assert all(b["ok"] for b in batch[:6])
    for req, got in zip(requests, batch[:6]):
        same_run(got["output"], json.loads(E.run_json(json.dumps(req))))
    assert not batch[6]["ok"]
    with pytest.raises(ValueError) as err:
        E.run_json(json.dumps(bad))
    assert batch[6]["error"] == str(err.value)
    assert batch[7]["ok"] and batch[7]["output"]["status"] == "budget_exhausted"
    assert len(batch[7]["output"]["frames"]) == 1
    # Resume the starved job inside a new batch, beside a fresh job.
    starved["max_wall_time_s"] = 60.0
    again = S.run_batched(E, [starved, requests[0]], [batch[7]["output"]["checkpoint"], None])
    assert again[0]["ok"] and again[0]["output"]["status"] == "completed"
    # A resumed segment's frames are the suffix after the checkpoint, as run_json gives them.
    same_run(again[0]["output"],
             json.loads(E.run_json(json.dumps(starved), json.dumps(batch[7]["output"]["checkpoint"]))))
    same_run(again[1]["output"], batch[0]["output"])


def test_batched_refinement_study(engine):
    E = engine
    S = sweep_module(E)
    base = golden_request("research_slump")
    base["duration_s"] = 0.02
    cases = S.refinement_cases(base, spacings=[0.005, 0.0025], spatial_dt=1e-4, dts=[2e-4, 1e-4],
                               temporal_spacing=0.005)
    out = S.run_batched(E, [c[4] for c in cases])
    assert all(o["ok"] for o in out)
    for case, got in zip(cases, out):
        single = json.loads(E.run_json(json.dumps(case[4])))
        same_run(got["output"], single)
        print(case[0], case[2], case[3], "steps", got["output"]["observables"]["step_count"], flush=True)
