"""Batched research runs (mpm_run_json_batch) against run_json one job at a time
and against the f64 oracle.

Jobs in one batch share the engine's launches; each ends exactly as its own
run_json does -- frames, observables and checkpoint bit for bit, since an
instance's trajectory depends on its own state alone -- and a batched job
matches the oracle as a single run does.
"""
import copy
import json

import numpy as np
import pytest

import oracle as O
from test_research import golden_request, small_request
from test_sweep import sweep_module

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def same_run(a, b):
    assert a["status"] == b["status"]
    assert a["frames"] == b["frames"]
    assert a["observables"] == b["observables"]
    assert a["checkpoint"] == b["checkpoint"]


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


def test_batched_job_matches_the_oracle(engine):
    # The reference's research_slump request cut to 100 ms, as one job of a
    # three-job sweep over the yield stress; the job with the request's own
    # yield stress is checked against the oracle at the single-run bars.
    E = engine
    S = sweep_module(E)
    req = golden_request("research_slump")
    req["duration_s"] = 0.1
    base = req["research"]["yield_stress_pa"]
    jobs = S.expand_jobs(req, [("research.yield_stress_pa", [0.5 * base, base, 2.0 * base])], seeds=[0])
    batch = S.run_batched(E, [j[2] for j in jobs])
    assert all(b["ok"] for b in batch)
    out = batch[1]["output"]
    m = O.paste(1000.0, 1000.0, 0.2, 30.0, 10.0, 0.6, shear_units=True)
    cfg = O.config(wall_friction=0.4, cfl=0.3, max_dt=1e-4, min_dt=1e-12, stiffness=1000.0, restitution=0.2,
                   contact_friction=0.4)
    osim = O.Sim.from_fixture(O.fixture("slump", m, 0.005, [0.06] * 3, [0.02] * 3, cfg, seed=7))
    k, interval = 1, req["record_interval_s"]
    while osim.history().time < req["duration_s"]:
        osim.advance_to(min(k * interval, req["duration_s"]))
        k += 1
    h, obs = osim.history(), out["observables"]
    print("batched job steps", obs["step_count"], "oracle", h.step_count, flush=True)
    assert obs["step_count"] == h.step_count and obs["rejected_steps"] == h.rejected_steps
    st = out["checkpoint"]["state"]["particles"]
    o_gpu = np.argsort(np.array(st["id"]))
    ost = osim.state()
    o_ref = np.argsort(ost.id)
    assert rel(np.array(st["position"])[o_gpu], ost.pos[o_ref]) < 1e-4
    assert rel(np.array(st["velocity"])[o_gpu], ost.vel[o_ref]) < 1e-3
    led = out["checkpoint"]["state"]["ledger"]
    assert abs(led["plastic_dissipation"] - osim.ledger().plastic_dissipation) <= \
        1e-3 * osim.ledger().plastic_dissipation
    # The other two jobs differ from it (the sweep did vary the yield stress).
    assert batch[0]["output"]["checkpoint"] != out["checkpoint"] != batch[2]["output"]["checkpoint"]
