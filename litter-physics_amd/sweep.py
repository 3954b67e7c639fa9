"""Parameter sweeps and refinement studies as one batched run.

The reference expands a base research request into jobs by dotted-path
overrides (python/litter_physics/sweep.py:75-205, verification.py:497-547) and
runs them one after another through run_json.  Here the same expansion feeds
run_json_batch, which steps every job as an instance of one batch engine.
"""
from __future__ import annotations

import copy
import itertools
import json
from typing import Any

MAX_SWEEP_JOBS = 10_000
GRID_SPACING_PATH = "research.grid_spacing_m"
DT_PATH = "dt_s"


def _index(part: str, path: str) -> int:
    if not part.isdigit():
        raise TypeError(f"path '{path}' indexes a list with non-integer '{part}'")
    return int(part)


def set_by_path(payload: dict[str, Any], path: str, value: Any) -> None:
    """Assign into a nested JSON tree by dotted path; integer segments index lists
    (sweep.py:75-130).  Raises KeyError, IndexError or TypeError like the reference."""
    parts = path.split(".")
    node: Any = payload
    for part in parts[:-1]:
        if isinstance(node, list):
            node = node[_index(part, path)]
        elif isinstance(node, dict):
            if part not in node:
                raise KeyError(f"path '{path}' names unknown key '{part}'")
            node = node[part]
        else:
            raise TypeError(f"path '{path}' descends into a scalar at '{part}'")
    last = parts[-1]
    if isinstance(node, list):
        node[_index(last, path)] = value
    elif isinstance(node, dict):
        if last not in node:
            raise KeyError(f"path '{path}' names unknown key '{last}'")
        node[last] = value
    else:
        raise TypeError(f"path '{path}' descends into a scalar at '{last}'")


def expand_jobs(base: dict[str, Any], axes: list[tuple[str, list[Any]]], seeds: list[int] = (0,)):
    """Grid product of the axes times the seed offsets (sweep.py:170-205):
    [(overrides, seed, request)], in the reference's order."""
    combos = list(itertools.product(*[values for _, values in axes]))
    total = len(combos) * len(seeds)
    if total > MAX_SWEEP_JOBS:
        raise ValueError(f"sweep expands to {total} jobs, above the cap {MAX_SWEEP_JOBS}")
    jobs = []
    for combo in combos:
        for offset in seeds:
            payload = copy.deepcopy(base)
            overrides = {path: value for (path, _), value in zip(axes, combo)}
            for path, value in overrides.items():
                set_by_path(payload, path, value)
            seed = int(base["seed"]) + offset
            payload["seed"] = seed
            jobs.append((overrides, seed, payload))
    return jobs


def refinement_cases(base: dict[str, Any], spacings=(), spatial_dt=None, dts=(), temporal_spacing=None):
    """Spatial cases then temporal cases, coarse to fine (verification.py:497-547):
    [(axis, index, grid_spacing_m, dt_s, request)]."""
    if base.get("mode") != "research" or not isinstance(base.get("research"), dict):
        raise ValueError("the base request must be a research request")
    pairs = [("spatial", i, h, spatial_dt) for i, h in enumerate(spacings)]
    pairs += [("temporal", i, temporal_spacing, dt) for i, dt in enumerate(dts)]
    cases = []
    for axis, index, h, dt in pairs:
        payload = copy.deepcopy(base)
        set_by_path(payload, GRID_SPACING_PATH, h)
        set_by_path(payload, DT_PATH, dt)
        cases.append((axis, index, h, dt, payload))
    return cases


def run_batched(engine, requests: list[dict[str, Any]], resumes=None) -> list[dict[str, Any]]:
    """Run every request as one batch (engine: the litter_physics_amd module);
    one {"ok", "output" | "error"} dict per request."""
    texts = [json.dumps(r) for r in requests]
    res = None if resumes is None else [None if r is None else json.dumps(r) for r in resumes]
    return engine.run_json_batch(texts, res)
