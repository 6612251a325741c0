"""Autotuner for the launch shapes of plans (and the single-mesh drivers).

The tile kernels are bound by FP64 (or FP32) VALU issue; what the launch
shape changes is how well the persistent grid hides LDS and memory latency
(workgroups per CU, against the kernels' register use) and how much per-tile
overhead (staging, the ragged last chunk) each tile carries (cells per tile).
The best shape depends on the mesh sizes, the point counts and the
precision, so it is measured, not modelled.

`tune()` runs a coordinate descent over the given tunables: for each key in
turn, every candidate value is tried with the other keys at their current
best.  Each trial builds the plan with the context's tunables set, runs it
once untimed and `repeats` times timed with the library's HIP-event timer,
and keeps the fastest.  Results are cached per (compute units, workload
signature) in a JSON file, so a workload seen before is not tuned again.
"""
from __future__ import annotations

import json
import os
import time
from typing import Callable, Optional

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "bielcc_amd", "tune.json")

FIELD_SPACE = {"plan_wg_per_cu": [1, 2, 4, 8], "plan_field_cells": [32, 64, 128, 256]}
ASSEMBLE_SPACE = {"plan_wg_per_cu": [1, 2, 4, 8], "plan_asm_cells": [32, 64, 128, 256]}


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path: str, data: dict) -> None:
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + f".tmp{os.getpid()}"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def measure(ctx, build: Callable, timer: str, repeats: int) -> float:
    """Mean milliseconds of one run of the plan build() returns (HIP events
    around each run on the context's stream)."""
    plan = build()
    try:
        plan.run()
        ctx.sync()
        ctx.timing_enable(True)
        ctx.timing_reset()
        for _ in range(repeats):
            plan.run()
        ctx.sync()
        ms, cnt = ctx.timing(timer)
        ctx.timing_enable(False)
        return ms / max(cnt, 1)
    finally:
        plan.close()


def tune(ctx, build: Callable, space: dict, timer: str, signature: str, repeats: int = 2,
         cache_path: Optional[str] = DEFAULT_CACHE, measure_fn: Optional[Callable] = None, log=None) -> dict:
    """Pick the tunables in `space` (key -> candidate values) that minimise the
    plan's run time.  Sets the winners on ctx and returns
    {"best": {key: value}, "ms": best time, "trials": [(candidate, ms)], "cached": bool}."""
    measure_fn = measure_fn or (lambda: measure(ctx, build, timer, repeats))
    n_cu = ctx.get_tuning("n_cu")
    key = f"{n_cu}cu|{signature}"
    if cache_path:
        hit = _load(cache_path).get(key)
        if hit and set(hit["best"]) == set(space):
            for k, v in hit["best"].items():
                ctx.set_tuning(k, int(v))
            return {"best": dict(hit["best"]), "ms": hit["ms"], "trials": [], "cached": True}
    best = {k: ctx.get_tuning(k) for k in space}
    for k in best:
        ctx.set_tuning(k, best[k])
    trials = []
    best_ms = None
    t0 = time.perf_counter()
    for k, values in space.items():
        for v in values:
            cand = dict(best)
            cand[k] = v
            if best_ms is not None and cand == best:
                continue
            for kk, vv in cand.items():
                ctx.set_tuning(kk, int(vv))
            ms = float(measure_fn())
            trials.append((dict(cand), ms))
            if log:
                log(f"[tune] {signature}: {cand} -> {ms:.3f} ms")
            if best_ms is None or ms < best_ms:
                best_ms, best = ms, dict(cand)
        for kk, vv in best.items():
            ctx.set_tuning(kk, int(vv))
    out = {"best": best, "ms": best_ms, "trials": trials, "cached": False,
           "tuning_s": round(time.perf_counter() - t0, 3)}
    if cache_path:
        data = _load(cache_path)
        data[key] = {"best": best, "ms": best_ms}
        _store(cache_path, data)
    return out
