"""Autotuner logic on the host: coordinate descent, the cache and its key."""
import bielcc_amd.tune as T


class FakeCtx:
    def __init__(self, n_cu=256):
        self.t = {"plan_wg_per_cu": 4, "plan_field_cells": 64, "n_cu": n_cu}

    def get_tuning(self, k):
        return self.t[k]

    def set_tuning(self, k, v):
        self.t[k] = v


def _cost(ctx):
    # minimum at wg = 2, cells = 128; separable, so one descent pass finds it
    wg, c = ctx.t["plan_wg_per_cu"], ctx.t["plan_field_cells"]
    return 10.0 + (wg - 2) ** 2 + ((c - 128) / 32.0) ** 2


def test_descent_finds_the_minimum_and_sets_it(tmp_path):
    ctx = FakeCtx()
    path = str(tmp_path / "tune.json")
    res = T.tune(ctx, None, T.FIELD_SPACE, "plan_field", "sig", cache_path=path, measure_fn=lambda: _cost(ctx))
    assert res["best"] == {"plan_wg_per_cu": 2, "plan_field_cells": 128}
    assert res["ms"] == 10.0 and not res["cached"]
    assert ctx.t["plan_wg_per_cu"] == 2 and ctx.t["plan_field_cells"] == 128
    # every candidate of every key was tried once, the shared one not twice
    assert len(res["trials"]) == 4 + 3


def test_cache_hit_skips_measuring(tmp_path):
    path = str(tmp_path / "tune.json")
    ctx = FakeCtx()
    T.tune(ctx, None, T.FIELD_SPACE, "plan_field", "sig", cache_path=path, measure_fn=lambda: _cost(ctx))
    ctx2 = FakeCtx()
    calls = []
    res = T.tune(ctx2, None, T.FIELD_SPACE, "plan_field", "sig", cache_path=path,
                 measure_fn=lambda: calls.append(1) or 1.0)
    assert res["cached"] and not calls
    assert ctx2.t["plan_wg_per_cu"] == 2 and ctx2.t["plan_field_cells"] == 128
    # another device size or workload signature is a miss
    ctx3 = FakeCtx(n_cu=304)
    res3 = T.tune(ctx3, None, T.FIELD_SPACE, "plan_field", "sig", cache_path=path, measure_fn=lambda: _cost(ctx3))
    assert not res3["cached"]
    res4 = T.tune(ctx2, None, T.FIELD_SPACE, "plan_field", "other", cache_path=path, measure_fn=lambda: _cost(ctx2))
    assert not res4["cached"]
