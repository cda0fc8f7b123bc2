"""Autotuner search logic (CPU, with a synthetic cost) and, on the GPU, a tuned
render that equals the default one bit for bit."""
import json

import numpy as np
import pytest

from ear_amd import autotune, scenes


def test_coordinate_descent_finds_separable_minimum():
    space = {"a": [1, 2, 3], "b": [10, 20], "c": [0, 5]}
    cost = {"a": {0: 9, 1: 3, 2: 1, 3: 2}, "b": {0: 5, 10: 4, 20: 0}, "c": {0: 1, 5: 2}}

    def measure(cfg):
        return sum(cost[k][cfg[k]] for k in cfg)

    best, t, trials = autotune.coordinate_descent(measure, space, {"a": 0, "b": 0, "c": 0}, rounds=2)
    assert best == {"a": 2, "b": 20, "c": 0} and t == 2
    # no configuration is measured twice
    keys = [json.dumps(c, sort_keys=True) for c, _ in trials]
    assert len(keys) == len(set(keys))


def test_coordinate_descent_keeps_start_when_nothing_is_better():
    best, t, trials = autotune.coordinate_descent(lambda c: 1.0 + c["x"], {"x": [1, 2]}, {"x": 0})
    assert best == {"x": 0} and t == 1.0 and len(trials) == 3


def test_load_rejects_unknown_version(tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"version": 99, "tuning": {}}))
    with pytest.raises(ValueError, match="version"):
        autotune.load(None, str(p))


@pytest.mark.gpu
def test_tuned_render_equals_default(gpu_ctx, tmp_path):
    from ear_amd import api

    s = scenes.room(seed=7, target_tris=3000)
    kw = dict(recorders=s["recorders"], src_pos=s["src"], air=s["air"], n_rays=20_000, max_bounces=80,
              cutoff=1e-7, n_bins=44100)
    res = autotune.tune(gpu_ctx, s["tris"], s["mat"], s["materials"], kw, sample_rays=4000, repeats=1)
    path = tmp_path / "tune.json"
    autotune.save(res, str(path))
    gpu_ctx.set_tuning()
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        ref = api.render(gpu_ctx, sc, **kw)
    opts = autotune.load(gpu_ctx, str(path))
    assert set(opts) == {"builder", "width", "leaf_size"}
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"], **opts) as sc:
        r = api.render(gpu_ctx, sc, **kw)
    gpu_ctx.set_tuning()
    np.testing.assert_array_equal(ref.raw, r.raw)
    assert res["segments_per_s"] > 0 and len(res["trials"]) >= 2
