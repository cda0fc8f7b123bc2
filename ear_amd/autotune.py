"""Autotuner for the launch configuration and the BVH shape.

Every knob here changes speed only: the histograms are integer sums, so a
render is bitwise identical under any setting (the GPU tests hold that).  The
tuner therefore times short renders of the caller's own workload and keeps the
fastest setting, by coordinate descent: starting from the library defaults it
tries each value of one knob with the others fixed, keeps the best, and moves
on to the next knob, `rounds` times over.  The result is a JSON sidecar that
`load` applies to a context (and that records the scene options to build with).
"""
from __future__ import annotations

import json
import time

# knob -> candidate values (0 = library default)
LAUNCH_SPACE = {
    "traversal": [1, 2, 3],
    "refill": [16, 24, 32, 48],
    "fetch_segments": [16, 64],
    "block_extend": [128, 256],
    "block_shadow": [128, 256],
    "stack_lds": [16, 32],
    "ray_chunk": [1 << 21, 1 << 22, 1 << 23],
    "graph_iters": [0, 8],
    "splat": [1, 2],
}
SCENE_SPACE = {
    "builder": [1, 2, 3, 4],
    "width": [2, 4],
    "leaf_size": [1, 2, 4],
}


def coordinate_descent(measure, space: dict, start: dict, rounds: int = 1):
    """Minimises measure(config) -> seconds over the product space, one
    coordinate at a time.  Returns (best config, best seconds, trials)."""
    best = dict(start)
    best_t = measure(best)
    trials = [(dict(best), best_t)]
    seen = {json.dumps(best, sort_keys=True): best_t}
    for _ in range(rounds):
        improved = False
        for knob, values in space.items():
            for v in values:
                if best.get(knob) == v:
                    continue
                cand = dict(best)
                cand[knob] = v
                key = json.dumps(cand, sort_keys=True)
                if key in seen:
                    t = seen[key]
                else:
                    t = measure(cand)
                    seen[key] = t
                    trials.append((dict(cand), t))
                if t < best_t:
                    best, best_t, improved = cand, t, True
        if not improved:
            break
    return best, best_t, trials


def _timed_render(ctx, scene, kw, repeats):
    from . import api

    api.render(ctx, scene, **kw)  # warm-up (first-touch allocations)
    ts = []
    for _ in range(repeats):
        r = api.render(ctx, scene, **kw)
        ts.append(r.wall_s)
    return min(ts), r.stats["segments"]


def tune(ctx, tris, mat, materials, render_kw: dict, *, sample_rays: int | None = None, repeats: int = 2,
         rounds: int = 1, tune_scene: bool = True) -> dict:
    """Tunes the BVH shape (width, leaf size) and then the launch knobs for
    the render `render_kw` (api.render keywords) of the given scene, timing
    renders of `sample_rays` rays per band (default: the render's own).
    Returns the sidecar dict; the context is left at the tuned setting."""
    from . import api

    kw = dict(render_kw)
    kw.setdefault("want_raw", False)
    if sample_rays:
        kw["n_rays"] = sample_rays
    ctx.set_tuning()
    scene_opts = {"builder": 4, "width": 4, "leaf_size": 2}
    scene_trials = []
    if tune_scene:
        def measure_scene(opts):
            if opts["builder"] in (3, 4) and opts["width"] != 4:
                return float("inf")  # the device PLOC and SAH builds make 4-wide trees only
            with api.Scene(ctx, tris, mat, materials, **opts) as sc:
                t, _ = _timed_render(ctx, sc, kw, repeats)
            return t

        scene_opts, _, scene_trials = coordinate_descent(measure_scene, SCENE_SPACE, scene_opts, rounds)
    with api.Scene(ctx, tris, mat, materials, **scene_opts) as sc:
        segs = {}

        def measure_launch(cfg):
            ctx.set_tuning(**cfg)
            t, n = _timed_render(ctx, sc, kw, repeats)
            segs[json.dumps(cfg, sort_keys=True)] = n
            return t

        start = {k: 0 for k in LAUNCH_SPACE}
        best, best_t, trials = coordinate_descent(measure_launch, LAUNCH_SPACE, start, rounds)
        n_seg = segs[json.dumps(best, sort_keys=True)]
    ctx.set_tuning(**best)
    return {
        "version": 1,
        "created": time.strftime("%Y-%m-%dT%H:%M:%S"),
        "workload": {"n_tris": int(len(tris)), "rays_per_band": int(kw["n_rays"]),
                     "max_bounces": int(kw.get("max_bounces", 999)), "bands": list(kw.get("bands", (0, 1, 2)))},
        "scene": scene_opts,
        "tuning": best,
        "seconds": best_t,
        "segments_per_s": n_seg / best_t if best_t > 0 else None,
        "trials": [{"config": c, "seconds": t} for c, t in scene_trials + trials],
    }


def save(result: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load(ctx, path: str) -> dict:
    """Applies a sidecar's launch tuning to ctx; returns its scene options
    (builder, width, leaf_size) for api.Scene."""
    with open(path) as f:
        d = json.load(f)
    if d.get("version") != 1:
        raise ValueError(f"{path}: unknown autotune sidecar version {d.get('version')}")
    known = set(LAUNCH_SPACE) | {"block_shade", "waves_per_cu", "fetch_chunk"}
    ctx.set_tuning(**{k: int(v) for k, v in d["tuning"].items() if k in known})
    return dict(d.get("scene", {}))
