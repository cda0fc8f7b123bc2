#!/usr/bin/env python3
"""Benchmark: ray-segments/s of the MI355X tracer on the 200k-triangle hall.

Workload (one step = one full render, i.e. Scene::Render of src/Scene.cpp:111-318
for the three band contexts of one sound, fed by synthetic seeded input):
  procedural concert hall of ~200k fp32 triangles, 8 materials x 3 bands with
  alpha ~ U(0.02, 0.6), one point source, stereo recorder (head 0.17 m),
  10M rays per band, <= 100 segments per ray, intensity cut-off 1e-8,
  3 s histograms at 44.1 kHz (132,300 bins per channel and band).
The scene and its BVH (built on the device) are resident in HBM before the timed region; each step
emits, traces and splats every ray and finalizes the histograms.

Prints one JSON line (rank 0, N=1).  `rooflines` prices each bounce-loop
kernel (k_extend_f, k_shade, k_shadow_f) at algorithmic bytes per unit of work
over its HIP-event time (kernel_rooflines below), with node visits V and
triangle tests T per query counted by an instrumented pass over the same rays;
`roofline` is the dominant one (the largest time per render), and
`render_roofline` the whole render at bytes per segment.  `traffic` is
rocprofv3's FETCH_SIZE + WRITE_SIZE per launch where profiles/ holds it.
`cpu_baseline` is the CPU oracle (a brute-force restatement of the reference
algorithm) timed on this host on the first K rays of the same Philox streams.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rays", type=int, default=10_000_000, help="rays per band")
    ap.add_argument("--tris", type=int, default=200_000)
    ap.add_argument("--max-bounces", type=int, default=100)
    ap.add_argument("--cutoff", type=float, default=1e-8)
    ap.add_argument("--seconds", type=float, default=3.0, help="histogram length")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-instrument", action="store_true")
    ap.add_argument("--no-checks", action="store_true",
                    help="skip the 4x-ray accuracy render and the graph replay (counter runs: every render of the "
                         "main leg is then a default one)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batch (256 rooms), forward + adjoint, animated-hall and fitting legs")
    ap.add_argument("--legs", default="main,batch,adjoint,animated,fit",
                    help="comma list of legs to run: main (the hall), batch, adjoint, animated, fit "
                         "(for profiling one leg)")
    ap.add_argument("--splat", choices=["default", "atomic", "tiled"], default="default",
                    help="the shadow kernels' splat (ear_tuning.splat) for the main leg")
    ap.add_argument("--stack-lds", type=int, default=0, help="traversal stack entries in LDS (16 or 32; 0: default)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (PMC calibration) and exit")
    ap.add_argument("--traffic-dir", default=os.path.join(ROOT, "profiles"),
                    help="folder of traffic_<kernel>_hall200k.json (rocprofv3 FETCH_SIZE + WRITE_SIZE per launch)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_probe(nbytes: int) -> None:
    """A plain device-to-device copy of nbytes (HIP's blit kernel), for scaling
    the FETCH_SIZE / WRITE_SIZE counters: it reads and writes exactly nbytes."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this tracer runs on one GPU; --gpus must be 1 (sharding the ray ranges "
                                   "over GPUs, with one integer all-reduce of the histograms, is not built)"}))
        sys.exit(2)
    import numpy as np

    from ear_amd import api, scenes

    legs = set(a.legs.split(","))
    if "main" not in legs:
        ctx = api.Context(0)
        extras = {}
        if "batch" in legs:
            extras["batch"] = bench_batch(ctx)
        if "adjoint" in legs:
            extras["adjoint"] = bench_adjoint(ctx)
        if "animated" in legs:
            extras["animated"] = bench_animated(ctx)
        if "fit" in legs:
            extras["fit"] = bench_fit(ctx)
        ctx.close()
        print(json.dumps({"legs": sorted(legs), "extras": extras}), flush=True)
        return

    t0 = time.time()
    s = scenes.hall(target_tris=a.tris)
    n_bins = int(round(44100 * a.seconds))
    log(f"scene {s['name']} generated in {time.time() - t0:.1f}s")
    ctx = api.Context(0)
    sc = api.Scene(ctx, s["tris"], s["mat"], s["materials"])
    info = sc.info()
    log(f"BVH: {info['n_nodes']} nodes, depth {info['max_depth']}, build {info['build_ms']:.2f} ms")
    kw = dict(recorders=s["recorders"], n_rays=a.rays, src_pos=s["src"], air=s["air"], max_bounces=a.max_bounces,
              cutoff=a.cutoff, n_bins=n_bins, abs_hist=True, want_raw=False, want_value=True)
    ctx.set_tuning(splat={"default": 0, "atomic": 1, "tiled": 2}[a.splat], stack_lds=a.stack_lds)

    for i in range(a.warmup):
        r = api.render(ctx, sc, **kw)
        log(f"warmup {i}: {r.wall_s:.3f}s, {r.stats['segments']} segments")
    walls, segs = [], 0
    # per-kernel sums over the timed renders (HIP events around every launch)
    acc = {k: 0 for k in ("extend_ms", "shade_ms", "shadow_ms", "splat_ms", "trace_ms", "extend_queries",
                          "shadow_queries", "extend_launches", "shadow_launches", "shadow_contributions",
                          "shadow_ramps", "segments", "rays")}
    t_start = time.perf_counter()
    for i in range(a.steps):
        r = api.render(ctx, sc, **kw)
        walls.append(r.wall_s)
        segs += r.stats["segments"]
        for k in acc:
            acc[k] += r.stats[k]
        log(f"step {i}: {r.wall_s:.3f}s, {r.stats['segments']} segments, trace {r.stats['trace_ms']:.1f} ms "
            f"(extend {r.stats['extend_ms']:.1f}, shade {r.stats['shade_ms']:.1f}, shadow {r.stats['shadow_ms']:.1f})")
    elapsed = time.perf_counter() - t_start
    value = segs / elapsed

    # EDC / T60 of the last render (unsigned mid band, left channel)
    from ear_amd import acoustics

    edc_t60 = None
    try:
        edc_t60 = acoustics.t60_iso(r.abs_value[1, 0])
    except Exception as e:  # pragma: no cover - reporting only
        log(f"t60 failed: {e}")
    # accuracy: the EDC (every band and ear, to -60 dB) and the mid band's T60
    # against a render of 4x the rays on an independent random stream (whose
    # own error is about half this render's)
    accuracy = None
    try:
        if a.no_checks:
            raise RuntimeError("skipped (--no-checks)")
        r4 = api.render(ctx, sc, **dict(kw, n_rays=4 * a.rays, seed=0x5EED))
        edc_err = max(acoustics.edc_error(r.abs_value[b, c], r4.abs_value[b, c]) for b in range(3)
                      for c in range(r.abs_value.shape[1]))
        t60_b = acoustics.t60_iso(r4.abs_value[1, 0])
        accuracy = {"edc_error_vs_4x_rays": round(edc_err, 5), "t60_mid_s_4x_rays": t60_b,
                    "t60_rel_error_vs_4x_rays": round(abs(edc_t60 / t60_b - 1.0), 5)}
        log(f"accuracy: EDC error against 4x the rays {edc_err:.4f}, T60 {edc_t60:.4f} / {t60_b:.4f} s")
    except Exception as e:  # pragma: no cover - reporting only
        log(f"accuracy check failed: {e}")
    # the bounce loop replayed as a captured HIP graph (same histograms)
    graph = None
    try:
        if a.no_checks:
            raise RuntimeError("skipped (--no-checks)")
        ctx.set_tuning(splat={"default": 0, "atomic": 1, "tiled": 2}[a.splat], stack_lds=a.stack_lds, graph_iters=8)
        api.render(ctx, sc, **kw)
        t0 = time.perf_counter()
        rg = api.render(ctx, sc, **kw)
        graph = {"graph_iters": 8, "ms_per_render": round(1e3 * (time.perf_counter() - t0), 3),
                 "segments": int(rg.stats["segments"])}
        ctx.set_tuning(splat={"default": 0, "atomic": 1, "tiled": 2}[a.splat], stack_lds=a.stack_lds)
        log(f"graph replay: {graph['ms_per_render']} ms per render")
    except Exception as e:  # pragma: no cover - reporting only
        log(f"graph replay failed: {e}")

    roof = rooflines = render_roof = None
    if not a.no_instrument:
        ri = api.render(ctx, sc, instrument=True, **kw)
        rooflines, render_roof = kernel_rooflines(acc, ri.stats, info, a.traffic_dir, a.steps)
        # the dominant kernel (largest share of the timed renders' kernel time)
        dom = max(rooflines, key=lambda k: rooflines[k]["ms_per_render"])
        roof = dict(rooflines[dom], kernel=dom)
        log(f"dominant kernel {dom}: {roof['achieved']:.0f} GB/s algorithmic, frac {roof['frac']:.3f}")

    ctx.set_tuning()  # the other legs run with the library defaults
    extras = None
    if not a.no_extras and legs & {"batch", "adjoint", "animated", "fit"}:
        extras = {}
        if "batch" in legs:
            extras["batch"] = bench_batch(ctx)
        if "adjoint" in legs:
            extras["adjoint"] = bench_adjoint(ctx)
        if "animated" in legs:
            extras["animated"] = bench_animated(ctx)
        if "fit" in legs:
            extras["fit"] = bench_fit(ctx)

    cpu = cpu_ref = None
    if not a.no_cpu:
        cpu = cpu_baseline(s, kw, a.cpu_seconds)
        cpu_ref = cpu_reference_threading(s, kw, 200)

    sc.close()
    ctx.close()
    out = {
        "metric": "ray-segments/s",
        "value": round(value, 1),
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * elapsed / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"hall {info['n_tris']} tris, stereo, 3 bands x {a.rays} rays, "
                               f"<= {a.max_bounces} segments, cutoff {a.cutoff:g}, {a.seconds:g} s histograms",
                   "n_tris": info["n_tris"], "rays_per_band": a.rays, "bands": 3, "max_bounces": a.max_bounces,
                   "hist_bins": n_bins, "parallelism": "single GPU"},
        "segments_per_step": segs // max(1, a.steps),
        "bvh_build_ms": round(info["build_ms"], 3),
        "t60_mid_s": edc_t60,
        "accuracy": accuracy,
        "graph_replay": graph,
        "roofline": roof,
        "rooflines": rooflines,
        "render_roofline": render_roof,
        "cpu_baseline": cpu,
        "cpu_baseline_reference_threading": cpu_ref,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)


def kernel_rooflines(acc, ist, info, traffic_dir, steps):
    """Algorithmic-byte rooflines of the bounce loop's kernels over the timed
    renders (acc: summed ear_stats), with node visits V and triangle tests T
    per query from the instrumented render `ist`, and the whole render's
    bytes per segment.  Bytes per unit:
      k_extend_f per closest-hit query: 24 (o, d) + 8 (hit) + B V + 36 T
      k_shade per shaded segment: 48 (queue entry) + 8 (hit) + 16 (normal and
        material) + 16 (material row) + 48 x (segments that continue) + 32 x
        (shadow requests written)
      k_shadow_f per shadow query: 32 (request) + B V_sh + 36 T_sh, plus per
        ramp 3 x 16 (128-bit points) with direct atomics, or with the
        time-tiled splat 4 per record slot + 8 per record
    B = 64 bytes per node visit.  `traffic` is rocprofv3's FETCH_SIZE +
    WRITE_SIZE per launch of that kernel (scaled by a plain copy), where
    profiles/ holds it for the kernel; measured_frac = traffic per average
    launch time over the peak."""
    B = info["node_bytes"]
    qe, qs = max(1, acc["extend_queries"]), max(1, acc["shadow_queries"])
    V = ist["node_visits_closest"] / max(1, ist["extend_queries"])
    T = ist["tri_tests_closest"] / max(1, ist["extend_queries"])
    Vs = ist["node_visits_shadow"] / max(1, ist["shadow_queries"])
    Ts = ist["tri_tests_shadow"] / max(1, ist["shadow_queries"])
    ramps_per_q = acc["shadow_ramps"] / qs
    tiled = acc["splat_ms"] > 0
    stride = 2 if acc["shadow_ramps"] > 1.5 * acc["shadow_contributions"] else 1
    cont = 1.0 - acc["rays"] / max(1, acc["segments"])  # shaded segments whose ray goes on
    per = {
        "k_extend_f": (24 + 8 + B * V + 36 * T, qe, acc["extend_ms"], acc["extend_launches"], "closest-hit query"),
        "k_shade": (48 + 8 + 16 + 16 + 48 * cont + 32 * qs / qe, qe, acc["shade_ms"], acc["extend_launches"],
                    "shaded segment"),
        "k_shadow_f": (32 + B * Vs + 36 * Ts + (4 * stride + 8 * ramps_per_q if tiled else 48 * ramps_per_q), qs,
                       acc["shadow_ms"], acc["shadow_launches"], "shadow query"),
    }
    out = {}
    for k, (bpu, units, ms, launches, unit) in per.items():
        achieved = bpu * units / max(1e-9, ms * 1e-3) / 1e9
        avg_ms = ms / max(1, launches)
        traffic = None
        tf = os.path.join(traffic_dir, f"traffic_{k}_hall200k.json")
        if os.path.exists(tf):
            with open(tf) as f:
                traffic = json.load(f).get("bytes_per_launch")
        out[k] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                  "measured_frac": (round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None),
                  "bytes_per_unit": round(bpu, 1), "unit_of_work": unit,
                  "units_per_launch": round(units / max(1, launches), 1), "avg_launch_ms": round(avg_ms, 4),
                  "launches": launches, "ms_per_render": round(ms / steps, 2)}
    out["k_extend_f"].update(node_visits_per_query=round(V, 3), tri_tests_per_query=round(T, 3))
    out["k_shadow_f"].update(node_visits_per_query=round(Vs, 3), tri_tests_per_query=round(Ts, 3),
                             ramps_per_query=round(ramps_per_q, 4), splat="tiled" if tiled else "atomic")
    if tiled:
        out["k_tile"] = {"ms_per_render": round(acc["splat_ms"] / steps, 2),
                         "note": "k_tile_count + k_tile_scan + k_tile_scatter + k_tile_splat, per bounce iteration"}
    # the whole render (all kernels, trace_ms): bytes per segment with every
    # shadow query and ramp of the render
    seg = max(1, acc["segments"])
    b_seg = (96 + 20 + B * (V + Vs * qs / seg) + 36 * (T + Ts * qs / seg) + 48 * acc["shadow_ramps"] / seg)
    ach = b_seg * seg / max(1e-9, acc["trace_ms"] * 1e-3) / 1e9
    render_roof = {"bytes_per_segment": round(b_seg, 1), "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                   "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                   "note": "96 ray state + 20 material + 64 (V + V_sh) + 36 (T + T_sh) + 48 per ramp, per segment"}
    return out, render_roof


def bench_batch(ctx, n_rooms=256, seed=3):
    """Batching: n_rooms box or L rooms of 12..50k triangles (log-uniform),
    10k..2M rays each (log-uniform), 3 bands, 1 s histograms, in one
    ear_render_batch, beside the same renders one call each."""
    from ear_amd import api, scenes

    t0 = time.time()
    specs = scenes.batch_rooms(n_rooms, seed=seed)
    log(f"batch: {n_rooms} rooms generated in {time.time() - t0:.1f}s")
    t0 = time.perf_counter()
    scs = [api.Scene(ctx, rm["tris"], rm["mat"], rm["materials"]) for rm, _ in specs]
    upload_ms = 1e3 * (time.perf_counter() - t0)
    items = [(sc, dict(recorders=rm["recorders"], src_pos=rm["src"], air=rm["air"], n_rays=rays, max_bounces=999,
                       cutoff=1e-8, n_bins=44100, instance=i, abs_hist=False, want_raw=False))
             for i, (sc, (rm, rays)) in enumerate(zip(scs, specs))]
    api.render_batch(ctx, items[:16])  # warm-up (first-touch allocations)
    _, st, wall = api.render_batch(ctx, items)
    log(f"batch: {st['segments']} segments in {wall:.3f}s ({st['segments'] / wall:.3e} segments/s)")
    t0 = time.perf_counter()
    segs1 = 0
    for sc, kw in items:
        segs1 += api.render(ctx, sc, **kw).stats["segments"]
    single = time.perf_counter() - t0
    log(f"batch: one call each {single:.3f}s ({segs1 / single:.3e} segments/s)")
    for sc in scs:
        sc.close()
    return {"workload": f"{n_rooms} box/L rooms, 12..50k triangles, 1e4..2e6 rays each, 3 bands, 1 s histograms",
            "triangles": int(sum(len(rm["tris"]) for rm, _ in specs)),
            "rays": int(3 * sum(r for _, r in specs)), "segments": int(st["segments"]),
            "upload_ms": round(upload_ms, 1), "batch_s": round(wall, 4),
            "batch_segments_per_s": round(st["segments"] / wall, 1), "single_calls_s": round(single, 4),
            "single_calls_segments_per_s": round(segs1 / single, 1), "speedup": round(single / wall, 3)}


def bench_adjoint(ctx, n_rooms=32, rays_accurate=4_000_000):
    """Differentiation: the 12 x 9 x 5 m shoebox (6 walls x 3 bands
    absorption + 6 specularities), 1M rays, 4096-bin log-EDC loss against a
    target rendered at the true parameters, evaluated at parameters off by
    +-20%: forward, adjoint (the score form; the local form timed beside it)
    and their rate beside the forward alone; the EDC
    error of the 1M-ray forward against 8M rays; then n_rooms rooms of
    1k..100k triangles through the batched forward and adjoint at 250k rays
    each and at rays_accurate each (the accuracy point: the EDC error of the
    rooms least accurate at 250k rays against 4x the rays)."""
    import numpy as np

    from ear_amd import acoustics, api, fit, scenes

    s = scenes.shoebox(12.0, 9.0, 5.0, spec=0.3, alpha_range=(0.1, 0.3), src=(3.0, 4.0, 2.0), rec=(8.0, 5.0, 1.7))
    n_mat = len(s["materials"])
    alpha_t, spec_t = s["alpha"], np.full(n_mat, 0.3)
    sign = np.random.default_rng(5).choice([-1, 1], size=(n_mat, 4))
    alpha0, spec0 = alpha_t * (1 + 0.2 * sign[:, :3]), spec_t * (1 + 0.2 * sign[:, 3])
    kw = dict(recorders=s["recorders"], src_pos=s["src"], air=s["air"], n_rays=1_000_000, max_bounces=999,
              cutoff=1e-8, n_bins=4096 * 11, abs_hist=True, want_raw=False)
    out = {}
    with api.Scene(ctx, s["tris"], s["mat"], fit.materials(alpha_t, spec_t)) as sc:
        target = fit.log_edc(fit.coarse_bins(api.render(ctx, sc, **kw).abs_value))
        mask = fit.decay_mask(target)
        sc.set_materials(fit.materials(alpha0, spec0))
        api.render(ctx, sc, **kw)  # warm-up
        t0 = time.perf_counter()
        r = api.render(ctx, sc, **kw)
        t_fwd = time.perf_counter() - t0
        loss, adj = fit.log_edc_loss(r.abs_value, target, mask)
        api.render_backward(ctx, sc, adj_abs=adj, score=True, **kw)  # warm-up
        t0 = time.perf_counter()
        api.render_backward(ctx, sc, adj_abs=adj, score=True, **kw)
        t_bwd = time.perf_counter() - t0
        api.render_backward(ctx, sc, adj_abs=adj, local=True, **kw)  # warm-up
        t0 = time.perf_counter()
        api.render_backward(ctx, sc, adj_abs=adj, local=True, **kw)
        t_bwd_local = time.perf_counter() - t0
        ref = api.render(ctx, sc, **dict(kw, n_rays=8_000_000, seed=0x5EED))
        edc_err = max(acoustics.edc_error(r.abs_value[b, 0], ref.abs_value[b, 0]) for b in range(3))
        segs = r.stats["segments"]
        # the same forward + adjoint at 8M rays, where the EDC error is within 1%
        kw4 = dict(kw, n_rays=8_000_000)
        t0 = time.perf_counter()
        r4 = api.render(ctx, sc, **kw4)
        t4f = time.perf_counter() - t0
        _, adj4 = fit.log_edc_loss(r4.abs_value, target, mask)
        t0 = time.perf_counter()
        api.render_backward(ctx, sc, adj_abs=adj4, score=True, **kw4)
        t4b = time.perf_counter() - t0
        ref16 = api.render(ctx, sc, **dict(kw, n_rays=16_000_000, seed=0x5EED))
        edc4 = max(acoustics.edc_error(r4.abs_value[b, 0], ref16.abs_value[b, 0]) for b in range(3))
        segs4 = r4.stats["segments"]
    log(f"adjoint: forward {t_fwd * 1e3:.1f} ms, backward {t_bwd * 1e3:.1f} ms (score form; local form "
        f"{t_bwd_local * 1e3:.1f} ms), {segs} segments, "
        f"EDC error {edc_err:.4f}")
    out.update({"workload": "shoebox 12x9x5 m, 18 absorption + 6 specularity parameters, 1M rays, 4096-bin log-EDC "
                            "loss, parameters +-20% off the target's",
                "adjoint_form": "score (EAR_FLAG_ADJ_SCORE)",
                "segments": int(segs), "forward_ms": round(1e3 * t_fwd, 3), "backward_ms": round(1e3 * t_bwd, 3),
                "backward_local_form_ms": round(1e3 * t_bwd_local, 3),
                "forward_segments_per_s": round(segs / t_fwd, 1),
                "forward_adjoint_segments_per_s": round(segs / (t_fwd + t_bwd), 1),
                "forward_adjoint_vs_forward": round(t_fwd / (t_fwd + t_bwd), 4),
                "edc_error_vs_8M_rays": round(edc_err, 5), "loss": round(loss, 6),
                "at_8M_rays": {"segments": int(segs4), "forward_ms": round(1e3 * t4f, 3),
                               "backward_ms": round(1e3 * t4b, 3),
                               "forward_segments_per_s": round(segs4 / t4f, 1),
                               "forward_adjoint_segments_per_s": round(segs4 / (t4f + t4b), 1),
                               "edc_error_vs_16M_rays": round(edc4, 5)}})
    log(f"adjoint at 8M rays: forward {t4f * 1e3:.1f} ms, backward {t4b * 1e3:.1f} ms, EDC error {edc4:.4f}")

    rng = np.random.default_rng(17)
    rooms = [scenes.room(seed=5000 + i, target_tris=int(np.exp(rng.uniform(np.log(1000), np.log(100_000)))))
             for i in range(n_rooms)]
    scs = [api.Scene(ctx, rm["tris"], rm["mat"], rm["materials"]) for rm in rooms]
    kws = [dict(recorders=rm["recorders"], src_pos=rm["src"], air=rm["air"], n_rays=250_000, max_bounces=999,
                cutoff=1e-8, n_bins=4096 * 11, abs_hist=True, want_raw=False) for rm in rooms]
    fwd_items = [(sc, k) for sc, k in zip(scs, kws)]
    res, _, _ = api.render_batch(ctx, fwd_items)  # warm-up; its histograms serve as targets
    targets = [fit.log_edc(fit.coarse_bins(x.abs_value)) for x in res]
    t0 = time.perf_counter()
    res, st, _ = api.render_batch(ctx, fwd_items)
    adjs = [fit.log_edc_loss(x.abs_value * 1.1, t, fit.decay_mask(t))[1] for x, t in zip(res, targets)]
    t1 = time.perf_counter()
    api.render_backward_batch(ctx, [(sc, dict(k, adj_abs=a, score=True)) for sc, k, a in zip(scs, kws, adjs)])
    t2 = time.perf_counter()
    # accuracy of the 250k-ray rooms: EDC error against independent 2M-ray renders
    refs, _, _ = api.render_batch(ctx, [(sc, dict(k, n_rays=2_000_000, seed=0x5EED)) for sc, k in fwd_items])
    errs = [max(acoustics.edc_error(x.abs_value[b, 0], y.abs_value[b, 0]) for b in range(3))
            for x, y in zip(res, refs)]
    # the accuracy point: the same batch at R rays per room, forward + adjoint
    # timed, and each room's EDC error (every band, to -60 dB) against an
    # independent render of 4R rays, for the 8 rooms whose 250k-ray error
    # was largest (the 4R render's own error is about half the R render's)
    R = rays_accurate
    kwsR = [dict(k, n_rays=R) for k in kws]
    t0b = time.perf_counter()
    resR, stR, _ = api.render_batch(ctx, [(sc, k) for sc, k in zip(scs, kwsR)])
    adjsR = [fit.log_edc_loss(x.abs_value * 1.1, t, fit.decay_mask(t))[1] for x, t in zip(resR, targets)]
    t1b = time.perf_counter()
    api.render_backward_batch(ctx, [(sc, dict(k, adj_abs=a, score=True)) for sc, k, a in zip(scs, kwsR, adjsR)])
    t2b = time.perf_counter()
    worst = [int(i) for i in np.argsort(errs)[::-1][:8]]
    refs4, _, _ = api.render_batch(ctx, [(scs[i], dict(kws[i], n_rays=4 * R, seed=0x5EED)) for i in worst])
    errsR = [max(acoustics.edc_error(resR[i].abs_value[b, 0], y.abs_value[b, 0]) for b in range(3))
             for i, y in zip(worst, refs4)]
    for sc in scs:
        sc.close()
    bsegs = st["segments"]
    log(f"adjoint batch: {n_rooms} rooms, {bsegs} segments, forward {t1 - t0:.3f}s, backward {t2 - t1:.3f}s")
    log(f"adjoint batch at {R} rays per room: forward {t1b - t0b:.3f}s, backward {t2b - t1b:.3f}s, EDC error "
        f"max {max(errsR):.4f} over the 8 rooms checked")
    out["batch"] = {"workload": f"{n_rooms} box/L rooms, 1k..100k triangles, 250k rays each, 3 bands, 4096-bin "
                                f"log-EDC loss", "segments": int(bsegs), "forward_s": round(t1 - t0, 4),
                    "backward_s": round(t2 - t1, 4),
                    "forward_adjoint_segments_per_s": round(bsegs / (t2 - t0), 1),
                    "forward_segments_per_s": round(bsegs / (t1 - t0), 1),
                    "edc_error_vs_2M_rays_median": round(float(np.median(errs)), 5),
                    "edc_error_vs_2M_rays_max": round(float(np.max(errs)), 5),
                    f"at_{R // 1_000_000}M_rays": {
                        "segments": int(stR["segments"]), "forward_s": round(t1b - t0b, 4),
                        "backward_s": round(t2b - t1b, 4),
                        "forward_segments_per_s": round(stR["segments"] / (t1b - t0b), 1),
                        "forward_adjoint_segments_per_s": round(stR["segments"] / (t2b - t0b), 1),
                        "rooms_checked": worst,
                        f"edc_error_vs_{4 * R // 1_000_000}M_rays_max": round(float(np.max(errsR)), 5),
                        f"edc_error_vs_{4 * R // 1_000_000}M_rays_median": round(float(np.median(errsR)), 5)}}
    return out


def bench_animated(ctx, n_keys=64, rays_total=50_000_000, n_tris=500_000):
    """Animated source: the ~500k-triangle hall with the source moving over the
    stage along n_keys keyframes (one render each, src/Animated.h:89-91), a
    stereo recorder, rays_total rays per band over all keyframes, 3 bands,
    100 segments, 3 s histograms; all keyframes in one ear_render_batch over
    the shared scene, beside one ear_render per keyframe."""
    from ear_amd import api, scenes

    t0 = time.time()
    h = scenes.hall(target_tris=n_tris)
    path = scenes.stage_path(n_keys)
    rays = rays_total // n_keys
    log(f"animated: {h['name']} generated in {time.time() - t0:.1f}s")
    sc = api.Scene(ctx, h["tris"], h["mat"], h["materials"])
    info = sc.info()
    items = [(sc, dict(recorders=h["recorders"], src_pos=path[k], air=h["air"], n_rays=rays, max_bounces=100,
                       cutoff=1e-8, n_bins=132_300, keyframe=k, abs_hist=False, want_raw=False))
             for k in range(n_keys)]
    api.render_batch(ctx, items[:2])  # warm-up (first-touch allocations)
    _, st, wall = api.render_batch(ctx, items)
    log(f"animated: {st['segments']} segments in {wall:.3f}s ({st['segments'] / wall:.3e} segments/s)")
    t0 = time.perf_counter()
    segs1 = 0
    for _, kw in items:
        segs1 += api.render(ctx, sc, **kw).stats["segments"]
    single = time.perf_counter() - t0
    log(f"animated: one call per keyframe {single:.3f}s ({segs1 / single:.3e} segments/s)")
    sc.close()
    return {"workload": f"hall {info['n_tris']} tris, source on a {n_keys}-keyframe path, stereo, 3 bands x "
                        f"{rays} rays per keyframe ({n_keys * rays} per band), <= 100 segments, 3 s histograms",
            "segments": int(st["segments"]), "bvh_build_ms": round(info["build_ms"], 1), "batch_s": round(wall, 4),
            "batch_segments_per_s": round(st["segments"] / wall, 1), "single_calls_s": round(single, 4),
            "single_calls_segments_per_s": round(segs1 / single, 1), "speedup": round(single / wall, 3)}


def bench_fit(ctx, steps=200, n_rooms=32, room_steps=10):
    """Inverse fitting by the adjoint.  The 12 x 9 x 5 m shoebox: 18 absorption
    + 6 specularity parameters started +-20% off, 1M rays, `steps` Adam steps
    on the 4096-bin log-EDC loss against a target rendered at the true
    parameters with another seed.  Then n_rooms rooms of 1k..100k triangles
    (box rooms 6 materials, L rooms 3; absorption per band + specularity)
    fitted together by batched forward and adjoint renders, 250k rays each."""
    import numpy as np

    from ear_amd import api, fit, scenes

    s = scenes.shoebox(12.0, 9.0, 5.0, spec=0.3, alpha_range=(0.1, 0.3), src=(3.0, 4.0, 2.0), rec=(8.0, 5.0, 1.7))
    n_mat = len(s["materials"])
    alpha_t, spec_t = s["alpha"], np.full(n_mat, 0.3)
    sign = np.random.default_rng(5).choice([-1, 1], size=(n_mat, 4))
    alpha0, spec0 = alpha_t * (1 + 0.2 * sign[:, :3]), spec_t * (1 + 0.2 * sign[:, 3])
    kw = dict(recorders=s["recorders"], src_pos=s["src"], air=s["air"], n_rays=1_000_000, max_bounces=999,
              cutoff=1e-8, n_bins=4096 * 11, abs_hist=True, want_raw=False)
    with api.Scene(ctx, s["tris"], s["mat"], fit.materials(alpha_t, spec_t)) as sc:
        target = api.render(ctx, sc, **dict(kw, seed=0x5EED)).abs_value
    t0 = time.perf_counter()
    st = fit.fit_room(ctx, s, target, alpha0, spec0, steps=steps, n_rays=1_000_000, max_bounces=999, cutoff=1e-8)
    wall = time.perf_counter() - t0

    # one receiver's decay fixes the area-weighted absorption of each band, not
    # the six walls apart: report that quantity's relative error (max over bands)
    tr = np.asarray(s["tris"], np.float64)
    area_t = 0.5 * np.linalg.norm(np.cross(tr[:, 1] - tr[:, 0], tr[:, 2] - tr[:, 0]), axis=1)
    area = np.array([area_t[np.asarray(s["mat"]) == m].sum() for m in range(n_mat)])

    def err(a):
        return float(np.max(np.abs(area @ a / (area @ alpha_t) - 1.0)))

    out = {"workload": f"shoebox 12x9x5 m, 24 parameters +-20% off, 1M rays, {steps} Adam steps on the 4096-bin "
                       f"log-EDC loss (target: the true parameters, another seed)",
           "steps": steps, "wall_s": round(wall, 3), "ms_per_step": round(1e3 * wall / steps, 2),
           "loss_first": round(st.history[0][0], 6), "loss_last": round(st.history[-1][0], 6),
           "area_weighted_alpha_rel_error_first": round(err(st.history[0][1]), 5),
           "area_weighted_alpha_rel_error_last": round(err(st.history[-1][1]), 5),
           "spec_mean_abs_error_first": round(float(np.abs(st.history[0][2] - spec_t).mean()), 5),
           "spec_mean_abs_error_last": round(float(np.abs(st.history[-1][2] - spec_t).mean()), 5)}
    log(f"fit: {steps} steps in {wall:.2f}s, loss {out['loss_first']} -> {out['loss_last']}, area-weighted alpha "
        f"error {out['area_weighted_alpha_rel_error_first']} -> {out['area_weighted_alpha_rel_error_last']}, "
        f"spec error {out['spec_mean_abs_error_first']} -> {out['spec_mean_abs_error_last']}")

    rng = np.random.default_rng(17)
    rooms = [scenes.room(seed=5000 + i, target_tris=int(np.exp(rng.uniform(np.log(1000), np.log(100_000)))))
             for i in range(n_rooms)]
    targets, a0s, s0s = [], [], []
    for i, rm in enumerate(rooms):
        nm = len(rm["materials"])
        a_t = np.array([[1.0 - x for x in m[0]] for m in rm["materials"]])
        sg = np.random.default_rng(100 + i).choice([-1, 1], size=(nm, 4))
        with api.Scene(ctx, rm["tris"], rm["mat"], fit.materials(a_t, np.full(nm, 0.2))) as sc:
            targets.append(api.render(ctx, sc, recorders=rm["recorders"], src_pos=rm["src"], air=rm["air"],
                                      n_rays=250_000, max_bounces=999, cutoff=1e-8, n_bins=4096 * 11,
                                      want_raw=False, seed=0x5EED).abs_value)
        a0s.append(np.clip(a_t * (1 + 0.2 * sg[:, :3]), 0.01, 0.95))
        s0s.append(np.full(nm, 0.2) * (1 + 0.2 * sg[:, 3]))
    t0 = time.perf_counter()
    sts = fit.fit_rooms(ctx, rooms, targets, a0s, s0s, steps=room_steps, n_rays=250_000)
    wall = time.perf_counter() - t0
    ratio = [x.history[-1][0] / max(x.history[0][0], 1e-30) for x in sts]
    n_par = int(sum(4 * len(rm["materials"]) for rm in rooms))
    out["rooms"] = {"workload": f"{n_rooms} box/L rooms, 1k..100k triangles, {n_par} parameters in all (+-20% off), "
                                f"250k rays each, {room_steps} steps of batched forward + adjoint",
                    "steps": room_steps, "wall_s": round(wall, 3), "s_per_step": round(wall / room_steps, 3),
                    "loss_ratio_median": round(float(np.median(ratio)), 4),
                    "loss_ratio_max": round(float(np.max(ratio)), 4)}
    log(f"fit rooms: {n_rooms} rooms x {room_steps} steps in {wall:.2f}s, last / first loss: median "
        f"{out['rooms']['loss_ratio_median']}, max {out['rooms']['loss_ratio_max']}")
    return out


def cpu_model() -> str | None:
    """The host CPU's model name (as lscpu reports it), from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(s, kw, target_s):
    """The CPU oracle (port of the reference's brute-force tracer) on the first
    K rays of each band, all host threads; K is calibrated to ~target_s."""
    from oracle import oracle as O

    nthreads = min(os.cpu_count() or 1, 64)
    osc = O.OracleScene(s["tris"], s["mat"], s["materials"])

    def run(k):
        t = 0.0
        segs = 0
        for b in (0, 1, 2):
            r = O.render(osc, band=b, air=s["air"][b], recorders=s["recorders"], n_rays=kw["n_rays"],
                         ray_end=k, src_pos=s["src"], max_bounces=kw["max_bounces"], cutoff=kw["cutoff"],
                         n_bins=kw["n_bins"], n_threads=nthreads, want_abs=False)
            t += r["seconds"]
            segs += int(r["segments"])
        return t, segs

    k = nthreads
    t, segs = run(k)
    if t < target_s / 4:
        k = max(k, int(k * target_s / max(t, 1e-3)))
        t, segs = run(k)
    log(f"cpu baseline: {3 * k} rays, {segs} segments in {t:.2f}s on {nthreads} threads")
    return {"value": round(segs / t, 2), "unit": "segments/s", "cores": nthreads, "kind": "port",
            "cpu_model": cpu_model(), "cpus_online": os.cpu_count(),
            "sample": f"first {k} rays of each of the 3 band contexts of the same workload ({segs} segments, "
                      f"{t:.1f} s), brute-force triangle scans as the reference"}


def cpu_reference_threading(s, kw, k):
    """The same CPU port with the reference's threading: one thread per band
    context (src/EAR.cpp:193-207), the three contexts at once, each on its
    first k rays."""
    import threading

    from oracle import oracle as O

    osc = O.OracleScene(s["tris"], s["mat"], s["materials"])
    segs = [0, 0, 0]

    def one(b):
        r = O.render(osc, band=b, air=s["air"][b], recorders=s["recorders"], n_rays=kw["n_rays"], ray_end=k,
                     src_pos=s["src"], max_bounces=kw["max_bounces"], cutoff=kw["cutoff"], n_bins=kw["n_bins"],
                     n_threads=1, want_abs=False)
        segs[b] = int(r["segments"])

    th = [threading.Thread(target=one, args=(b,)) for b in range(3)]
    t0 = time.perf_counter()
    for x in th:
        x.start()
    for x in th:
        x.join()
    t = time.perf_counter() - t0
    log(f"cpu baseline, reference threading: {3 * k} rays, {sum(segs)} segments in {t:.2f}s on 3 threads")
    return {"value": round(sum(segs) / t, 2), "unit": "segments/s", "cores": 3, "kind": "port",
            "sample": f"first {k} rays of each band context, one thread per context ({sum(segs)} segments, "
                      f"{t:.1f} s)"}


if __name__ == "__main__":
    main()
