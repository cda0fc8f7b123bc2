#!/usr/bin/env python3
"""Benchmark: ray-segments/s of the MI355X tracer on the 200k-triangle hall.

Workload (one step = one full render, i.e. Scene::Render of src/Scene.cpp:111-318
for the three band contexts of one sound, fed by synthetic seeded input):
  procedural concert hall of ~200k fp32 triangles, 8 materials x 3 bands with
  alpha ~ U(0.02, 0.6), one point source, stereo recorder (head 0.17 m),
  10M rays per band, <= 100 segments per ray, intensity cut-off 1e-8,
  3 s histograms at 44.1 kHz (132,300 bins per channel and band).
The scene and its LBVH are resident in HBM before the timed region; each step
emits, traces and splats every ray and finalizes the histograms.

Prints one JSON line (rank 0, N=1).  `roofline` is for the closest-hit kernel
(k_extend), priced at algorithmic bytes per query
    24 (ray origin + direction) + 8 (hit record) + B * V + 36 * T
with V / T the mean BVH node visits / triangle tests per query counted by an
instrumented pass over the same rays, and B the bytes of one node that a visit
reads (112 for the 4-wide BVH: four child boxes and four child references; 56
for the binary one: two child boxes and two child references).  `cpu_baseline` is the CPU oracle (a
brute-force restatement of the reference algorithm) timed on this host on the
first K rays of the same Philox streams.
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
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (PMC calibration) and exit")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic_k_extend_hall200k.json"))
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
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md, not built)"}))
        sys.exit(2)
    import numpy as np

    from ear_amd import api, scenes

    t0 = time.time()
    s = scenes.hall(target_tris=a.tris)
    n_bins = int(round(44100 * a.seconds))
    log(f"scene {s['name']} generated in {time.time() - t0:.1f}s")
    ctx = api.Context(0)
    sc = api.Scene(ctx, s["tris"], s["mat"], s["materials"])
    info = sc.info()
    log(f"LBVH: {info['n_nodes']} nodes, depth {info['max_depth']}, build {info['build_ms']:.2f} ms")
    kw = dict(recorders=s["recorders"], n_rays=a.rays, src_pos=s["src"], air=s["air"], max_bounces=a.max_bounces,
              cutoff=a.cutoff, n_bins=n_bins, abs_hist=True, want_raw=False, want_value=True)

    for i in range(a.warmup):
        r = api.render(ctx, sc, **kw)
        log(f"warmup {i}: {r.wall_s:.3f}s, {r.stats['segments']} segments")
    walls, segs, ext_ms, ext_q, ext_launch, stats = [], 0, 0.0, 0, 0, None
    t_start = time.perf_counter()
    for i in range(a.steps):
        r = api.render(ctx, sc, **kw)
        walls.append(r.wall_s)
        segs += r.stats["segments"]
        ext_ms += r.stats["extend_ms"]
        ext_q += r.stats["extend_queries"]
        ext_launch += r.stats["extend_launches"]
        stats = r.stats
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

    roof = None
    if not a.no_instrument:
        ri = api.render(ctx, sc, instrument=True, **kw)
        st = ri.stats
        V = st["node_visits_closest"] / max(1, st["extend_queries"])
        T = st["tri_tests_closest"] / max(1, st["extend_queries"])
        bpq = 24 + 8 + info["node_bytes"] * V + 36 * T
        achieved = bpq * ext_q / (ext_ms * 1e-3) / 1e9
        traffic = None
        if os.path.exists(a.traffic_file):
            with open(a.traffic_file) as f:
                traffic = json.load(f).get("bytes_per_launch")
        roof = {"bound": "hbm", "kernel": "k_extend", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_query": round(bpq, 1), "node_bytes": info["node_bytes"],
                "node_visits_per_query": round(V, 3),
                "tri_tests_per_query": round(T, 3),
                "shadow_node_visits_per_query": round(st["node_visits_shadow"] / max(1, st["shadow_queries"]), 3),
                "shadow_tri_tests_per_query": round(st["tri_tests_shadow"] / max(1, st["shadow_queries"]), 3),
                "avg_launch_ms": round(ext_ms / max(1, ext_launch), 4), "launches": ext_launch,
                "queries_per_launch": round(ext_q / max(1, ext_launch), 1)}

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(s, kw, a.cpu_seconds)

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
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


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
            "sample": f"first {k} rays of each of the 3 band contexts of the same workload ({segs} segments, "
                      f"{t:.1f} s), brute-force triangle scans as the reference"}


if __name__ == "__main__":
    main()
