"""The batching and animated workloads at their full sizes, through the C ABI:
the 256-room batch (box and L rooms of 12..50k triangles, 1e4..2e6 rays each,
1 s histograms) and the 500k-triangle hall with a source on a 64-keyframe
path (stereo, 3 s histograms).  Batched renders equal per-item renders bit
for bit, and sampled items equal the FP64 brute-force oracle on the same
Philox streams (hit sequences exactly, histograms within the parity gate)."""
import numpy as np
import pytest

from ear_amd import api, scenes

pytestmark = pytest.mark.gpu

RTOL, QUANTUM = 1e-5, 2.0 ** -80


def _gate(g, o, n_splats):
    """|H_gpu - H_fp64| <= 1e-5 sum|c| + n 2^-80, and zero where the oracle
    splatted nothing."""
    gs, ga = api.raw_to_float(g["raw"]), api.raw_to_float(g["abs_raw"])
    tol = RTOL * o["hist_abs"] + n_splats * QUANTUM
    assert not (np.abs(gs - o["hist_signed"]) > tol).any()
    assert not (np.abs(ga - o["hist_abs"]) > tol).any()
    assert np.all(gs[o["hist_abs"] == 0] == 0)


@pytest.fixture(scope="module")
def rooms(gpu_ctx):
    specs = scenes.batch_rooms(256)
    print("256 rooms,", sum(len(r["tris"]) for r, _ in specs), "triangles", flush=True)
    objs = [api.Scene(gpu_ctx, r["tris"], r["mat"], r["materials"]) for r, _ in specs]
    yield specs, objs
    for o in objs:
        o.close()


def _room_kw(r, n_rays, i, **extra):
    kw = dict(recorders=r["recorders"], src_pos=r["src"], air=r["air"], n_rays=n_rays, max_bounces=999,
              cutoff=1e-8, n_bins=44100, instance=i)
    kw.update(extra)
    return kw


def test_room_batch_full_size_equals_single_renders(gpu_ctx, rooms):
    """All 256 rooms at their full ray counts in one ear_render_batch: the
    segment, shadow-ray and contribution totals equal the sums over one
    ear_render per room, and 16 sampled rooms (the smallest, the largest and
    14 spread between) have bitwise equal histograms."""
    specs, objs = rooms
    order = np.argsort([len(r["tris"]) for r, _ in specs])
    sample = set(int(order[k]) for k in np.linspace(0, len(order) - 1, 16).round().astype(int))
    items = [(sc, _room_kw(r, n, i, abs_hist=i in sample, want_raw=i in sample, want_value=False))
             for i, (sc, (r, n)) in enumerate(zip(objs, specs))]
    res, st, _ = api.render_batch(gpu_ctx, items)
    print("batch:", st["segments"], "segments", flush=True)
    tot = dict(segments=0, shadow_rays=0, contributions=0)
    for i, (sc, kw) in enumerate(items):
        r = api.render(gpu_ctx, sc, **kw)
        for k in tot:
            tot[k] += r.stats[k]
        if i in sample:
            np.testing.assert_array_equal(res[i].raw, r.raw)
            np.testing.assert_array_equal(res[i].abs_raw, r.abs_raw)
    for k, v in tot.items():
        assert st[k] == v, (k, st[k], v)
    assert st["segments"] > 1e10


@pytest.mark.parametrize("which", ["smallest", "median", "largest"])
def test_room_batch_items_against_the_oracle(gpu_ctx, oracle, rooms, which):
    """Three rooms of the batch (fewest, median and most triangles) rendered
    inside a batch of 12 with 1000 rays each, against the oracle fed the same
    instance word: every logged hit sequence, the counts and the histograms."""
    specs, objs = rooms
    order = np.argsort([len(r["tris"]) for r, _ in specs])
    i = int(order[{"smallest": 0, "median": len(order) // 2, "largest": -1}[which]])
    r, _ = specs[i]
    n_rays, tr, ts = 1000, 128, 64
    others = [int(j) for j in order[::23][:11] if int(j) != i][:11]
    items = [(objs[j], _room_kw(specs[j][0], 1000, j)) for j in others]
    items.insert(5, (objs[i], _room_kw(r, n_rays, i, trace_rays=tr, trace_segs=ts)))
    res, _, _ = api.render_batch(gpu_ctx, items)
    g = res[5]
    print(which, "room", i, len(r["tris"]), "triangles", flush=True)
    osc = oracle.OracleScene(r["tris"], r["mat"], r["materials"])
    seg = con = 0
    single = api.render(gpu_ctx, objs[i], **_room_kw(r, n_rays, i))
    for b in (0, 1, 2):
        o = oracle.render(osc, band=b, air=r["air"][b], recorders=r["recorders"], n_rays=n_rays, src_pos=r["src"],
                          max_bounces=999, cutoff=1e-8, n_bins=44100, instance=i, n_threads=16, trace_rays=tr,
                          trace_segs=ts)
        seg += int(o["segments"])
        con += int(o["contributions"])
        np.testing.assert_array_equal(g.trace_hits[b], o["trace_hits"])
        _gate(dict(raw=g.raw[b], abs_raw=g.abs_raw[b]), o, n_rays * 1000)
    assert single.stats["segments"] == seg and single.stats["contributions"] == con
    np.testing.assert_array_equal(single.raw, g.raw)


@pytest.fixture(scope="module")
def animated(gpu_ctx):
    h = scenes.hall(target_tris=500_000)
    path = scenes.stage_path(64)
    print("animated hall,", len(h["tris"]), "triangles", flush=True)
    with api.Scene(gpu_ctx, h["tris"], h["mat"], h["materials"]) as sc:
        yield h, path, sc


def _key_kw(h, path, k, n_rays, **extra):
    kw = dict(recorders=h["recorders"], src_pos=path[k], air=h["air"], n_rays=n_rays, max_bounces=100, cutoff=1e-8,
              n_bins=132_300, keyframe=k)
    kw.update(extra)
    return kw


def test_animated_hall_keyframes_against_the_oracle(gpu_ctx, oracle, animated):
    """Keyframes 0, 31 and 63 of the 500k-triangle hall's 64-keyframe path,
    100 rays each, inside one batch: hit sequences, counts and histograms
    against the oracle fed the keyframe's context word (band | (k + 1) << 16,
    src/Animated.h:89-91 selects the keyframe's source position)."""
    h, path, sc = animated
    keys, n_rays, tr, ts = (0, 31, 63), 100, 100, 100
    items = [(sc, _key_kw(h, path, k, n_rays, trace_rays=tr, trace_segs=ts)) for k in keys]
    res, st, _ = api.render_batch(gpu_ctx, items)
    osc = oracle.OracleScene(h["tris"], h["mat"], h["materials"])
    seg = 0
    for k, g in zip(keys, res):
        for b in (0, 1, 2):
            o = oracle.render(osc, band=b, air=h["air"][b], recorders=h["recorders"], n_rays=n_rays,
                              src_pos=path[k], max_bounces=100, cutoff=1e-8, n_bins=132_300, ctx=b | (k + 1) << 16,
                              n_threads=16, trace_rays=tr, trace_segs=ts)
            seg += int(o["segments"])
            np.testing.assert_array_equal(g.trace_hits[b], o["trace_hits"])
            _gate(dict(raw=g.raw[b], abs_raw=g.abs_raw[b]), o, n_rays * 101)
    assert st["segments"] == seg


def test_animated_hall_64_keyframe_batch_equals_single_renders(gpu_ctx, animated):
    """The 64 keyframes in one ear_render_batch on the shared scene (20,000
    rays each) equal one ear_render per keyframe bit for bit."""
    h, path, sc = animated
    items = [(sc, _key_kw(h, path, k, 20_000)) for k in range(64)]
    res, st, _ = api.render_batch(gpu_ctx, items)
    seg = 0
    for (_, kw), rb in zip(items, res):
        r = api.render(gpu_ctx, sc, **kw)
        seg += r.stats["segments"]
        np.testing.assert_array_equal(rb.raw, r.raw)
        np.testing.assert_array_equal(rb.abs_raw, r.abs_raw)
    assert st["segments"] == seg
