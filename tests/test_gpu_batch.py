"""Batched renders (many scenes in one wavefront) against single renders, and
full-size configuration checks, through the C ABI."""
import numpy as np
import pytest

from ear_amd import api, scenes

pytestmark = pytest.mark.gpu


def _kw(s, **extra):
    kw = dict(recorders=s["recorders"], src_pos=s["src"], air=s["air"], max_bounces=60, cutoff=1e-7, n_bins=44100)
    kw.update(extra)
    return kw


def test_batch_equals_single_renders(gpu_ctx):
    rooms = [scenes.room(seed=i, target_tris=[12, 400, 3000, 9000][i % 4]) for i in range(7)]
    rooms.append(scenes.shoebox(stereo=True))
    src_tris = np.array([[[1.0, 1.0, 1.0], [2.0, 1.0, 1.0], [1.0, 2.0, 1.2]]], np.float32)
    objs = [api.Scene(gpu_ctx, r["tris"], r["mat"], r["materials"]) for r in rooms]
    try:
        items = []
        for i, (r, sc) in enumerate(zip(rooms, objs)):
            extra = dict(n_rays=[500, 3000, 1200, 10000, 77, 2500, 4000, 6000][i], instance=i, seed=0xEA2 + i,
                         bands=[(0, 1, 2), (1,), (0, 2), (0, 1, 2)][i % 4])
            if i == 5:
                extra["src_tris"] = src_tris
            items.append((sc, _kw(r, **extra)))
        res, stats, _ = api.render_batch(gpu_ctx, items)
        seg = 0
        for (sc, kw), rb in zip(items, res):
            rs = api.render(gpu_ctx, sc, **kw)
            seg += rs.stats["segments"]
            np.testing.assert_array_equal(rb.raw, rs.raw)
            np.testing.assert_array_equal(rb.abs_raw, rs.abs_raw)
            np.testing.assert_array_equal(rb.value, rs.value)
        assert stats["segments"] == seg
    finally:
        for o in objs:
            o.close()


def test_batch_with_small_chunks_is_identical(gpu_ctx):
    rooms = [scenes.room(seed=10 + i, target_tris=600) for i in range(4)]
    objs = [api.Scene(gpu_ctx, r["tris"], r["mat"], r["materials"]) for r in rooms]
    try:
        items = [(sc, _kw(r, n_rays=1500 + 100 * i, instance=i)) for i, (r, sc) in enumerate(zip(rooms, objs))]
        a, _, _ = api.render_batch(gpu_ctx, items)
        gpu_ctx.set_tuning(ray_chunk=701)
        b, _, _ = api.render_batch(gpu_ctx, items)
        gpu_ctx.set_tuning()
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x.raw, y.raw)
    finally:
        for o in objs:
            o.close()


def test_batch_under_graph_replay_with_the_tail(gpu_ctx):
    """A batch of rooms of 600 to 9000 triangles with the bounce loop replayed
    as a captured graph (4 iterations, small queues so that it replays many
    times), the last live rays handed to the tail launch after a replay,
    equals the eager batch, on two calls in a row."""
    rooms = [scenes.room(seed=20 + i, target_tris=[600, 3000, 9000][i % 3]) for i in range(5)]
    objs = [api.Scene(gpu_ctx, r["tris"], r["mat"], r["materials"]) for r in rooms]
    try:
        items = [(sc, _kw(r, n_rays=2000 + 700 * i, instance=i)) for i, (r, sc) in enumerate(zip(rooms, objs))]
        ref, st_ref, _ = api.render_batch(gpu_ctx, items)
        gpu_ctx.set_tuning(graph_iters=4, ray_chunk=3000, tail_rays=4000)
        try:
            for _ in range(2):
                res, st, _ = api.render_batch(gpu_ctx, items)
                assert st["segments"] == st_ref["segments"]
                for x, y in zip(ref, res):
                    np.testing.assert_array_equal(x.raw, y.raw)
                    np.testing.assert_array_equal(x.abs_raw, y.abs_raw)
        finally:
            gpu_ctx.set_tuning()
    finally:
        for o in objs:
            o.close()


def test_config1_full_size_against_fp64_oracle(gpu_ctx, oracle):
    """Shoebox 10x8x4, 12 triangles, 100k rays, <= 50 bounces, cut-off 1e-6,
    88,200 bins per band: histograms within 1e-5 of the FP64 oracle."""
    s = scenes.shoebox()
    kw = dict(max_bounces=50, cutoff=1e-6, n_bins=88200)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        g = api.render(gpu_ctx, sc, recorders=s["recorders"], n_rays=100_000, src_pos=s["src"], air=s["air"], **kw)
    osc = oracle.OracleScene(s["tris"], s["mat"], s["materials"])
    seg = 0
    for b in (0, 1, 2):
        o = oracle.render(osc, band=b, air=s["air"][b], recorders=s["recorders"], n_rays=100_000, src_pos=s["src"],
                          n_threads=8, **kw)
        seg += o["segments"]
        gs = api.raw_to_float(g.raw[b])
        tol = 1e-5 * o["hist_abs"] + 100_000 * 51 * api.FRAC_SCALE
        assert np.all(np.abs(gs - o["hist_signed"]) <= tol)
        assert np.all(gs[o["hist_abs"] == 0] == 0)
    assert g.stats["segments"] == seg


def test_keyframes_in_one_batch_share_a_scene(gpu_ctx):
    """A moving source: one item per keyframe, all on one scene.  Each item
    equals its own ear_render, and the keyframe index alone selects the random
    stream: two keyframes at the same position give different paths."""
    h = scenes.hall(target_tris=20_000)
    path = scenes.stage_path(6)
    path[5] = path[4]
    with api.Scene(gpu_ctx, h["tris"], h["mat"], h["materials"]) as sc:
        items = [(sc, _kw(h, n_rays=3000, keyframe=k, src_pos=path[k], n_bins=132_300)) for k in range(6)]
        res, stats, _ = api.render_batch(gpu_ctx, items)
        seg = 0
        for (_, kw), rb in zip(items, res):
            rs = api.render(gpu_ctx, sc, **kw)
            seg += rs.stats["segments"]
            np.testing.assert_array_equal(rb.raw, rs.raw)
            np.testing.assert_array_equal(rb.abs_raw, rs.abs_raw)
        assert stats["segments"] == seg
        assert not np.array_equal(res[4].raw, res[5].raw)
        # same position, same expected response: the totals agree statistically
        e4, e5 = res[4].abs_value.sum(), res[5].abs_value.sum()
        assert abs(e4 / e5 - 1.0) < 0.1
