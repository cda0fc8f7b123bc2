"""The time-tiled splat (records binned by (row, time tile), accumulated per
tile in LDS) against the direct splat (128-bit global atomics per ramp point):
the histograms are the same integer sums, so every raw bin, the finished
values and the counters must be bitwise equal, through the C ABI."""
import numpy as np
import pytest

from ear_amd import api, scenes

pytestmark = pytest.mark.gpu

ATOMIC, TILED = 1, 2


def _both(ctx, fn):
    """fn() under the direct and the tiled splat; the tail launch is off so
    that every connection goes through the shadow kernel."""
    out = []
    for mode in (ATOMIC, TILED):
        ctx.set_tuning(splat=mode, tail_rays=1)
        try:
            out.append(fn())
        finally:
            ctx.set_tuning()
    return out


def _same(a, b):
    for k in ("raw", "abs_raw", "value", "abs_value"):
        x, y = getattr(a, k), getattr(b, k)
        if x is not None:
            np.testing.assert_array_equal(x, y, err_msg=k)
    for k in ("segments", "shadow_rays", "contributions", "dropped_ramps", "bins_needed", "shadow_contributions",
              "shadow_ramps"):
        assert a.stats[k] == b.stats[k], (k, a.stats[k], b.stats[k])


def test_tiled_splat_equals_direct_on_the_hall(gpu_ctx):
    """20k-triangle hall, stereo, signed and |c| rows, 3 s: 12 rows x 517
    tiles of 256 bins = 6204 keys, counted in LDS."""
    s = scenes.hall(target_tris=20_000)
    kw = dict(recorders=s["recorders"], n_rays=300_000, src_pos=s["src"], air=s["air"], max_bounces=100,
              cutoff=1e-8, n_bins=132_300)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        a, t = _both(gpu_ctx, lambda: api.render(gpu_ctx, sc, **kw))
    _same(a, t)
    assert t.stats["shadow_ramps"] == 2 * t.stats["shadow_contributions"] > 1_000_000
    assert t.stats["splat_ms"] > 0 and a.stats["splat_ms"] == 0


def test_tiled_splat_wide_ramps_and_cut_ones(gpu_ctx):
    """A live shoebox (absorption 0.01) with a 25 s histogram: late ramps are
    wider than the tile halo (ceil(sqrt(d)) > 62 past d = 3844 m) and go
    straight to HBM, and a 2 s copy cuts ramps at its end; both equal the
    direct splat, dropped counts included."""
    s = scenes.shoebox(stereo=True, alpha_range=(0.01, 0.01))
    kw = dict(recorders=s["recorders"], n_rays=20_000, src_pos=s["src"], air=(1.0, 0.9999, 0.999), max_bounces=2000,
              cutoff=1e-9)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        for n_bins in (25 * 44100, 2 * 44100):
            a, t = _both(gpu_ctx, lambda: api.render(gpu_ctx, sc, n_bins=n_bins, **kw))
            _same(a, t)
            if n_bins > 44100 * 20:
                late = np.nonzero(t.raw[0, 0, :, 0] | t.raw[0, 0, :, 1])[0]
                assert late.max() > 3844 / 343 * 44100  # ramps past the halo's width were added
            else:
                assert t.stats["dropped_ramps"] > 0


def test_tiled_splat_in_a_batch_with_many_keys(gpu_ctx):
    """24 rooms, mono and stereo, mesh and point sources, some bands only:
    more than 8192 keys, so the records are counted with global atomics."""
    rooms = [scenes.room(seed=40 + i, target_tris=[12, 300, 2000][i % 3]) for i in range(24)]
    emitter = np.array([[[1.0, 1.0, 1.0], [2.0, 1.0, 1.0], [1.0, 2.0, 1.2]]], np.float32)
    objs = [api.Scene(gpu_ctx, r["tris"], r["mat"], r["materials"]) for r in rooms]
    try:
        items = []
        for i, (r, sc) in enumerate(zip(rooms, objs)):
            recs = r["recorders"] if i % 2 else [scenes.stereo_recorder(r["recorders"][0]["pos"])]
            kw = dict(recorders=recs, n_rays=4000 + 500 * i, src_pos=r["src"], air=r["air"], max_bounces=60,
                      cutoff=1e-7, n_bins=88200, instance=i, bands=[(0, 1, 2), (1,), (0, 2)][i % 3])
            if i % 5 == 0:
                kw["src_tris"] = emitter
            items.append((sc, kw))
        (ra, sa, _), (rt, st, _) = _both(gpu_ctx, lambda: api.render_batch(gpu_ctx, items))
        for x, y in zip(ra, rt):
            np.testing.assert_array_equal(x.raw, y.raw)
            np.testing.assert_array_equal(x.abs_raw, y.abs_raw)
        for k in ("segments", "contributions", "shadow_ramps", "dropped_ramps"):
            assert sa[k] == st[k], k
    finally:
        for o in objs:
            o.close()


def test_tiled_splat_under_graph_replay(gpu_ctx):
    """The bounce loop captured as a HIP graph and replayed (no memset nodes:
    the key counts are cleared by the scan kernel) equals the eager direct
    splat, on several renders in a row."""
    s = scenes.shoebox(stereo=True)
    kw = dict(recorders=s["recorders"], n_rays=50_000, src_pos=s["src"], air=s["air"], max_bounces=200,
              cutoff=1e-8, n_bins=88200)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        gpu_ctx.set_tuning(splat=ATOMIC, tail_rays=1)
        try:
            ref = api.render(gpu_ctx, sc, **kw)
            gpu_ctx.set_tuning(splat=TILED, tail_rays=1, graph_iters=4, ray_chunk=20_000)
            for _ in range(3):
                _same(api.render(gpu_ctx, sc, **kw), ref)
        finally:
            gpu_ctx.set_tuning()
