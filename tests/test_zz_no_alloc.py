"""ear_render_noalloc, the render that allocates no device memory, against
ear_render through the C ABI: the same histograms, paths and counts on two
calls in a row (the second smaller, on the scratch region the first left
behind), no device allocation at all (a context whose memory cap refuses
every allocation still renders), and a clear capacity error for a render the
fixed region cannot hold."""
import numpy as np
import pytest

from ear_amd import api, scenes

pytestmark = pytest.mark.gpu

SCRATCH_MIB = 256  # EAR_NOALLOC_SCRATCH_BYTES


def _same(a, b):
    for k in ("raw", "abs_raw", "value", "abs_value", "trace_hits", "trace_nseg"):
        x, y = getattr(a, k), getattr(b, k)
        assert (x is None) == (y is None), k
        if x is not None:
            np.testing.assert_array_equal(x, y, err_msg=k)
    for k in ("segments", "shadow_rays", "contributions", "rays", "dropped_ramps", "bins_needed"):
        assert a.stats[k] == b.stats[k], (k, a.stats[k], b.stats[k])


def test_noalloc_equals_render_on_two_calls_the_second_smaller(gpu_ctx):
    """A 20k-triangle hall, stereo, 3 bands x 600k rays, 3 s |c| and signed
    histograms (its queues are cut to what the region holds, so rays are
    regenerated many times), then a mono shoebox render of 2000 rays in one
    band with a shorter histogram, a trace log and a mesh source: each equals
    ear_render bit for bit, and the hall again afterwards."""
    hall = scenes.hall(target_tris=20_000)
    box = scenes.shoebox()
    emitter = np.array([[[4.0, 3.0, 1.0], [5.0, 3.0, 1.0], [4.0, 4.0, 1.2]]], np.float32)
    big = dict(recorders=hall["recorders"], n_rays=600_000, src_pos=hall["src"], air=hall["air"], max_bounces=100,
               cutoff=1e-8, n_bins=132_300)
    small = dict(recorders=box["recorders"], n_rays=2000, bands=(1,), src_pos=box["src"], src_tris=emitter,
                 air=box["air"], max_bounces=50, cutoff=1e-6, n_bins=30_000, trace_rays=64, trace_segs=16)
    with api.Scene(gpu_ctx, hall["tris"], hall["mat"], hall["materials"]) as sh, \
            api.Scene(gpu_ctx, box["tris"], box["mat"], box["materials"]) as sb:
        a1 = api.render_noalloc(gpu_ctx, sh, **big)
        b1 = api.render(gpu_ctx, sh, **big)
        _same(a1, b1)
        assert a1.stats["segments"] > 10 * 3 * 600_000
        a2 = api.render_noalloc(gpu_ctx, sb, **small)
        b2 = api.render(gpu_ctx, sb, **small)
        _same(a2, b2)
        assert a2.stats["segments"] > 2000
        assert np.all(a2.raw[(0, 2), ...] == 0)  # bands outside the mask stay untouched
        # and the larger render once more, on the region the smaller one used last
        _same(api.render_noalloc(gpu_ctx, sh, **big), b1)


def test_noalloc_makes_no_device_allocation(gpu_ctx):
    """With a memory cap that refuses every further device allocation of the
    context, ear_render fails with EAR_ERR_OUT_OF_MEMORY, while
    ear_render_noalloc renders the same histograms it gave without the cap."""
    box = scenes.shoebox(stereo=True)
    kw = dict(recorders=box["recorders"], n_rays=20_000, src_pos=box["src"], air=box["air"], max_bounces=200,
              cutoff=1e-8, n_bins=88_200)
    with api.Scene(gpu_ctx, box["tris"], box["mat"], box["materials"]) as sc:
        ref = api.render(gpu_ctx, sc, **kw)
        gpu_ctx.set_memory_cap(1)
        try:
            with pytest.raises(api.EarError) as e:
                api.render(gpu_ctx, sc, **kw)
            assert e.value.status_name == "EAR_ERR_OUT_OF_MEMORY"
            got = api.render_noalloc(gpu_ctx, sc, **kw)
        finally:
            gpu_ctx.set_memory_cap(0)
        _same(got, ref)


def test_noalloc_rejects_a_render_the_region_cannot_hold(gpu_ctx):
    """16 stereo recorders (32 channels) with 600k-bin signed and |c|
    histograms need 3 x 32 x 2 x 600k x 32 B = 3.7 GB of accumulators: the
    call fails with EAR_ERR_CAPACITY and names the region's size, and the
    next render on the region is unaffected."""
    box = scenes.shoebox()
    recs = [scenes.stereo_recorder((1.0 + 0.5 * i, 4.0, 1.7)) for i in range(16)]
    with api.Scene(gpu_ctx, box["tris"], box["mat"], box["materials"]) as sc:
        with pytest.raises(api.EarError) as e:
            api.render_noalloc(gpu_ctx, sc, recorders=recs, n_rays=100, src_pos=box["src"], air=box["air"],
                               max_bounces=10, cutoff=1e-6, n_bins=600_000)
        assert e.value.status_name == "EAR_ERR_CAPACITY"
        assert f"{SCRATCH_MIB} MiB" in str(e.value)
        kw = dict(recorders=box["recorders"], n_rays=3000, src_pos=box["src"], air=box["air"], max_bounces=50,
                  cutoff=1e-6, n_bins=44_100)
        _same(api.render_noalloc(gpu_ctx, sc, **kw), api.render(gpu_ctx, sc, **kw))
