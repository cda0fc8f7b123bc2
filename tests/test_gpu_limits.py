"""Limits and the finishing step of a render, through the C ABI: gain^2 on the
finished track, the bounce-number field, and ramps cut at the end of the
histogram (counted, with the length that keeps them whole)."""
import numpy as np
import pytest

from ear_amd import api, scenes

pytestmark = pytest.mark.gpu


def test_gain_squares_onto_the_finished_track(gpu_ctx, oracle):
    """Scene::Render scales each track by 1/N, adds the direct sound, then
    multiplies by gain * gain (src/Scene.cpp:286-316): value = (raw / N +
    direct) * gain^2, against the oracle's raw sums and direct ramps, for a
    stereo recorder and a gain of 0.5 and 1.7."""
    s = scenes.shoebox(stereo=True)
    n_rays, n_bins = 2000, 44100
    kw = dict(max_bounces=40, cutoff=1e-6, n_bins=n_bins)
    osc = oracle.OracleScene(s["tris"], s["mat"], s["materials"])
    ora = {b: oracle.render(osc, band=b, air=s["air"][b], recorders=s["recorders"], n_rays=n_rays,
                            src_pos=s["src"], n_threads=8, **kw) for b in (0, 1, 2)}
    finv = float(np.float32(1.0) / np.float32(n_rays))
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        for gain in (0.5, 1.7):
            g = api.render(gpu_ctx, sc, recorders=s["recorders"], n_rays=n_rays, src_pos=s["src"], air=s["air"],
                           gain=gain, **kw)
            g2 = float(np.float32(gain) * np.float32(gain))  # the product in float, as the reference forms it
            for b in (0, 1, 2):
                o = ora[b]
                exp = (o["hist_signed"] * finv + oracle.direct_ramp(o["direct"], n_bins)) * g2
                expa = (o["hist_abs"] * finv + np.abs(oracle.direct_ramp(o["direct"], n_bins))) * g2
                np.testing.assert_allclose(g.value[b], exp, rtol=1e-9, atol=1e-12 * np.abs(exp).max())
                np.testing.assert_allclose(g.abs_value[b], expa, rtol=1e-9, atol=1e-12 * np.abs(expa).max())
                assert np.abs(o["direct"]["a"]).max() > 0  # the direct sound is in the scaled sum


def test_max_bounces_is_bounded_by_its_field(gpu_ctx):
    """The bounce number travels in 16 bits of a queue entry: 65535 is the
    largest accepted max_bounces; 65536 is refused before any launch."""
    s = scenes.shoebox()
    kw = dict(recorders=s["recorders"], n_rays=200, src_pos=s["src"], air=s["air"], cutoff=1e-6, n_bins=44100)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        with pytest.raises(api.EarError) as e:
            api.render(gpu_ctx, sc, max_bounces=65536, **kw)
        assert e.value.status_name == "EAR_ERR_CAPACITY"
        with pytest.raises(api.EarError):
            api.render_backward(gpu_ctx, sc, adj_value=np.ones((3, 1, 44100)), max_bounces=1 << 20, **kw)
        r = api.render(gpu_ctx, sc, max_bounces=65535, **kw)
        assert r.stats["segments"] > 200


@pytest.mark.parametrize("tail", [1, 0])
def test_cut_ramps_are_counted_with_the_length_that_keeps_them(gpu_ctx, tail):
    """A histogram too short for the late contributions: their ramps are cut
    (the reference would grow its buffer, src/Recorder.cpp:52-59), counted in
    dropped_ramps, and bins_needed is the length at which none is cut.  The
    bins both lengths hold are bitwise equal (same paths, same splats), for
    the bounce loop (tail 1) and the tail kernel (default for 12 triangles)."""
    s = scenes.shoebox(stereo=True)
    kw = dict(recorders=s["recorders"], n_rays=3000, src_pos=s["src"], air=s["air"], max_bounces=50, cutoff=1e-6)
    gpu_ctx.set_tuning(tail_rays=tail)
    try:
        with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
            short = api.render(gpu_ctx, sc, n_bins=3000, **kw)
            assert short.stats["dropped_ramps"] > 0
            need = short.stats["bins_needed"]
            assert need > 3000
            full = api.render(gpu_ctx, sc, n_bins=need, **kw)
            assert full.stats["dropped_ramps"] == 0 and full.stats["bins_needed"] == 0
            almost = api.render(gpu_ctx, sc, n_bins=need - 1, **kw)
            assert almost.stats["dropped_ramps"] >= 1 and almost.stats["bins_needed"] == need
            np.testing.assert_array_equal(full.raw[:, :, :3000], short.raw)
            np.testing.assert_array_equal(full.abs_raw[:, :, :3000], short.abs_raw)
            for x in (short, full, almost):
                assert x.stats["segments"] == full.stats["segments"]
    finally:
        gpu_ctx.set_tuning()


def test_cut_direct_sound_is_counted(gpu_ctx):
    """Free field (no triangles): the direct sound is the only ramp
    (src/Scene.cpp:299-311).  It starts at bin floor(|SR| / 343 * 44100) = 692
    and is ceil(sqrt(|SR|)) = 3 bins wide: 694 bins cut it, 695 hold it."""
    kw = dict(recorders=[dict(pos=(7.0, 5.0, 1.7))], n_rays=100, src_pos=(2.0, 3.0, 1.5), air=(1.0, 1.0, 1.0),
              max_bounces=10, cutoff=1e-6, bands=(1,))
    with api.Scene(gpu_ctx, np.zeros((0, 3, 3), np.float32), np.zeros(0, np.uint16), [([0.5] * 3, None, None)]) as sc:
        cut = api.render(gpu_ctx, sc, n_bins=694, **kw)
        whole = api.render(gpu_ctx, sc, n_bins=695, **kw)
    assert cut.stats["dropped_ramps"] == 1 and cut.stats["bins_needed"] == 695
    assert whole.stats["dropped_ramps"] == 0
    assert np.count_nonzero(whole.value[1, 0]) == 3 and whole.value[1, 0, 692] > 0
    np.testing.assert_array_equal(cut.value[1, 0], whole.value[1, 0, :694])
