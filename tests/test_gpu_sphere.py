"""The receiver-sphere estimator (EAR_EST_SPHERE) on the device against the
oracle, and against the reference's point estimator on a room's decay.

Parity bar as for the NEE estimator: the same Philox paths, contribution
counts exact, raw histograms within 1e-5 of the FP64 sums per bin, empty bins
exactly 0."""
import numpy as np
import pytest

from ear_amd import acoustics, api, scenes
from ear_amd import _ffi as F

pytestmark = pytest.mark.gpu

RTOL = 1e-5
QUANTUM = 2.0 ** -F.EAR_FRAC_BITS


def with_radius(recs, r):
    return [dict(rec, radius=r) for rec in recs]


def check_parity(g, ora, bands, n_rays, n_bins, dry_scale_note=""):
    con = sum(int(o["contributions"]) for o in ora.values())
    assert g.stats["contributions"] == con, dry_scale_note
    assert g.stats["shadow_rays"] == 0
    finv = float(np.float32(1.0) / np.float32(n_rays))
    for b in bands:
        o = ora[b]
        gs = api.raw_to_float(g.raw[b])
        ga = api.raw_to_float(g.abs_raw[b])
        tol = RTOL * o["hist_abs"] + n_rays * 60 * QUANTUM
        assert not (np.abs(gs - o["hist_signed"]) > tol).any()
        assert not (np.abs(ga - o["hist_abs"]) > tol).any()
        assert np.all(gs[o["hist_abs"] == 0] == 0)
        # no analytic direct sound: value = raw / N
        np.testing.assert_allclose(g.value[b], o["hist_signed"] * finv, rtol=1e-9,
                                   atol=1e-12 * max(np.abs(g.value[b]).max(), 1e-30))


@pytest.mark.parametrize("stereo,mesh,dry", [(False, False, 1.0), (False, False, 0.5), (True, False, 1.0),
                                             (True, True, 1.0)])
def test_sphere_parity(gpu_ctx, oracle, stereo, mesh, dry):
    s = scenes.shoebox(stereo=stereo, spec=0.3)
    recs = with_radius(s["recorders"], 0.6)
    src_tris = np.array([[[4.0, 3.0, 1.0], [5.0, 3.0, 1.0], [4.0, 4.0, 1.2]]], np.float32) if mesh else None
    n_rays, n_bins = 20000, 44100
    kw = dict(max_bounces=60, cutoff=1e-6, n_bins=n_bins)
    print(f"sphere parity stereo={stereo} mesh={mesh} dry={dry}", flush=True)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        g = api.render(gpu_ctx, sc, recorders=recs, n_rays=n_rays, src_pos=s["src"], air=s["air"], dry=dry,
                       src_tris=src_tris, estimator=F.EAR_EST_SPHERE, **kw)
    osc = oracle.OracleScene(s["tris"], s["mat"], s["materials"])
    ora = {b: oracle.render(osc, band=b, air=s["air"][b], recorders=recs, n_rays=n_rays, src_pos=s["src"], ctx=b,
                            dry=dry, src_tris=src_tris, n_threads=8, estimator=1, **kw) for b in (0, 1, 2)}
    check_parity(g, ora, (0, 1, 2), n_rays, n_bins)
    assert g.stats["contributions"] > 1000


def test_sphere_requires_radius(gpu_ctx):
    s = scenes.shoebox()
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        with pytest.raises(api.EarError):
            api.render(gpu_ctx, sc, recorders=s["recorders"], n_rays=100, src_pos=s["src"], n_bins=1024,
                       estimator=F.EAR_EST_SPHERE)
        with pytest.raises(api.EarError):
            api.render_backward(gpu_ctx, sc, recorders=with_radius(s["recorders"], 0.5), n_rays=100,
                                src_pos=s["src"], n_bins=1024, estimator=F.EAR_EST_SPHERE,
                                adj_value=np.zeros((3, 1, 1024)))


def test_sphere_and_point_estimators_agree_on_decay(gpu_ctx):
    # a diffuse room: both estimators see the same exponential decay of the
    # reverberant field; their levels differ (the point estimator's weight is
    # the reference's lobe, not a physical BRDF), so only decay rates compare
    s = scenes.shoebox(spec=0.0)
    n_rays = 400_000
    kw = dict(max_bounces=400, cutoff=1e-9, n_bins=88200, bands=(1,), src_pos=s["src"], air=s["air"])
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        a = api.render(gpu_ctx, sc, recorders=s["recorders"], n_rays=n_rays, **kw)
        b = api.render(gpu_ctx, sc, recorders=with_radius(s["recorders"], 0.5), n_rays=n_rays,
                       estimator=F.EAR_EST_SPHERE, **kw)
    t_nee = acoustics.t60_iso(a.abs_value[1, 0])
    t_sph = acoustics.t60_iso(b.abs_value[1, 0])
    print("T60 (T30 fit): point", t_nee, "sphere", t_sph, flush=True)
    assert abs(t_sph / t_nee - 1.0) < 0.08
