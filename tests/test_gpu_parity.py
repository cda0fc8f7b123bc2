"""HIP tracer vs the CPU oracle, through the C ABI (libear_hip.so).

Parity bar (all on the same Philox stream):
  * per-ray hit-triangle sequence and segment count: bit-exact
  * segment / contribution counts: exact
  * histograms: |H_gpu - H_fp64| <= 1e-5 * sum|c| per bin (+ the 2^-88
    fixed-point quantum per splat), and bins with no contribution exactly 0
"""
import numpy as np
import pytest

from ear_amd import api, scenes

pytestmark = pytest.mark.gpu

RTOL = 1e-5
QUANTUM = 2.0 ** -88


def oracle_bands(O, s, bands, n_rays, **kw):
    sc = O.OracleScene(s["tris"], s["mat"], s["materials"])
    out = {}
    for b in bands:
        out[b] = O.render(sc, band=b, air=s["air"][b], recorders=s["recorders"], n_rays=n_rays,
                          src_pos=s["src"], ctx=b, n_threads=8, **kw)
    return out


def assert_hist_parity(gpu_signed, gpu_abs, ora_signed, ora_abs, n_splats_bound):
    tol = RTOL * ora_abs + n_splats_bound * QUANTUM
    err = np.abs(gpu_signed - ora_signed)
    bad = err > tol
    assert not bad.any(), f"{bad.sum()} bins out of tolerance; worst {np.max(err - tol)}"
    err_a = np.abs(gpu_abs - ora_abs)
    assert not (err_a > tol).any()
    assert np.all(gpu_signed[ora_abs == 0] == 0)
    assert np.all(gpu_abs[ora_abs == 0] == 0)


@pytest.fixture(scope="module")
def shoebox():
    return scenes.shoebox()


def test_shoebox_parity_bitexact_paths(gpu_ctx, oracle, shoebox):
    s = shoebox
    n_rays, tr, ts = 3000, 256, 64
    kw = dict(max_bounces=50, cutoff=1e-6, n_bins=88200)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        g = api.render(gpu_ctx, sc, recorders=s["recorders"], n_rays=n_rays, src_pos=s["src"], air=s["air"],
                       trace_rays=tr, trace_segs=ts, **kw)
    ora = oracle_bands(oracle, s, (0, 1, 2), n_rays, trace_rays=tr, trace_segs=ts, **kw)
    seg = sum(int(o["segments"]) for o in ora.values())
    con = sum(int(o["contributions"]) for o in ora.values())
    assert g.stats["segments"] == seg
    assert g.stats["contributions"] == con
    for b in (0, 1, 2):
        o = ora[b]
        np.testing.assert_array_equal(g.trace_hits[b], o["trace_hits"])
        np.testing.assert_array_equal(g.trace_nseg[b], o["trace_nseg"])
        gs = api.raw_to_float(g.raw[b])
        ga = api.raw_to_float(g.abs_raw[b])
        assert_hist_parity(gs, ga, o["hist_signed"], o["hist_abs"], n_rays * 51)
        # finalized value: raw * (1/N as float) + direct ramp, times gain^2
        finv = float(np.float32(1.0) / np.float32(n_rays))
        exp = o["hist_signed"] * finv + oracle.direct_ramp(o["direct"], kw["n_bins"])
        expa = o["hist_abs"] * finv + np.abs(oracle.direct_ramp(o["direct"], kw["n_bins"]))
        np.testing.assert_allclose(g.value[b], exp, rtol=1e-9, atol=1e-12 * np.abs(exp).max())
        np.testing.assert_allclose(g.abs_value[b], expa, rtol=1e-9, atol=1e-12 * np.abs(expa).max())


@pytest.mark.parametrize("width", [2, 4])
def test_hall_queries_match_bruteforce(gpu_ctx, oracle, width):
    s = scenes.hall(target_tris=20000)
    rng = np.random.default_rng(7)
    n = 4000
    lo = s["tris"].reshape(-1, 3).min(0)
    hi = s["tris"].reshape(-1, 3).max(0)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    # a quarter of the rays start exactly on a wall or axis-aligned (robust slab test)
    o[: n // 8, 2] = 0.0
    d[n // 8: n // 4] = np.eye(3, dtype=np.float32)[rng.integers(0, 3, n // 8)] * rng.choice([-1, 1], (n // 8, 1))
    osc = oracle.OracleScene(s["tris"], s["mat"], s["materials"])
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"], leaf_size=4, width=width) as sc:
        assert sc.info()["width"] == width
        t, idx = sc.query_closest(o, d)
        q = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
        occ = sc.query_occluded(o, q)
    for i in range(0, n, 7):
        ot, oi = osc.closest_hit(o[i], d[i])
        assert idx[i] == oi, (i, idx[i], oi)
        if oi >= 0:
            assert np.float32(t[i]) == np.float32(ot)
        assert occ[i] == osc.any_hit(o[i], q[i])


@pytest.mark.parametrize("leaf,width", [(1, 4), (2, 2), (8, 4), (16, 2)])
def test_bvh_shape_does_not_change_results(gpu_ctx, shoebox, leaf, width):
    s = scenes.shoebox(tess=(6, 6, 3))
    kw = dict(recorders=s["recorders"], n_rays=2000, src_pos=s["src"], air=s["air"], max_bounces=50, cutoff=1e-6,
              n_bins=44100)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"], leaf_size=4) as a:
        ra = api.render(gpu_ctx, a, **kw)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"], leaf_size=leaf, width=width) as b:
        rb = api.render(gpu_ctx, b, **kw)
    np.testing.assert_array_equal(ra.raw, rb.raw)
    np.testing.assert_array_equal(ra.abs_raw, rb.abs_raw)


def test_launch_shape_determinism(gpu_ctx, shoebox):
    s = shoebox
    kw = dict(recorders=s["recorders"], n_rays=5000, src_pos=s["src"], air=s["air"], max_bounces=50, cutoff=1e-6,
              n_bins=88200)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        gpu_ctx.set_tuning()
        ra = api.render(gpu_ctx, sc, **kw)
        gpu_ctx.set_tuning(ray_chunk=777, waves_per_cu=3, stack_lds=32)
        rb = api.render(gpu_ctx, sc, **kw)
        gpu_ctx.set_tuning(ray_chunk=4099, waves_per_cu=64, stack_lds=16)
        rc = api.render(gpu_ctx, sc, **kw)
        gpu_ctx.set_tuning()
    for r in (rb, rc):
        np.testing.assert_array_equal(ra.raw, r.raw)
        np.testing.assert_array_equal(ra.abs_raw, r.abs_raw)
        assert ra.stats["segments"] == r.stats["segments"]


def test_stereo_and_mesh_source_parity(gpu_ctx, oracle):
    s = scenes.shoebox(stereo=True)
    src_tris = np.array([[[4.0, 3.0, 1.0], [5.0, 3.0, 1.0], [4.0, 4.0, 1.2]],
                         [[5.0, 3.0, 1.0], [5.0, 4.0, 1.2], [4.0, 4.0, 1.2]]], np.float32)
    n_rays = 2000
    kw = dict(max_bounces=40, cutoff=1e-6, n_bins=44100)
    for mesh in (False, True):
        with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
            g = api.render(gpu_ctx, sc, recorders=s["recorders"], n_rays=n_rays, src_pos=s["src"], air=s["air"],
                           src_tris=src_tris if mesh else None, trace_rays=64, trace_segs=40, **kw)
        ora = oracle_bands(oracle, s, (0, 1, 2), n_rays, trace_rays=64, trace_segs=40,
                           src_tris=src_tris if mesh else None, **kw)
        for b in (0, 1, 2):
            o = ora[b]
            np.testing.assert_array_equal(g.trace_hits[b], o["trace_hits"])
            assert_hist_parity(api.raw_to_float(g.raw[b]), api.raw_to_float(g.abs_raw[b]), o["hist_signed"],
                               o["hist_abs"], n_rays * 41)
