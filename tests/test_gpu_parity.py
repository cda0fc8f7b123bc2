"""HIP tracer vs the CPU oracle, through the C ABI (libear_hip.so).

Parity bar (all on the same Philox stream):
  * per-ray hit-triangle sequence and segment count: bit-exact
  * segment / contribution counts: exact
  * histograms: |H_gpu - H_fp64| <= 1e-5 * sum|c| per bin (+ the 2^-80
    fixed-point quantum per splat), and bins with no contribution exactly 0
"""
import numpy as np
import pytest

from ear_amd import _ffi as F
from ear_amd import api, scenes

pytestmark = pytest.mark.gpu

RTOL = 1e-5
QUANTUM = 2.0 ** -F.EAR_FRAC_BITS


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


@pytest.mark.parametrize("spec", [0.5, None])
def test_shoebox_parity_bitexact_paths(gpu_ctx, oracle, spec):
    # spec None: materials without specularity (the lobe term is exactly zero)
    s = scenes.shoebox(spec=spec)
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


@pytest.mark.parametrize("width,builder,leaf", [(2, 1, 4), (4, 1, 4), (2, 2, 1), (4, 2, 3), (4, 3, 1), (4, 3, 2),
                                               (4, 3, 5), (4, 4, 1), (4, 4, 2), (4, 4, 6)])
def test_hall_queries_match_bruteforce(gpu_ctx, oracle, width, builder, leaf):
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
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"], leaf_size=leaf, width=width, builder=builder) as sc:
        assert sc.info()["width"] == width and sc.info()["builder"] == builder
        t, idx = sc.query_closest(o, d)
        q = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
        occ = sc.query_occluded(o, q)
    for i in range(0, n, 7):
        ot, oi = osc.closest_hit(o[i], d[i])
        assert idx[i] == oi, (i, idx[i], oi)
        if oi >= 0:
            assert np.float32(t[i]) == np.float32(ot)
        assert occ[i] == osc.any_hit(o[i], q[i])


@pytest.mark.parametrize("builder,leaf", [(1, 2), (2, 1), (3, 2), (4, 2)])
@pytest.mark.parametrize("offset,scale", [((3.0e4, -1.5e4, 8.0e3), 1.0), ((0.0, 0.0, 0.0), 0.05)])
def test_far_or_tiny_scene_queries_match_bruteforce(gpu_ctx, oracle, builder, leaf, offset, scale):
    """The hall 30 km from the origin (float spacing there: 2 mm, against
    triangles of a few cm) and shrunk 20x to 2 m (smaller quantisation
    scales; much smaller scenes lose their triangles to the reference's
    |det| < 1e-5 test on both sides): the 4-wide tree's quantised boxes and
    the box padding still keep every closest and any hit of the brute-force
    scan."""
    s = scenes.hall(target_tris=20000)
    tris = (s["tris"].astype(np.float64) * scale + np.asarray(offset)).astype(np.float32)
    rng = np.random.default_rng(8)
    n = 3000
    lo = tris.reshape(-1, 3).min(0)
    hi = tris.reshape(-1, 3).max(0)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    d[: n // 8] = np.eye(3, dtype=np.float32)[rng.integers(0, 3, n // 8)] * rng.choice([-1, 1], (n // 8, 1))
    osc = oracle.OracleScene(tris, s["mat"], s["materials"])
    with api.Scene(gpu_ctx, tris, s["mat"], s["materials"], leaf_size=leaf, width=4, builder=builder) as sc:
        t, idx = sc.query_closest(o, d)
        q = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
        occ = sc.query_occluded(o, q)
    hits = 0
    for i in range(0, n, 5):
        ot, oi = osc.closest_hit(o[i], d[i])
        assert idx[i] == oi, (i, idx[i], oi)
        if oi >= 0:
            hits += 1
            assert np.float32(t[i]) == np.float32(ot)
        assert occ[i] == osc.any_hit(o[i], q[i])
    assert hits > 100


@pytest.mark.parametrize("leaf,width,builder", [(1, 4, 1), (2, 2, 1), (8, 4, 1), (16, 2, 1), (1, 4, 2), (3, 2, 2),
                                               (16, 4, 2), (1, 4, 3), (2, 4, 3), (16, 4, 3), (1, 4, 4), (2, 4, 4),
                                               (16, 4, 4)])
def test_bvh_shape_does_not_change_results(gpu_ctx, shoebox, leaf, width, builder):
    s = scenes.shoebox(tess=(6, 6, 3))
    kw = dict(recorders=s["recorders"], n_rays=2000, src_pos=s["src"], air=s["air"], max_bounces=50, cutoff=1e-6,
              n_bins=44100)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"], leaf_size=4) as a:
        ra = api.render(gpu_ctx, a, **kw)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"], leaf_size=leaf, width=width, builder=builder) as b:
        rb = api.render(gpu_ctx, b, **kw)
    np.testing.assert_array_equal(ra.raw, rb.raw)
    np.testing.assert_array_equal(ra.abs_raw, rb.abs_raw)


@pytest.mark.parametrize("builder", [3, 4])
@pytest.mark.parametrize("case", ["one", "two", "three", "coincident", "within_one_leaf", "random_soup"])
def test_device_builds_small_and_degenerate_scenes_match_bruteforce(gpu_ctx, oracle, case, builder):
    """The device builds (PLOC, device SAH) on the inputs their clustering,
    splitting and collapse treat specially: 1, 2 and 3 triangles, many
    triangles with one centroid (every Morton code equal; every SAH bin
    empty but one, so ranges split by count), a scene of leaf_size triangles
    (the root's range is already a leaf), and a soup of 3000 random
    triangles: every closest and any hit equals the brute-force scan."""
    rng = np.random.default_rng({"one": 1, "two": 2, "three": 3, "coincident": 4, "within_one_leaf": 5,
                                 "random_soup": 6}[case])
    n = {"one": 1, "two": 2, "three": 3, "coincident": 40, "within_one_leaf": 8, "random_soup": 3000}[case]
    if case == "coincident":
        c = np.array([1.0, 2.0, 3.0])
        tris = c + rng.normal(size=(n, 3, 3)) * np.array([1.0, 1.0, 0.0])
        tris -= tris.mean(axis=1, keepdims=True) - c  # every centroid exactly at c
    else:
        base = rng.uniform(-5, 5, size=(n, 1, 3))
        tris = base + rng.normal(size=(n, 3, 3))
    tris = tris.astype(np.float32)
    mat = np.zeros(n, np.uint16)
    mats = [([0.7] * 3, None, None)]
    leaf = 8 if case == "within_one_leaf" else 2
    with api.Scene(gpu_ctx, tris, mat, mats, leaf_size=leaf, width=4, builder=builder) as sc:
        info = sc.info()
        assert info["builder"] == builder and info["width"] == 4
        m = 2000
        o = rng.uniform(-8, 8, size=(m, 3)).astype(np.float32)
        tgt = tris[rng.integers(0, n, m)].mean(axis=1)
        d = tgt - o
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        t, idx = sc.query_closest(o, d)
        q = rng.uniform(-8, 8, size=(m, 3)).astype(np.float32)
        occ = sc.query_occluded(o, q)
    osc = oracle.OracleScene(tris, mat, mats)
    hits = 0
    for i in range(m):
        ot, oi = osc.closest_hit(o[i], d[i])
        assert idx[i] == oi, (i, idx[i], oi)
        if oi >= 0:
            hits += 1
            assert np.float32(t[i]) == np.float32(ot)
        assert occ[i] == osc.any_hit(o[i], q[i])
    assert hits > m // 4


def test_device_sah_tree_quality_equals_host_sah(gpu_ctx):
    """The device SAH build of the 200k-triangle hall makes the host build's
    splits, so its 4-wide tree needs the node visits and triangle tests per
    query of the host tree (within 2%: the collapse's choice among children
    of equal area may differ; an instrumented render of the same rays counts
    them), and builds in under a quarter of the host build's time (without
    the host's 4-wide collapse, which the device build includes).  The PLOC
    tree's counts are printed beside them."""
    s = scenes.hall()
    kw = dict(recorders=s["recorders"], n_rays=200_000, src_pos=s["src"], air=s["air"], max_bounces=100,
              cutoff=1e-8, n_bins=132_300, instrument=True)
    out = {}
    for b in (2, 3, 4):
        with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"], builder=b) as sc:
            st = api.render(gpu_ctx, sc, **kw).stats
            out[b] = (st["node_visits_closest"] / st["extend_queries"], st["tri_tests_closest"] / st["extend_queries"],
                      st["node_visits_shadow"] / st["shadow_queries"], st["tri_tests_shadow"] / st["shadow_queries"],
                      sc.info()["build_ms"], sc.info()["n_nodes"])
    for b, name in ((2, "host SAH"), (3, "PLOC"), (4, "device SAH")):
        print("%-10s visits/tests closest %.3f %.3f shadow %.3f %.3f build %.1f ms nodes %d" % ((name,) + out[b]),
              flush=True)
    for k in range(4):
        assert out[4][k] <= 1.02 * out[2][k], (k, out[4], out[2])
    assert out[4][4] < 0.25 * out[2][4]


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


@pytest.mark.parametrize("case", ["stereo_shoebox", "mesh_source", "stereo_hall"])
def test_tail_kernel_against_the_oracle(gpu_ctx, oracle, case):
    """Renders that the tail kernel carries (no hit log, so the tail launch is
    on): the 12-triangle stereo shoebox and its mesh source (the whole render
    in k_tail), and the 20k-triangle hall with its stereo recorder (3000 rays
    per band: the live rays are below 2^20 from the first check, so k_tail
    takes them after four bounce iterations), against the oracle: segment
    and contribution counts exactly, both ears' histograms within the gate."""
    if case == "stereo_hall":
        s = scenes.hall(target_tris=20_000)
        n_rays, kw = 3000, dict(max_bounces=100, cutoff=1e-8, n_bins=132_300)
    else:
        s = scenes.shoebox(stereo=True)
        n_rays, kw = 20_000, dict(max_bounces=200, cutoff=1e-8, n_bins=88_200)
    src_tris = None
    if case == "mesh_source":
        src_tris = np.array([[[4.0, 3.0, 1.0], [5.0, 3.0, 1.0], [4.0, 4.0, 1.2]]], np.float32)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        g = api.render(gpu_ctx, sc, recorders=s["recorders"], n_rays=n_rays, src_pos=s["src"], air=s["air"],
                       src_tris=src_tris, **kw)
    assert g.stats["bounces_launched"] <= 8  # the tail launch took the rest
    ora = oracle_bands(oracle, s, (0, 1, 2), n_rays, src_tris=src_tris, **kw)
    assert g.stats["segments"] == sum(int(o["segments"]) for o in ora.values())
    assert g.stats["contributions"] == sum(int(o["contributions"]) for o in ora.values())
    for b in (0, 1, 2):
        o = ora[b]
        assert_hist_parity(api.raw_to_float(g.raw[b]), api.raw_to_float(g.abs_raw[b]), o["hist_signed"],
                           o["hist_abs"], n_rays * (kw["max_bounces"] + 1))


@pytest.mark.parametrize("width", [2, 4])
def test_fetch_traversal_is_bitwise_identical(gpu_ctx, width):
    """Persistent lanes that refill from the queue walk every ray exactly as
    one-ray-per-lane launches do: same histograms, same node visits and
    triangle tests, for both BVH widths, block sizes, stack splits, refill
    thresholds (1..64 idle lanes), claim sizes (1..4096 entries) and queue
    segments (1..64), with a
    stereo recorder and a mesh source (its k = 0 connections are shadow
    requests of the first iteration)."""
    s = scenes.hall(target_tris=20_000)
    src_tris = np.array([[[4.0, 13.0, 1.0], [5.0, 13.0, 1.0], [4.0, 14.0, 1.2]]], np.float32)
    kw = dict(recorders=s["recorders"], n_rays=20_000, src_pos=s["src"], air=s["air"], max_bounces=100,
              cutoff=1e-8, n_bins=132300, instrument=True)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"], width=width) as sc:
        res = []
        for trav, blk, lds, refill, chunk, segs in [(1, 256, 16, 0, 0, 0), (2, 256, 16, 1, 1, 1),
                                                    (3, 128, 32, 64, 4096, 64), (3, 256, 16, 0, 0, 0),
                                                    (3, 256, 16, 16, 0, 7)]:
            gpu_ctx.set_tuning(traversal=trav, block_extend=blk, block_shadow=blk, stack_lds=lds, refill=refill,
                               fetch_chunk=chunk, fetch_segments=segs)
            res.append(api.render(gpu_ctx, sc, **kw))
            res.append(api.render(gpu_ctx, sc, src_tris=src_tris, **kw))
        gpu_ctx.set_tuning()
    keys = ("segments", "shadow_rays", "contributions", "node_visits_closest", "tri_tests_closest",
            "node_visits_shadow", "tri_tests_shadow")
    for i in range(2, len(res)):
        ref = res[i % 2]
        np.testing.assert_array_equal(ref.raw, res[i].raw)
        np.testing.assert_array_equal(ref.abs_raw, res[i].abs_raw)
        assert all(ref.stats[k] == res[i].stats[k] for k in keys), (ref.stats, res[i].stats)


def test_fetch_traversal_batch(gpu_ctx):
    rooms = [scenes.room(seed=30 + i, target_tris=[12, 2000, 9000][i % 3]) for i in range(5)]
    objs = [api.Scene(gpu_ctx, r["tris"], r["mat"], r["materials"]) for r in rooms]
    try:
        items = [(sc, dict(recorders=r["recorders"], src_pos=r["src"], air=r["air"], max_bounces=60, cutoff=1e-7,
                           n_bins=44100, n_rays=3000 + 500 * i, instance=i))
                 for i, (r, sc) in enumerate(zip(rooms, objs))]
        gpu_ctx.set_tuning(traversal=1)
        a, _, _ = api.render_batch(gpu_ctx, items)
        gpu_ctx.set_tuning(traversal=3)
        b, _, _ = api.render_batch(gpu_ctx, items)
        gpu_ctx.set_tuning()
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x.raw, y.raw)
            np.testing.assert_array_equal(x.abs_raw, y.abs_raw)
    finally:
        for o in objs:
            o.close()


def test_graph_bounce_loop_is_bitwise_identical(gpu_ctx):
    """The bounce loop captured once as a HIP graph (2 or 6 iterations) and
    replayed until the queues drain, with a host synchronisation between
    replays, gives the eager loop's histograms on every render; a small
    queue (ray_chunk 3000) makes each render replay it many times and
    regenerate rays in most iterations.  With 2 iterations the replays run
    to the end (tail_rays = 1); with 6 the last live rays go to the tail
    launch after a replay, as in the eager loop."""
    s = scenes.hall(target_tris=20_000)
    src_tris = np.array([[[4.0, 13.0, 1.0], [5.0, 13.0, 1.0], [4.0, 14.0, 1.2]]], np.float32)
    kw = dict(recorders=s["recorders"], n_rays=20_000, src_pos=s["src"], air=s["air"], max_bounces=100,
              cutoff=1e-8, n_bins=132300)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        gpu_ctx.set_tuning()
        ref = [api.render(gpu_ctx, sc, **kw), api.render(gpu_ctx, sc, src_tris=src_tris, **kw)]
        for g, trav, tail in ((2, 1, 1), (6, 3, 0)):
            gpu_ctx.set_tuning(graph_iters=g, ray_chunk=3000, traversal=trav, tail_rays=tail)
            for rep in range(2):
                for j, extra in enumerate(({}, {"src_tris": src_tris})):
                    r = api.render(gpu_ctx, sc, **kw, **extra)
                    np.testing.assert_array_equal(ref[j].raw, r.raw)
                    np.testing.assert_array_equal(ref[j].abs_raw, r.abs_raw)
                    np.testing.assert_array_equal(ref[j].value, r.value)
                    assert r.stats["segments"] == ref[j].stats["segments"]
        gpu_ctx.set_tuning()


def test_hall_render_parity(gpu_ctx, oracle):
    """The 200k-triangle hall with its stereo recorder (the bench workload's
    scene) rendered through the C ABI against the brute-force oracle on the
    same rays: hit sequences, contribution counts and histograms."""
    s = scenes.hall()
    n_rays, tr, ts = 600, 200, 100
    kw = dict(max_bounces=100, cutoff=1e-8, n_bins=132300)
    print("hall render parity", len(s["tris"]), "triangles", flush=True)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        g = api.render(gpu_ctx, sc, recorders=s["recorders"], n_rays=n_rays, src_pos=s["src"], air=s["air"],
                       trace_rays=tr, trace_segs=ts, **kw)
    ora = oracle_bands(oracle, s, (0, 1, 2), n_rays, trace_rays=tr, trace_segs=ts, **kw)
    assert g.stats["segments"] == sum(int(o["segments"]) for o in ora.values())
    assert g.stats["contributions"] == sum(int(o["contributions"]) for o in ora.values())
    for b in (0, 1, 2):
        o = ora[b]
        np.testing.assert_array_equal(g.trace_hits[b], o["trace_hits"])
        assert_hist_parity(api.raw_to_float(g.raw[b]), api.raw_to_float(g.abs_raw[b]), o["hist_signed"],
                           o["hist_abs"], n_rays * 101)


def test_hall_full_size_is_launch_shape_independent(gpu_ctx):
    """The bench workload at full size (200k triangles, 3 x 10M rays): the
    raw histograms are bitwise the same under another traversal, queue size,
    refill setting and graph replay (integer sums do not depend on order)."""
    s = scenes.hall()
    kw = dict(recorders=s["recorders"], n_rays=10_000_000, src_pos=s["src"], air=s["air"], max_bounces=100,
              cutoff=1e-8, n_bins=132300, want_value=False)
    print("hall full size, two launch shapes", flush=True)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        a = api.render(gpu_ctx, sc, **kw)
        gpu_ctx.set_tuning(traversal=1, ray_chunk=1 << 21, refill=48, graph_iters=8)
        try:
            b = api.render(gpu_ctx, sc, **kw)
        finally:
            gpu_ctx.set_tuning()
    assert a.stats["segments"] == b.stats["segments"] > 9e8
    assert np.array_equal(a.raw, b.raw) and np.array_equal(a.abs_raw, b.abs_raw)


def test_keyframe_and_sound_select_the_stream(gpu_ctx, oracle):
    """A (sound, keyframe) context's paths against the oracle fed the same
    Philox context word, band | sound << 8 | (keyframe + 1) << 16: the hit
    sequences are bit-exact, so the device forms the word as the oracle does."""
    s = scenes.shoebox(stereo=True)
    n_rays, tr, ts = 1500, 128, 48
    kw = dict(max_bounces=50, cutoff=1e-6, n_bins=88200)
    sound, kf = 2, 5
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        g = api.render(gpu_ctx, sc, recorders=s["recorders"], n_rays=n_rays, src_pos=s["src"], air=s["air"],
                       trace_rays=tr, trace_segs=ts, sound=sound, keyframe=kf, **kw)
    osc = oracle.OracleScene(s["tris"], s["mat"], s["materials"])
    for b in (0, 1, 2):
        o = oracle.render(osc, band=b, air=s["air"][b], recorders=s["recorders"], n_rays=n_rays, src_pos=s["src"],
                          ctx=b | sound << 8 | (kf + 1) << 16, n_threads=8, trace_rays=tr, trace_segs=ts, **kw)
        np.testing.assert_array_equal(g.trace_hits[b], o["trace_hits"])
        np.testing.assert_array_equal(g.trace_nseg[b], o["trace_nseg"])
        assert_hist_parity(api.raw_to_float(g.raw[b]), api.raw_to_float(g.abs_raw[b]), o["hist_signed"],
                           o["hist_abs"], n_rays * 51)


@pytest.mark.parametrize("tail", [1 << 30, 50_000, 0])
def test_tail_kernel_is_bitwise_identical(gpu_ctx, tail):
    """tail_rays: once every ray is emitted, the live ones traced to their ends
    in one launch.  The histograms are the bounce loop's (tail_rays = 1: no
    tail launch) bit for bit, and so are the segment, shadow-ray and
    contribution counts (single renders with up to 999 bounces, and a batch).
    2^30 hands almost the whole render to the tail kernel, 50,000 switches
    late, 0 is the default (the whole render for the 12-triangle boxes, the
    last 2^20 live rays for the hall and the rooms of up to 9000 triangles)."""
    emitter = np.array([[[4.0, 3.0, 1.0], [5.0, 3.0, 1.0], [4.0, 4.0, 1.2]]], np.float32)
    cases = [(scenes.shoebox(stereo=True, spec=0.5), dict(n_rays=30_000, max_bounces=999, cutoff=1e-8, n_bins=88200)),
             (scenes.shoebox(spec=0.2), dict(n_rays=20_000, max_bounces=999, cutoff=1e-8, n_bins=88200,
                                             src_tris=emitter)),
             (scenes.hall(target_tris=20_000), dict(n_rays=20_000, max_bounces=100, cutoff=1e-8, n_bins=132_300))]
    for s, kw in cases:
        print(f"tail {tail}: {s['name']}", flush=True)
        with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
            args = dict(recorders=s["recorders"], src_pos=s["src"], air=s["air"], **kw)
            gpu_ctx.set_tuning(tail_rays=1)
            a = api.render(gpu_ctx, sc, **args)
            gpu_ctx.set_tuning(tail_rays=tail)
            try:
                b = api.render(gpu_ctx, sc, **args)
            finally:
                gpu_ctx.set_tuning()
        np.testing.assert_array_equal(a.raw, b.raw)
        np.testing.assert_array_equal(a.abs_raw, b.abs_raw)
        for k in ("segments", "shadow_rays", "contributions"):
            assert a.stats[k] == b.stats[k], k
        assert b.stats["extend_queries"] < b.stats["segments"]  # the tail's segments are its own
    rooms = [scenes.room(seed=40 + i, target_tris=[12, 2000, 9000][i]) for i in range(3)]
    objs = [api.Scene(gpu_ctx, r["tris"], r["mat"], r["materials"]) for r in rooms]
    try:
        items = [(o, dict(recorders=r["recorders"], src_pos=r["src"], air=r["air"], n_rays=[3000, 8000, 5000][i],
                          instance=i, max_bounces=300, cutoff=1e-8, n_bins=44100))
                 for i, (r, o) in enumerate(zip(rooms, objs))]
        gpu_ctx.set_tuning(tail_rays=1)
        a, sa, _ = api.render_batch(gpu_ctx, items)
        gpu_ctx.set_tuning(tail_rays=tail)
        try:
            b, sb, _ = api.render_batch(gpu_ctx, items)
        finally:
            gpu_ctx.set_tuning()
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x.raw, y.raw)
        assert sa["segments"] == sb["segments"] and sa["contributions"] == sb["contributions"]
    finally:
        for o in objs:
            o.close()
