"""The adjoint render (ear_render_backward) against the FP64 oracle gradient and
against central finite differences of ear_render with common random numbers.

Scene: box 12 x 9 x 5 m, one material per wall, alpha ~ U(0.1, 0.3),
specularity 0.3, point source (3, 4, 2), recorder (8, 5, 1.7).  The energy
cut-off is off (paths end on the bounce count), so a parameter change moves
no path end; refraction is off, so the reflect/refract choice never flips."""
import numpy as np
import pytest

from ear_amd import api, scenes

pytestmark = pytest.mark.gpu

NB = 44100


def _scene(stereo=False):
    return scenes.shoebox(12.0, 9.0, 5.0, spec=0.3, alpha_range=(0.1, 0.3), src=(3.0, 4.0, 2.0),
                          rec=(8.0, 5.0, 1.7), stereo=stereo)


def _adjoints(nch):
    t = np.arange(NB) / 9000.0
    a_v = np.stack([np.stack([np.cos(t + b + c) for c in range(nch)]) for b in range(3)])
    a_a = np.stack([np.stack([np.exp(-t * (1 + b)) * (1 + 0.5 * c) for c in range(nch)]) for b in range(3)])
    return a_v, a_a


def _kw(s, **extra):
    kw = dict(recorders=s["recorders"], src_pos=s["src"], air=s["air"], cutoff=0.0, n_bins=NB, abs_hist=True)
    kw.update(extra)
    return kw


@pytest.mark.parametrize("stereo,max_bounces,local", [(False, 3, False), (False, 12, False), (True, 6, False),
                                                      (False, 40, True)])
def test_backward_matches_fp64_oracle(gpu_ctx, oracle, stereo, max_bounces, local):
    s = _scene(stereo)
    n_rays = 3000
    nch = 2 if stereo else 1
    a_v, a_a = _adjoints(nch)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        g_r, g_s = api.render_backward(gpu_ctx, sc, adj_value=a_v, adj_abs=a_a, local=local,
                                       **_kw(s, n_rays=n_rays, max_bounces=max_bounces))
    osc = oracle.OracleScene(s["tris"], s["mat"], s["materials"])
    for b in range(3):
        o_r, o_s = oracle.gradient(osc, band=b, air=s["air"][b], recorders=s["recorders"], n_rays=n_rays,
                                   adj_signed=a_v[b], adj_abs=a_a[b], src_pos=s["src"], max_bounces=max_bounces,
                                   cutoff=0.0, n_bins=NB, local=local)
        # the oracle's gradient is of the raw sums; value = raw / N
        o_r, o_s = o_r / n_rays, o_s / n_rays
        # same paths, tangents in binary32 on the device: 1e-4 of the largest entry
        np.testing.assert_allclose(g_r[:, b], o_r, rtol=0, atol=1e-4 * np.abs(o_r).max())
        np.testing.assert_allclose(g_s[:, b], o_s, rtol=0, atol=1e-4 * np.abs(o_s).max())
        assert np.abs(o_s).max() > 0 and np.abs(o_r).max() > 0


def _loss(ctx, s, mats, a_v, a_a, **kw):
    with api.Scene(ctx, s["tris"], s["mat"], mats) as sc:
        r = api.render(ctx, sc, **_kw(s, **kw))
    return float((a_v * r.value).sum() + (a_a * r.abs_value).sum())


def _perturbed(mats, m, which, h):
    out = [(list(a), list(b) if b is not None else None, list(c) if c is not None else None) for a, b, c in mats]
    for band in range(3):
        out[m][which][band] += h
    return out


def test_refl_gradient_matches_finite_differences(gpu_ctx):
    """Retention is linear in refl and no path moves: the central difference
    (h = 1e-3, common random numbers) agrees to 1e-3."""
    s = _scene()
    kw = dict(n_rays=4000, max_bounces=10)
    a_v, a_a = _adjoints(1)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        g_r, _ = api.render_backward(gpu_ctx, sc, adj_value=a_v, adj_abs=a_a, **_kw(s, **kw))
    h = 1e-3
    for m in (0, 3, 5):
        fd = (_loss(gpu_ctx, s, _perturbed(s["materials"], m, 0, h), a_v, a_a, **kw) -
              _loss(gpu_ctx, s, _perturbed(s["materials"], m, 0, -h), a_v, a_a, **kw)) / (2 * h)
        assert abs(fd - g_r[m].sum()) <= 1e-3 * abs(fd), (m, fd, g_r[m])


@pytest.mark.parametrize("max_bounces", [1, 2])
def test_spec_gradient_matches_finite_differences(gpu_ctx, max_bounces):
    """Specularity moves the lobe and the reparameterised directions; with one
    or two bounces no hit leaves its wall for h = 1e-3 (few enough that the
    ignored visibility terms stay below 1%)."""
    s = _scene()
    kw = dict(n_rays=4000, max_bounces=max_bounces)
    a_v, a_a = _adjoints(1)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        _, g_s = api.render_backward(gpu_ctx, sc, adj_abs=a_a, **_kw(s, **kw))
    h = 1e-3
    for m in (0, 1, 4):
        fd = (_loss(gpu_ctx, s, _perturbed(s["materials"], m, 2, h), 0 * a_v, a_a, **kw) -
              _loss(gpu_ctx, s, _perturbed(s["materials"], m, 2, -h), 0 * a_v, a_a, **kw)) / (2 * h)
        assert abs(fd - g_s[m].sum()) <= 1e-2 * abs(fd), (m, fd, g_s[m])


def test_backward_argument_errors(gpu_ctx):
    s = _scene()
    a_v, _ = _adjoints(1)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        with pytest.raises(api.EarError, match="exclude"):
            api.render_backward(gpu_ctx, sc, adj_value=a_v, local=True, score=True, **_kw(s, n_rays=10, max_bounces=2))
        sc.set_materials([(a, b, [0.3, 0.5, 0.3]) for a, b, _ in s["materials"]])
        with pytest.raises(api.EarError, match="below 0.5") as e:
            api.render_backward(gpu_ctx, sc, adj_value=a_v, score=True, **_kw(s, n_rays=10, max_bounces=2))
        assert e.value.status_name == "EAR_ERR_INVALID_ARGUMENT"
        # the band whose specularity is 0.5 is not rendered: accepted
        api.render_backward(gpu_ctx, sc, adj_value=a_v, score=True, bands=(0, 2), **_kw(s, n_rays=10, max_bounces=2))
        sc.set_materials(s["materials"])
        with pytest.raises(api.EarError, match="point sources"):
            api.render_backward(gpu_ctx, sc, adj_value=a_v, src_tris=np.eye(3, dtype=np.float32)[None],
                                **_kw(s, n_rays=10, max_bounces=2))
        with pytest.raises(api.EarError, match="no adjoint"):
            api.render_backward(gpu_ctx, sc, **_kw(s, n_rays=10, max_bounces=2))
    mats = s["materials"] + s["materials"][:3]
    mat_ids = np.arange(len(s["mat"])) % 9
    with api.Scene(gpu_ctx, s["tris"], mat_ids, mats) as sc:
        with pytest.raises(api.EarError) as e:
            api.render_backward(gpu_ctx, sc, adj_value=a_v, **_kw(s, n_rays=10, max_bounces=2))
        assert e.value.status_name == "EAR_ERR_CAPACITY"


def test_fit_room_recovers_area_weighted_absorption(gpu_ctx):
    """fit.fit_room on the device (ear_render + local ear_render_backward per
    step), 50k rays, 12 steps from alpha and spec off by +-20%: the log-EDC loss
    falls 10x and the area-weighted absorption of every band moves 2.5x
    closer to the truth."""
    from ear_amd import fit

    s = _scene()
    alpha_t, spec_t = s["alpha"], np.full(6, 0.3)
    kw = dict(recorders=s["recorders"], src_pos=s["src"], air=s["air"], n_rays=50_000, max_bounces=100,
              cutoff=1e-8, n_bins=4096 * 11, abs_hist=True)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], fit.materials(alpha_t, spec_t)) as sc:
        target = api.render(gpu_ctx, sc, **kw).abs_value
    sign = np.random.default_rng(5).choice([-1, 1], size=(6, 4))
    a0, s0 = alpha_t * (1 + 0.2 * sign[:, :3]), spec_t * (1 + 0.2 * sign[:, 3])
    st = fit.fit_room(gpu_ctx, s, target, a0, s0, steps=12, n_rays=50_000, max_bounces=100, cutoff=1e-8)
    tr = np.asarray(s["tris"], np.float64)
    area_t = 0.5 * np.linalg.norm(np.cross(tr[:, 1] - tr[:, 0], tr[:, 2] - tr[:, 0]), axis=1)
    area = np.array([area_t[np.asarray(s["mat"]) == m].sum() for m in range(6)])
    err0 = area @ (st.history[0][1] - alpha_t) / area.sum()
    err1 = area @ (st.history[-1][1] - alpha_t) / area.sum()
    print("fit losses", [round(h[0], 5) for h in st.history], "area-weighted alpha error", err0, err1)
    assert st.history[-1][0] < st.history[0][0] / 10
    assert np.abs(err1).max() < np.abs(err0).max() / 2.5


def test_backward_batch_equals_single(gpu_ctx):
    """ear_render_backward_batch over rooms of different shapes, material
    counts, ray counts and bands: each item's gradient equals its own
    ear_render_backward (the per-wave partials are summed in another order,
    so to 1e-9 relative)."""
    rooms = [scenes.room(seed=11, target_tris=12, kind="box"), scenes.room(seed=12, target_tris=400, kind="L"),
             _scene(), scenes.room(seed=13, target_tris=2000, kind="box")]
    specs = [dict(n_rays=30_000, bands=(0, 1, 2)), dict(n_rays=12_345, bands=(1,)), dict(n_rays=50_000, bands=(0, 2)),
             dict(n_rays=7_000, bands=(0, 1, 2))]
    rng = np.random.default_rng(9)
    handles, items, singles = [], [], []
    try:
        for r, sp in zip(rooms, specs):
            sc = api.Scene(gpu_ctx, r["tris"], r["mat"], r["materials"])
            handles.append(sc)
            nb = 4096
            kw = dict(recorders=r["recorders"], src_pos=r["src"], air=r["air"], max_bounces=60, cutoff=1e-8,
                      n_bins=nb, abs_hist=True, **sp)
            a_v = rng.standard_normal((3, 1, nb))
            a_a = rng.standard_normal((3, 1, nb))
            items.append((sc, dict(kw, adj_value=a_v, adj_abs=a_a)))
            singles.append(api.render_backward(gpu_ctx, sc, adj_value=a_v, adj_abs=a_a, **kw))
        print("backward batch of", len(items), "rooms", flush=True)
        got = api.render_backward_batch(gpu_ctx, items)
        for (gr, gs), (wr, ws), sp in zip(got, singles, specs):
            for g, w in ((gr, wr), (gs, ws)):
                assert np.any(w[:, list(sp["bands"])] != 0)
                np.testing.assert_allclose(g, w, rtol=1e-9, atol=1e-12 * np.abs(w).max())
                off = [b for b in range(3) if b not in sp["bands"]]
                assert np.all(g[:, off] == 0)
    finally:
        for sc in handles:
            sc.close()


def test_set_materials_equals_fresh_upload(gpu_ctx):
    from ear_amd import fit

    s = _scene()
    kw = _kw(s, n_rays=20_000)
    m2 = fit.materials(s["alpha"] * 1.3, np.full(6, 0.4))
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        sc.set_materials(m2)
        a = api.render(gpu_ctx, sc, **kw)
        with pytest.raises(api.EarError):
            sc.set_materials(m2[:3])
    with api.Scene(gpu_ctx, s["tris"], s["mat"], m2) as sc:
        b = api.render(gpu_ctx, sc, **kw)
    assert np.array_equal(a.raw, b.raw) and np.array_equal(a.abs_raw, b.abs_raw)


def test_fit_rooms_batched(gpu_ctx):
    """fit.fit_rooms: two rooms fitted in lockstep by batched forward and
    backward renders; a room fitted alone follows the same trajectory, and
    every room's loss falls."""
    from ear_amd import fit

    rooms = [_scene(), scenes.room(seed=21, target_tris=200, kind="L")]
    kw = dict(n_rays=40_000, max_bounces=100, cutoff=1e-8)
    targets, a0s, s0s = [], [], []
    for i, r in enumerate(rooms):
        n_mat = len(r["materials"])
        spec_t = np.full(n_mat, 0.3)
        with api.Scene(gpu_ctx, r["tris"], r["mat"], fit.materials(r["alpha"], spec_t)) as sc:
            targets.append(api.render(gpu_ctx, sc, recorders=r["recorders"], src_pos=r["src"], air=r["air"],
                                      n_bins=4096 * 11, abs_hist=True, **kw).abs_value)
        sign = np.random.default_rng(30 + i).choice([-1, 1], size=(n_mat, 4))
        a0s.append(r["alpha"] * (1 + 0.2 * sign[:, :3]))
        s0s.append(spec_t * (1 + 0.2 * sign[:, 3]))
    states = fit.fit_rooms(gpu_ctx, rooms, targets, a0s, s0s, steps=8, **kw)
    alone = fit.fit_room(gpu_ctx, rooms[0], targets[0], a0s[0], s0s[0], steps=8, **kw)
    print("batched fit losses", [[round(h[0], 5) for h in st.history] for st in states], flush=True)
    # 8 steps from +-20%: the shoebox's loss falls ~5x, the L room's ~2x
    for st in states:
        assert st.history[-1][0] < 0.6 * st.history[0][0]
    for (l1, a1, s1), (l2, a2, s2) in zip(states[0].history, alone.history):
        assert abs(l1 - l2) <= 1e-6 * max(abs(l2), 1e-12)
        np.testing.assert_allclose(a1, a2, rtol=1e-6)
        np.testing.assert_allclose(s1, s2, rtol=1e-6, atol=1e-9)


def test_backward_is_linear_in_the_adjoint_full_size(gpu_ctx):
    """The differentiation workload at full size (12 x 9 x 5 m shoebox, 1M
    rays, 4096-bin energy histograms): the gradient is linear in the
    adjoint, g(a1 + 2 a2) = g(a1) + 2 g(a2), for the full and the local form.
    The local form queues each connection's weights in single precision, so
    the identity holds to ~1e-8 relative (measured 8.8e-9), not to the
    summation order alone."""
    s = _scene()
    nb = 4096 * 11
    kw = dict(recorders=s["recorders"], src_pos=s["src"], air=s["air"], n_rays=1_000_000, max_bounces=999,
              cutoff=1e-8, n_bins=nb, abs_hist=True)
    t = np.arange(nb) / 44100.0
    a1 = np.broadcast_to(np.exp(-3.0 * t), (3, 1, nb)).copy()
    a2 = np.broadcast_to(np.cos(40.0 * t), (3, 1, nb)).copy()
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        for local in (True, False):
            print("linearity, local =", local, flush=True)
            g1 = api.render_backward(gpu_ctx, sc, adj_abs=a1, local=local, **kw)
            g2 = api.render_backward(gpu_ctx, sc, adj_abs=a2, local=local, **kw)
            g12 = api.render_backward(gpu_ctx, sc, adj_abs=a1 + 2 * a2, local=local, **kw)
            for k in range(2):
                want = g1[k] + 2 * g2[k]
                np.testing.assert_allclose(g12[k], want, rtol=1e-7, atol=1e-9 * np.abs(want).max())


@pytest.mark.parametrize("stereo,max_bounces,cutoff,refract", [(False, 12, 0.0, False), (True, 6, 0.0, False),
                                                               (False, 999, 1e-8, False), (False, 30, 1e-8, True)])
def test_score_form_matches_fp64_oracle(gpu_ctx, oracle, stereo, max_bounces, cutoff, refract):
    """EAR_FLAG_ADJ_SCORE against the oracle's score form on the same rays
    and the same continuation paths (spawn decisions, rim directions and
    their transport come from the vertex streams both draw alike).  refract:
    two walls transmit 0.2 of each band (the blend then targets the incoming
    direction about the negated normal, src/Scene.cpp:65-70)."""
    s = _scene(stereo)
    if refract:
        s["materials"] = [([0.75 * x for x in a], [0.2, 0.2, 0.2] if m in (0, 3) else b, c)
                          for m, (a, b, c) in enumerate(s["materials"])]
    n_rays = 20_000
    nch = 2 if stereo else 1
    a_v, a_a = _adjoints(nch)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], s["materials"]) as sc:
        g_r, g_s = api.render_backward(gpu_ctx, sc, adj_value=a_v, adj_abs=a_a, score=True,
                                       **_kw(s, n_rays=n_rays, max_bounces=max_bounces, cutoff=cutoff))
    osc = oracle.OracleScene(s["tris"], s["mat"], s["materials"])
    for b in range(3):
        o_r, o_s = oracle.gradient(osc, band=b, air=s["air"][b], recorders=s["recorders"], n_rays=n_rays,
                                   adj_signed=a_v[b], adj_abs=a_a[b], src_pos=s["src"], max_bounces=max_bounces,
                                   cutoff=cutoff, n_bins=NB, score=True)
        o_r, o_s = o_r / n_rays, o_s / n_rays
        np.testing.assert_allclose(g_r[:, b], o_r, rtol=0, atol=1e-4 * np.abs(o_r).max())
        np.testing.assert_allclose(g_s[:, b], o_s, rtol=0, atol=1e-4 * np.abs(o_s).max())
        assert np.abs(o_s).max() > 0


def test_score_form_batch_equals_single(gpu_ctx):
    """ear_render_backward_batch in the score form: each item's gradient
    equals its own ear_render_backward (continuations included) up to the
    summation order."""
    rooms = [_scene(), scenes.room(seed=12, target_tris=400, kind="L"), scenes.room(seed=13, target_tris=2000)]
    specs = [dict(n_rays=40_000, bands=(0, 1, 2)), dict(n_rays=12_345, bands=(1,)), dict(n_rays=25_000, bands=(0, 2))]
    rng = np.random.default_rng(19)
    handles, items, singles = [], [], []
    try:
        for r, sp in zip(rooms, specs):
            sc = api.Scene(gpu_ctx, r["tris"], r["mat"], r["materials"])
            handles.append(sc)
            kw = dict(recorders=r["recorders"], src_pos=r["src"], air=r["air"], max_bounces=200, cutoff=1e-8,
                      n_bins=4096, abs_hist=True, **sp)
            a_a = rng.standard_normal((3, 1, 4096))
            items.append((sc, dict(kw, adj_abs=a_a, score=True)))
            singles.append(api.render_backward(gpu_ctx, sc, adj_abs=a_a, score=True, **kw))
        got = api.render_backward_batch(gpu_ctx, items)
        for (gr, gs), (wr, ws) in zip(got, singles):
            for g, w in ((gr, wr), (gs, ws)):
                np.testing.assert_allclose(g, w, rtol=1e-9, atol=1e-12 * np.abs(w).max())
        with pytest.raises(api.EarError, match="mixes"):
            api.render_backward_batch(gpu_ctx, [items[0], (items[1][0], dict(items[1][1], score=False))])
    finally:
        for sc in handles:
            sc.close()
