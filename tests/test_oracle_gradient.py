"""The FP64 oracle gradient (or_gradient) pinned by central finite differences
of the oracle render with common random numbers, and the log-EDC loss of the
fitting module pinned by finite differences of its own value.

Box 12 x 9 x 5 m, alpha ~ U(0.1, 0.3), specularity 0.3, cut-off off (paths end
on the bounce count), no refraction: a parameter change moves no path end and
flips no reflect/refract choice."""
import numpy as np
import pytest

from ear_amd import fit, scenes

NB = 44100


@pytest.fixture(scope="module")
def room():
    return scenes.shoebox(12.0, 9.0, 5.0, spec=0.3, alpha_range=(0.1, 0.3), src=(3.0, 4.0, 2.0),
                          rec=(8.0, 5.0, 1.7))


def _adj():
    return np.cos(np.arange(NB) / 9000.0)[None, :], np.exp(-np.arange(NB) / 20000.0)[None, :]


def _loss(oracle, s, mats, adj_s, adj_a, band, **kw):
    sc = oracle.OracleScene(s["tris"], s["mat"], mats)
    r = oracle.render(sc, band=band, air=s["air"][band], recorders=s["recorders"], src_pos=s["src"], cutoff=0.0,
                      n_bins=NB, n_threads=8, **kw)
    return float((adj_s * r["hist_signed"]).sum() + (adj_a * r["hist_abs"]).sum())


def _pert(mats, m, which, band, h):
    out = [(list(a), list(b) if b is not None else None, list(c) if c is not None else None) for a, b, c in mats]
    out[m][which][band] += h
    return out


def test_refl_gradient_equals_central_difference(oracle, room):
    """Retention = refl (+ refr) enters every contribution as a product, so the
    loss is a polynomial in refl: the central difference with h = 1e-3 is
    exact up to its O(h^2) term and binary32 rounding."""
    adj_s, adj_a = _adj()
    kw = dict(n_rays=1500, max_bounces=8)
    sc = oracle.OracleScene(room["tris"], room["mat"], room["materials"])
    band = 1
    g_r, _ = oracle.gradient(sc, band=band, air=room["air"][band], recorders=room["recorders"], adj_signed=adj_s,
                             adj_abs=adj_a, src_pos=room["src"], cutoff=0.0, n_bins=NB, **kw)
    h = 1e-3
    for m in (0, 2, 5):
        fd = (_loss(oracle, room, _pert(room["materials"], m, 0, band, h), adj_s, adj_a, band, **kw) -
              _loss(oracle, room, _pert(room["materials"], m, 0, band, -h), adj_s, adj_a, band, **kw)) / (2 * h)
        assert abs(fd - g_r[m]) <= 2e-4 * abs(fd), (m, fd, g_r[m])


@pytest.mark.parametrize("max_bounces,rtol", [(1, 1e-4), (2, 1e-2)])
def test_spec_gradient_equals_central_difference(oracle, room, max_bounces, rtol):
    """With one bounce only the connection's lobe blend depends on spec (no
    direction is carried to a later vertex): exact to 1e-4.  With two, the
    reparameterised direction moves the second hit point; walls switch for a
    few rays, which the path derivative ignores (below 1%)."""
    _, adj_a = _adj()
    zero = np.zeros_like(adj_a)
    kw = dict(n_rays=3000, max_bounces=max_bounces)
    sc = oracle.OracleScene(room["tris"], room["mat"], room["materials"])
    band = 2
    _, g_s = oracle.gradient(sc, band=band, air=room["air"][band], recorders=room["recorders"], adj_abs=adj_a,
                             src_pos=room["src"], cutoff=0.0, n_bins=NB, **kw)
    h = 1e-3
    for m in (0, 1, 4):
        fd = (_loss(oracle, room, _pert(room["materials"], m, 2, band, h), zero, adj_a, band, **kw) -
              _loss(oracle, room, _pert(room["materials"], m, 2, band, -h), zero, adj_a, band, **kw)) / (2 * h)
        assert abs(fd - g_s[m]) <= rtol * abs(fd), (m, fd, g_s[m])


@pytest.mark.parametrize("theta,s", [(0, 0.3), (40, 0.3), (75, 0.45), (20, 0.05), (60, 0.0)])
def test_score_terms_match_a_pathwise_derivative(oracle, theta, s):
    """The score form's closed forms against the mathematics they restate: for
    a smooth G the derivative of E[G(v)], v = normalize((1-s) h + s r) with h
    uniform on the hemisphere, is the common-random-number difference of the
    sample means (no discontinuity, so exact up to O(eps^2)); the score form
    writes it as E[G dlog p/ds] - F E[G(v_b)] with v_b drawn by the rim
    sampler.  E[dlog p/ds] = F (G = 1) pins the density and the flux
    together; a dropped rim term or a wrong flux is off by ~F E[G]."""
    rng = np.random.default_rng(11 + theta)
    nu = np.array([0.0, 0.0, 1.0])
    r = np.array([np.sin(np.radians(theta)), 0.0, np.cos(np.radians(theta))])
    n = 1_000_000
    h = rng.standard_normal((n, 3))
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    h[h @ nu < 0] *= -1

    def blend(t):
        x = (1 - t) * h + t * r
        return x / np.linalg.norm(x, axis=1, keepdims=True)

    a = np.array([0.7, -1.1, 0.4])

    def G(w):
        return np.exp(w @ a) * (1 + w[:, 0] ** 2)

    eps = 1e-4
    truth = (G(blend(s + eps)).mean() - G(blend(s - eps)).mean()) / (2 * eps)
    v = blend(s)
    dl = np.array([oracle.score_dlogp(c, s) for c in v @ r])
    F = oracle.score_flux(float(r @ nu), s)
    vb = oracle.score_rim_dirs(400_000, nu, r, s).astype(np.float64)
    assert abs(dl.mean() - F) < 4 * dl.std() / np.sqrt(n) + 1e-9
    np.testing.assert_allclose(np.linalg.norm(vb, axis=1), 1.0, atol=1e-6)
    # every rim direction lies on the image of the rim: h = ((t v - s r)/(1-s)) . nu = 0
    c = vb @ r
    t = s * c + np.sqrt((1 - s) ** 2 - s * s * (1 - c * c))
    np.testing.assert_allclose(((t[:, None] * vb - s * r) / (1 - s)) @ nu, 0.0, atol=2e-6)
    gv = G(v) * dl
    gb = G(vb)
    est = gv.mean() - F * gb.mean()
    se = np.hypot(gv.std() / np.sqrt(n), F * gb.std() / np.sqrt(len(gb)))
    assert abs(est - truth) < 4 * se, (est, truth, se)


def test_score_gradient_equals_central_differences_over_seeds(oracle, room):
    """The score form of d/dspec (continuations included) against central
    differences of the oracle render with common random numbers at h = 0.05,
    six bounces: both are unbiased estimates of the same derivative of the
    expected loss (the differences up to O(h^2)), so over 8 seeds they agree
    within their standard errors (at most 4 of them per wall, chi^2 over the
    six walls below 20)."""
    _, adj_a = _adj()
    band = 2
    kw = dict(band=band, air=room["air"][band], recorders=room["recorders"], src_pos=room["src"], cutoff=0.0,
              n_bins=NB, max_bounces=6)
    n, seeds, h = 200_000, 8, 0.05
    sc = oracle.OracleScene(room["tris"], room["mat"], room["materials"])

    def pert(m, d):
        out = [(list(a), list(b) if b is not None else None, list(c) if c is not None else None)
               for a, b, c in room["materials"]]
        out[m][2][band] += d
        return out

    def loss(mats, seed):
        r = oracle.render(oracle.OracleScene(room["tris"], room["mat"], mats), n_rays=n, seed=seed, n_threads=8, **kw)
        return float((adj_a * r["hist_abs"]).sum())

    g = np.array([oracle.gradient(sc, adj_abs=adj_a, n_rays=n, score=True, seed=500 + k, **kw)[1]
                  for k in range(seeds)])
    fd = np.array([[(loss(pert(m, h), 900 + k) - loss(pert(m, -h), 900 + k)) / (2 * h) for m in range(6)]
                   for k in range(seeds)])
    se = np.hypot(g.std(0, ddof=1), fd.std(0, ddof=1)) / np.sqrt(seeds)
    z = (g.mean(0) - fd.mean(0)) / se
    print("score", g.mean(0), "fd", fd.mean(0), "z", z)
    assert np.all(np.abs(z) < 4) and float((z ** 2).sum()) < 20, z


def test_score_form_keeps_refl_and_rejects_what_it_cannot_do(oracle, room):
    """d/drefl does not depend on the form; specularities of 0.5 and above
    (where the blend's direction map folds) and mesh sources are refused."""
    adj_s, adj_a = _adj()
    sc = oracle.OracleScene(room["tris"], room["mat"], room["materials"])
    args = dict(band=1, air=room["air"][1], recorders=room["recorders"], adj_signed=adj_s, adj_abs=adj_a,
                src_pos=room["src"], cutoff=0.0, n_bins=NB, n_rays=500, max_bounces=12)
    full = oracle.gradient(sc, **args)
    sco = oracle.gradient(sc, score=True, **args)
    np.testing.assert_allclose(sco[0], full[0], rtol=1e-12)
    # one bounce: no direction drawn at a vertex reaches a connection, so the
    # score form is the local one (the explicit lobe term only)
    one = dict(args, max_bounces=1)
    np.testing.assert_allclose(oracle.gradient(sc, score=True, **one)[1], oracle.gradient(sc, local=True, **one)[1],
                               rtol=1e-12, atol=0)
    mats = [(a, b, [0.5, 0.5, 0.5]) for a, b, _ in room["materials"]]
    with pytest.raises(RuntimeError, match="-3"):
        oracle.gradient(oracle.OracleScene(room["tris"], room["mat"], mats), score=True, **args)


def test_local_mode_drops_only_carried_tangents(oracle, room):
    """At one bounce no tangent is carried, so the local and full gradients
    agree exactly; at several bounces refl is untouched and spec differs."""
    adj_s, adj_a = _adj()
    sc = oracle.OracleScene(room["tris"], room["mat"], room["materials"])
    args = dict(band=0, air=room["air"][0], recorders=room["recorders"], adj_signed=adj_s, adj_abs=adj_a,
                src_pos=room["src"], cutoff=0.0, n_bins=NB, n_rays=800)
    f1 = oracle.gradient(sc, max_bounces=1, **args)
    l1 = oracle.gradient(sc, max_bounces=1, local=True, **args)
    np.testing.assert_array_equal(f1[0], l1[0])
    np.testing.assert_array_equal(f1[1], l1[1])
    f6 = oracle.gradient(sc, max_bounces=6, **args)
    l6 = oracle.gradient(sc, max_bounces=6, local=True, **args)
    np.testing.assert_array_equal(f6[0], l6[0])
    assert np.any(f6[1] != l6[1])


def test_log_edc_loss_adjoint_equals_finite_differences():
    rng = np.random.default_rng(1)
    n = 4096 * 3
    t = np.arange(n)
    base = np.exp(-t / 2000.0) * (1.0 + 0.3 * rng.random(n))
    x = np.stack([base, 1.3 * base, 0.7 * base])[:, None, :]
    target = fit.log_edc(fit.coarse_bins(x * (1.0 + 0.1 * rng.random(x.shape))))
    mask = fit.decay_mask(target)
    loss, adj = fit.log_edc_loss(x, target, mask)
    assert loss > 0
    for idx in [(0, 0, 5), (1, 0, 4000), (2, 0, 9000), (0, 0, n - 1)]:
        h = 1e-6 * x[idx]
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        fd = (fit.log_edc_loss(xp, target, mask)[0] - fit.log_edc_loss(xm, target, mask)[0]) / (2 * h)
        assert abs(fd - adj[idx]) <= 1e-5 * abs(adj[idx]) + 1e-12, (idx, fd, adj[idx])
    # the loss vanishes at the target
    same, _ = fit.log_edc_loss(x, fit.log_edc(fit.coarse_bins(x)), mask)
    assert same < 1e-25


def test_fit_loop_recovers_area_weighted_absorption(oracle, room):
    """The fitting loop over the oracle (forward and local backward), 12 Adam
    steps from alpha and spec off by +-20%: the log-EDC loss falls 10x and the
    area-weighted absorption (what one decay curve per band determines) moves
    2.5x closer to the truth.  Per-wall values are not identifiable."""
    nr, nb = 8000, 4096 * 11
    s = room

    def forward(mats, seed=0xEA2):
        sc = oracle.OracleScene(s["tris"], s["mat"], mats)
        out = np.zeros((3, 1, nb))
        for b in range(3):
            r = oracle.render(sc, band=b, air=s["air"][b], recorders=s["recorders"], n_rays=nr, src_pos=s["src"],
                              max_bounces=100, cutoff=1e-8, n_bins=nb, n_threads=8, seed=seed)
            out[b, 0] = r["hist_abs"] / nr
        return out

    def backward(mats, adj):
        sc = oracle.OracleScene(s["tris"], s["mat"], mats)
        gr, gs = np.zeros((6, 3)), np.zeros((6, 3))
        for b in range(3):
            r_, s_ = oracle.gradient(sc, band=b, air=s["air"][b], recorders=s["recorders"], n_rays=nr,
                                     adj_abs=adj[b], src_pos=s["src"], max_bounces=100, cutoff=1e-8, n_bins=nb,
                                     local=True)
            gr[:, b], gs[:, b] = r_ / nr, s_ / nr
        return gr, gs

    alpha_t, spec_t = s["alpha"], np.full(6, 0.3)
    target = fit.log_edc(fit.coarse_bins(forward(fit.materials(alpha_t, spec_t))))
    sign = np.random.default_rng(5).choice([-1, 1], size=(6, 4))
    a0, s0 = alpha_t * (1 + 0.2 * sign[:, :3]), spec_t * (1 + 0.2 * sign[:, 3])
    st = fit.fit_loop(forward, backward, target, a0, s0, steps=12)
    tr = np.asarray(s["tris"], np.float64)
    area_t = 0.5 * np.linalg.norm(np.cross(tr[:, 1] - tr[:, 0], tr[:, 2] - tr[:, 0]), axis=1)
    area = np.array([area_t[np.asarray(s["mat"]) == m].sum() for m in range(6)])
    err0 = area @ (st.history[0][1] - alpha_t) / area.sum()
    err1 = area @ (st.history[-1][1] - alpha_t) / area.sum()
    assert st.history[-1][0] < st.history[0][0] / 10, [h[0] for h in st.history]
    assert np.abs(err1).max() < np.abs(err0).max() / 2.5, (err0, err1)
