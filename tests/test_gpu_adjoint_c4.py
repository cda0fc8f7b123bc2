"""The adjoint at the differentiation workload itself: the 12 x 9 x 5 m
shoebox (6 walls x 3 bands absorption + 6 specularities), 1M rays, up to 999
bounces with the 1e-8 energy cut-off, a log-EDC loss on 4096 coarse bins
against a target rendered at the true parameters, evaluated at parameters
+-20% off the target's.

* d/drefl against central differences of the loss (h = 1e-3, common random
  numbers) with the cut-off on (paths whose energy crosses 1e-8 end one
  bounce earlier or later: a boundary term the adjoint does not carry) and
  off (999 bounces for every path: nothing moves).
* Both gradients against the FP64 oracle on the same 1M rays, all bands.
* d/dspec in the score form against central differences at h = 0.03 over
  independent seeds (the differences of one seed are dominated by paths that
  switch walls, so the truth is their mean), and the local form beside it."""
import numpy as np
import pytest

from ear_amd import api, fit, scenes

pytestmark = pytest.mark.gpu

N_MAT = 6


@pytest.fixture(scope="module")
def c4():
    s = scenes.shoebox(12.0, 9.0, 5.0, spec=0.3, alpha_range=(0.1, 0.3), src=(3.0, 4.0, 2.0), rec=(8.0, 5.0, 1.7))
    alpha_t, spec_t = s["alpha"], np.full(N_MAT, 0.3)
    sign = np.random.default_rng(5).choice([-1, 1], size=(N_MAT, 4))
    alpha0, spec0 = alpha_t * (1 + 0.2 * sign[:, :3]), spec_t * (1 + 0.2 * sign[:, 3])
    return dict(s=s, alpha_t=alpha_t, spec_t=spec_t, alpha0=alpha0, spec0=spec0)


def _kw(c, **extra):
    s = c["s"]
    kw = dict(recorders=s["recorders"], src_pos=s["src"], air=s["air"], n_rays=1_000_000, max_bounces=999,
              cutoff=1e-8, n_bins=4096 * 11, abs_hist=True, want_raw=False)
    kw.update(extra)
    return kw


class _Loss:
    """The C4 loss of one scene handle: target at the true parameters
    (16M rays, its own seed), the loss at any (alpha, spec)."""

    def __init__(self, ctx, c, **kw):
        self.ctx, self.c, self.kw = ctx, c, _kw(c, **kw)
        s = c["s"]
        self.sc = api.Scene(ctx, s["tris"], s["mat"], fit.materials(c["alpha_t"], c["spec_t"]))
        ref = api.render(ctx, self.sc, **dict(self.kw, n_rays=16_000_000, seed=0x5EED))
        self.target = fit.log_edc(fit.coarse_bins(ref.abs_value))
        self.mask = fit.decay_mask(self.target)

    def __call__(self, alpha, spec, **kw):
        self.sc.set_materials(fit.materials(alpha, spec))
        r = api.render(self.ctx, self.sc, **dict(self.kw, **kw))
        return fit.log_edc_loss(r.abs_value, self.target, self.mask)

    def backward(self, adj, **kw):
        self.sc.set_materials(fit.materials(self.c["alpha0"], self.c["spec0"]))
        return api.render_backward(self.ctx, self.sc, adj_abs=adj, **dict(self.kw, **kw))

    def close(self):
        self.sc.close()


def _refl_fd(L, c, h):
    a0, s0 = c["alpha0"], c["spec0"]
    fd = np.zeros((N_MAT, 3))
    for m in range(N_MAT):
        for b in range(3):
            ap, am = a0.copy(), a0.copy()
            ap[m, b] -= h  # refl = 1 - alpha
            am[m, b] += h
            fd[m, b] = (L(ap, s0)[0] - L(am, s0)[0]) / (2 * h)
    return fd


@pytest.mark.parametrize("cutoff,rtol", [(1e-8, 1e-2), (0.0, 1e-3)])
def test_c4_refl_gradient_matches_finite_differences(gpu_ctx, c4, cutoff, rtol):
    """All 18 reflection coefficients, the full-size loss.  Measured: 0.33%
    at most with the cut-off on (the mid band, whose paths are longest), 2e-4
    with it off."""
    L = _Loss(gpu_ctx, c4, cutoff=cutoff)
    try:
        _, adj = L(c4["alpha0"], c4["spec0"])
        g_r, _ = L.backward(adj, score=True)
        fd = _refl_fd(L, c4, 1e-3)
    finally:
        L.close()
    rel = np.abs(g_r - fd) / np.abs(fd)
    print("cutoff", cutoff, "max relative difference", rel.max())
    assert np.all(np.abs(fd) > 0)
    assert rel.max() <= rtol, rel


@pytest.mark.timeout(600)
def test_c4_gradient_matches_fp64_oracle(gpu_ctx, oracle, c4):
    """The device gradients (local and score form) of the full-size loss
    adjoint against the FP64 oracle on the same 1M rays and the same
    continuations, every band: the tangents and S accumulate in binary32 on
    the device, so to 1e-5 of the largest entry."""
    L = _Loss(gpu_ctx, c4)
    s = c4["s"]
    try:
        _, adj = L(c4["alpha0"], c4["spec0"])
        dev = {form: L.backward(adj, **{form: True}) for form in ("local", "score")}
    finally:
        L.close()
    osc = oracle.OracleScene(s["tris"], s["mat"], fit.materials(c4["alpha0"], c4["spec0"]))
    n = 1_000_000
    for form, (g_r, g_s) in dev.items():
        for b in range(3):
            print("oracle", form, "band", b, flush=True)
            o_r, o_s = oracle.gradient(osc, band=b, air=s["air"][b], recorders=s["recorders"], n_rays=n,
                                       adj_abs=adj[b], src_pos=s["src"], max_bounces=999, cutoff=1e-8,
                                       n_bins=4096 * 11, n_threads=16, **{form: True})
            o_r, o_s = o_r / n, o_s / n
            np.testing.assert_allclose(g_r[:, b], o_r, rtol=0, atol=1e-5 * np.abs(o_r).max())
            np.testing.assert_allclose(g_s[:, b], o_s, rtol=0, atol=1e-5 * np.abs(o_s).max())


@pytest.mark.timeout(600)
def test_c4_score_spec_gradient_matches_finite_differences(gpu_ctx, c4):
    """d loss / d spec of the 6 walls: the score form averaged over 8 seeds of
    4M rays against central differences at h = 0.03 over 8 other seeds (each
    difference of the loss with common random numbers).  Both estimate the
    derivative of the expected loss, so they agree within their standard
    errors (at most 4 per wall, chi^2 over the walls below 20).  The local
    form's estimate is printed beside them: it misses the score and rim
    terms, which carry almost all of the derivative here."""
    L = _Loss(gpu_ctx, c4, n_rays=4_000_000)
    a0, s0 = c4["alpha0"], c4["spec0"]
    seeds, h = 8, 0.03
    try:
        fd = np.zeros((seeds, N_MAT))
        for k in range(seeds):
            for m in range(N_MAT):
                sp, sm = s0.copy(), s0.copy()
                sp[m] += h
                sm[m] -= h
                fd[k, m] = (L(a0, sp, seed=100 + k)[0] - L(a0, sm, seed=100 + k)[0]) / (2 * h)
        gs, gl = np.zeros((seeds, N_MAT)), np.zeros((seeds, N_MAT))
        for k in range(seeds):
            _, adj = L(a0, s0, seed=300 + k)
            gs[k] = L.backward(adj, score=True, seed=300 + k)[1].sum(axis=1)
            gl[k] = L.backward(adj, local=True, seed=300 + k)[1].sum(axis=1)
    finally:
        L.close()
    se = np.hypot(gs.std(0, ddof=1), fd.std(0, ddof=1)) / np.sqrt(seeds)
    z = (gs.mean(0) - fd.mean(0)) / se
    cos = float(gl.mean(0) @ fd.mean(0) / np.linalg.norm(gl.mean(0)) / np.linalg.norm(fd.mean(0)))
    print("fd", fd.mean(0), "+-", fd.std(0, ddof=1) / np.sqrt(seeds))
    print("score", gs.mean(0), "+-", gs.std(0, ddof=1) / np.sqrt(seeds), "z", z)
    print("local", gl.mean(0), "cosine with fd", cos)
    assert np.all(np.abs(z) < 4) and float((z ** 2).sum()) < 20, z


def test_fit_by_the_score_form_matches_a_flutter_decay(gpu_ctx):
    """A long room (12 x 4 x 3 m) with absorbing floor and ceiling (alpha 0.7)
    and reflective side walls (0.03): specular paths flutter between the walls
    and miss the absorbers, so the decay depends on the specularities (the
    loss of a common specularity has its minimum at the target's 0.35).
    fit.fit_loop on the six walls' specularities alone (the absorptions held
    at the truth), from 0.15, 16 steps of 250k rays.  With the score form
    the loss falls more than 5x (measured 13x: 0.114 -> 0.0087, the target
    render's own noise level); the local form's gradient, which misses the
    score and rim terms, ends with a loss higher than it started (measured
    0.24).  The six specularities are not identified one by one (floor and
    ceiling end lower than the walls): the decay constrains their effect."""
    s = scenes.shoebox(12.0, 4.0, 3.0, spec=0.3, alpha_range=(0.1, 0.3), src=(2.0, 2.0, 1.5), rec=(10.0, 2.2, 1.4))
    alpha = np.array([[0.7] * 3, [0.7] * 3] + [[0.03] * 3] * 4)
    kw = dict(recorders=s["recorders"], src_pos=s["src"], air=s["air"], n_rays=250_000, max_bounces=999,
              cutoff=1e-8, n_bins=4096 * 11, abs_hist=True, want_raw=False)
    with api.Scene(gpu_ctx, s["tris"], s["mat"], fit.materials(alpha, np.full(N_MAT, 0.35))) as sc:
        target = fit.log_edc(fit.coarse_bins(api.render(gpu_ctx, sc, **dict(kw, n_rays=8_000_000,
                                                                             seed=0x5EED)).abs_value))

        def forward(mats):
            sc.set_materials(mats)
            return api.render(gpu_ctx, sc, **kw).abs_value

        result = {}
        for form in ("score", "local"):
            def backward(mats, adj, form=form):
                sc.set_materials(mats)
                g_r, g_s = api.render_backward(gpu_ctx, sc, adj_abs=adj, **{form: True}, **kw)
                return np.zeros_like(g_r), g_s

            st = fit.fit_loop(forward, backward, target, alpha, np.full(N_MAT, 0.15), steps=16, lr=0.02)
            result[form] = (st.history[0][0], st.history[-1][0])
            print(form, "loss", round(st.history[0][0], 5), "->", round(st.history[-1][0], 5), "specularities",
                  np.round(st.history[-1][2], 3), flush=True)
    (s0, s1), (l0, l1) = result["score"], result["local"]
    assert s1 < s0 / 5 and s1 < l1, result
