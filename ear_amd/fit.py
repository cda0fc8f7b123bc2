"""Inverse design through the adjoint render: fit each wall's absorption and
specularity so that a room's energy decay matches a target.

E.A.R renders forward only; this module has no counterpart in the reference.
The loss is the mean squared difference of log energy decay curves on 4096
coarse bins (consecutive 44.1 kHz bins summed), per band and channel, over the
first 60 dB of the target's decay:

    A[b, c, j] = sum of abs_value[b, c, f*j : f*(j+1)]      f = n_bins // 4096
    E[b, c, i] = sum_{j >= i} A[b, c, j]
    loss = mean over kept (b, c, i) of (log E - log E_target)^2

Its adjoint with respect to abs_value is exact (log_edc_loss); the adjoint
render turns it into d loss / d refl and d loss / d spec per material, and
Adam steps alpha = 1 - refl and a per-wall specularity shared by the bands.
The device fits use the score form of d / d spec (EAR_FLAG_ADJ_SCORE,
unbiased at any bounce count), which needs specularities below 0.5: the
specularity bounds default to [0, 0.45].  form="local" selects the local
form (EAR_FLAG_ADJ_LOCAL), whose d / d spec keeps only each connection's
explicit lobe term and is biased on long paths.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

N_COARSE = 4096


def coarse_bins(x: np.ndarray, n_coarse: int = N_COARSE) -> np.ndarray:
    """Sums consecutive bins of the last axis into n_coarse bins (the tail
    past n_coarse * (n // n_coarse) is dropped)."""
    x = np.asarray(x, np.float64)
    f = x.shape[-1] // n_coarse
    if f < 1:
        raise ValueError(f"{x.shape[-1]} bins cannot be summed into {n_coarse}")
    return x[..., : n_coarse * f].reshape(*x.shape[:-1], n_coarse, f).sum(-1)


def log_edc(energy_coarse: np.ndarray) -> np.ndarray:
    e = np.cumsum(np.asarray(energy_coarse, np.float64)[..., ::-1], axis=-1)[..., ::-1]
    with np.errstate(divide="ignore"):
        return np.log(e)


def decay_mask(target_log_edc: np.ndarray, range_db: float = 60.0) -> np.ndarray:
    """Bins whose target EDC lies within range_db of its start."""
    drop = range_db / 10.0 * np.log(10.0)
    start = target_log_edc[..., :1]
    return np.isfinite(target_log_edc) & (target_log_edc >= start - drop) & np.isfinite(start)


def log_edc_loss(abs_value: np.ndarray, target_log_edc: np.ndarray, mask: np.ndarray,
                 n_coarse: int = N_COARSE) -> tuple[float, np.ndarray]:
    """Loss and its gradient with respect to abs_value (same shape)."""
    abs_value = np.asarray(abs_value, np.float64)
    f = abs_value.shape[-1] // n_coarse
    A = coarse_bins(abs_value, n_coarse)
    E = np.cumsum(A[..., ::-1], axis=-1)[..., ::-1]
    keep = mask & (E > 0)
    cnt = max(1, int(mask.sum()))
    r = np.zeros_like(E)
    r[keep] = np.log(E[keep]) - target_log_edc[keep]
    # bins the target keeps but the render left empty: penalise by the target's level
    lost = mask & ~(E > 0)
    loss = float((r[keep] ** 2).sum() / cnt) + float(lost.sum()) * 100.0 / cnt
    dE = np.zeros_like(E)
    dE[keep] = 2.0 * r[keep] / cnt / E[keep]
    dA = np.cumsum(dE, axis=-1)  # E_i = sum_{j >= i} A_j  =>  dL/dA_j = sum_{i <= j} dL/dE_i
    adj = np.zeros_like(abs_value)
    adj[..., : n_coarse * f] = np.repeat(dA, f, axis=-1)
    return loss, adj


class Adam:
    def __init__(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        self.m = self.v = None
        self.t = 0

    def step(self, x: np.ndarray, g: np.ndarray, lr: float | None = None) -> np.ndarray:
        if self.m is None:
            self.m = np.zeros_like(x)
            self.v = np.zeros_like(x)
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return x - (self.lr if lr is None else lr) * mh / (np.sqrt(vh) + self.eps)


@dataclass
class FitState:
    alpha: np.ndarray  # (n_mat, 3)
    spec: np.ndarray   # (n_mat,)
    history: list = field(default_factory=list)  # (loss, alpha, spec) per step


def materials(alpha: np.ndarray, spec: np.ndarray) -> list:
    """MAT rows: refl = 1 - alpha, no refraction, one specularity per wall."""
    return [([float(1.0 - a) for a in alpha[m]], [0.0, 0.0, 0.0], [float(spec[m])] * 3)
            for m in range(len(alpha))]


def fit_loop(forward, backward, target_log_edc: np.ndarray, alpha0, spec0, *, steps: int, lr: float = 0.004,
             decay: float = 5.0, momentum: float = 0.5, alpha_bounds=(0.01, 0.95), spec_bounds=(0.0, 0.45),
             n_coarse: int = N_COARSE) -> FitState:
    """forward(mats) -> abs_value (3, n_ch, n_bins); backward(mats, adj_abs) ->
    (d/drefl, d/dspec), each (n_mat, 3).  Adam (first-moment decay
    `momentum`) on (alpha, spec) with the step size lr / (1 + t / decay)."""
    st = FitState(np.array(alpha0, np.float64), np.array(spec0, np.float64))
    mask = decay_mask(target_log_edc)
    opt = Adam(lr, b1=momentum)
    n_mat = len(st.alpha)
    for t in range(steps):
        mats = materials(st.alpha, st.spec)
        loss, adj = log_edc_loss(forward(mats), target_log_edc, mask, n_coarse)
        st.history.append((loss, st.alpha.copy(), st.spec.copy()))
        g_refl, g_spec = backward(mats, adj)
        g = np.concatenate([-np.asarray(g_refl).reshape(-1), np.asarray(g_spec).sum(axis=1)])
        x = opt.step(np.concatenate([st.alpha.reshape(-1), st.spec]), g, lr / (1.0 + t / decay))
        st.alpha = np.clip(x[: 3 * n_mat].reshape(n_mat, 3), *alpha_bounds)
        st.spec = np.clip(x[3 * n_mat:], *spec_bounds)
    mats = materials(st.alpha, st.spec)
    loss, _ = log_edc_loss(forward(mats), target_log_edc, mask, n_coarse)
    st.history.append((loss, st.alpha.copy(), st.spec.copy()))
    return st


def _form_flag(form: str) -> dict:
    if form not in ("score", "local"):
        raise ValueError(f"form must be 'score' or 'local', not {form!r}")
    return {form: True}


def fit_rooms(ctx, rooms: list, target_abs_values: list, alpha0s: list, spec0s: list, *, steps: int,
              lr: float = 0.004, decay: float = 5.0, momentum: float = 0.5, alpha_bounds=(0.01, 0.95),
              spec_bounds=(0.0, 0.45), n_rays=1_000_000, max_bounces: int = 999, cutoff: float = 1e-8,
              n_bins: int = 4096 * 11, seed: int = 0xEA2, n_coarse: int = N_COARSE, form: str = "score") -> list:
    """Many rooms fitted together: every step is one ear_render_batch of all
    rooms and one ear_render_backward_batch, then one Adam step per room (the
    same update as fit_loop).  The scenes are uploaded once; each step only
    replaces their material tables.  n_rays: one count or one per room.
    form: "score" or "local", the form of d / d spec.  Returns one FitState
    per room."""
    from . import api

    flag = _form_flag(form)

    k = len(rooms)
    rays = list(n_rays) if np.ndim(n_rays) else [int(n_rays)] * k
    kws = [dict(recorders=r["recorders"], src_pos=r["src"], air=r["air"], n_rays=rays[i], max_bounces=max_bounces,
                cutoff=cutoff, n_bins=n_bins, seed=seed, abs_hist=True) for i, r in enumerate(rooms)]
    targets = [log_edc(coarse_bins(t, n_coarse)) for t in target_abs_values]
    masks = [decay_mask(t) for t in targets]
    states = [FitState(np.array(a, np.float64), np.array(sp, np.float64)) for a, sp in zip(alpha0s, spec0s)]
    opts = [Adam(lr, b1=momentum) for _ in range(k)]
    scenes = [api.Scene(ctx, r["tris"], r["mat"], materials(st.alpha, st.spec)) for r, st in zip(rooms, states)]

    def losses():
        res, _, _ = api.render_batch(ctx, [(sc, dict(kw, want_raw=False)) for sc, kw in zip(scenes, kws)])
        return [log_edc_loss(r.abs_value, t, m, n_coarse) for r, t, m in zip(res, targets, masks)]

    try:
        for t in range(steps):
            for sc, st in zip(scenes, states):
                sc.set_materials(materials(st.alpha, st.spec))
            la = losses()
            grads = api.render_backward_batch(ctx, [(sc, dict(kw, adj_abs=adj, **flag))
                                                    for sc, kw, (_, adj) in zip(scenes, kws, la)])
            for i, st in enumerate(states):
                st.history.append((la[i][0], st.alpha.copy(), st.spec.copy()))
                g_refl, g_spec = grads[i]
                n_mat = len(st.alpha)
                g = np.concatenate([-np.asarray(g_refl).reshape(-1), np.asarray(g_spec).sum(axis=1)])
                x = opts[i].step(np.concatenate([st.alpha.reshape(-1), st.spec]), g, lr / (1.0 + t / decay))
                st.alpha = np.clip(x[: 3 * n_mat].reshape(n_mat, 3), *alpha_bounds)
                st.spec = np.clip(x[3 * n_mat:], *spec_bounds)
        for sc, st in zip(scenes, states):
            sc.set_materials(materials(st.alpha, st.spec))
        for st, (loss, _) in zip(states, losses()):
            st.history.append((loss, st.alpha.copy(), st.spec.copy()))
        return states
    finally:
        for sc in scenes:
            sc.close()


def fit_room(ctx, room: dict, target_abs_value: np.ndarray, alpha0, spec0, *, steps: int, lr: float = 0.004,
             decay: float = 5.0,
             n_rays: int = 1_000_000, max_bounces: int = 999, cutoff: float = 1e-8, n_bins: int = 4096 * 11,
             seed: int = 0xEA2, form: str = "score") -> FitState:
    """fit_loop over the device: every step is one ear_render and one
    ear_render_backward of the room (tris, mat, recorders, src, air) with
    the current materials (form: "score" or "local", the form of
    d / d spec)."""
    from . import api

    flag = _form_flag(form)

    kw = dict(recorders=room["recorders"], src_pos=room["src"], air=room["air"], n_rays=n_rays,
              max_bounces=max_bounces, cutoff=cutoff, n_bins=n_bins, seed=seed, abs_hist=True)
    target = log_edc(coarse_bins(target_abs_value))
    scenes = {}

    def scene(mats):
        # one upload (and BVH) for the whole fit; each step swaps the materials
        if "s" not in scenes:
            scenes["s"] = api.Scene(ctx, room["tris"], room["mat"], mats)
            scenes["m"] = repr(mats)
        elif scenes["m"] != repr(mats):
            scenes["s"].set_materials(mats)
            scenes["m"] = repr(mats)
        return scenes["s"]

    def forward(mats):
        return api.render(ctx, scene(mats), **kw).abs_value

    def backward(mats, adj):
        return api.render_backward(ctx, scene(mats), adj_abs=adj, **flag, **kw)

    try:
        return fit_loop(forward, backward, target, alpha0, spec0, steps=steps, lr=lr, decay=decay)
    finally:
        if "s" in scenes:
            scenes["s"].close()
