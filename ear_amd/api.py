"""Python front end of the C ABI (include/ear.h), used by tests and bench.py.

Mirrors the reference's render unit: one call renders the (sound, keyframe,
band) contexts that SceneContext::operator() runs through Scene::Render
(src/SceneContext.h:46-48, src/Scene.cpp:111-318), for several bands at once.
Errors come back as EarError carrying the ear_status, the way the reference's
exceptions reach main (src/EAR.cpp:404-408).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _ffi as F

FRAC_SCALE = 2.0 ** -F.EAR_FRAC_BITS


class EarError(RuntimeError):
    def __init__(self, status: int, message: str = ""):
        self.status = status
        self.status_name = F.STATUS_NAMES.get(status, str(status))
        super().__init__(f"{self.status_name}: {message}")


def _check(lib, status, ctx=None):
    if status != F.EAR_OK:
        msg = lib.ear_ctx_last_error(ctx).decode() if ctx else ""
        raise EarError(status, msg or lib.ear_status_string(status).decode())


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class Context:
    """ear_ctx: one HIP device and one stream."""

    def __init__(self, device: int = 0):
        self.lib = F.load()
        h = C.c_void_p()
        _check(self.lib, self.lib.ear_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.ear_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_memory_cap(self, nbytes: int):
        _check(self.lib, self.lib.ear_ctx_set_memory_cap(self.h, nbytes), self.h)

    def set_tuning(self, block_extend=0, block_shade=0, block_shadow=0, ray_chunk=0, waves_per_cu=0, stack_lds=0,
                   traversal=0, graph_iters=0, refill=0, fetch_chunk=0, fetch_segments=0, tail_rays=0, splat=0):
        """Launch configuration (0 = library default); traversal: 1 static
        lanes, 2 closest-hit lanes refilled from the queue, 3 both kernels
        refilled (default); refill / fetch_chunk: idle lanes that trigger a
        refill (default 24), entries claimed per atomic (default: one per
        idle lane) and queue segments with a counter each (default 64);
        graph_iters > 0 replays the bounce loop as a captured HIP graph;
        tail_rays: the last <= tail_rays live rays go to one launch that
        traces each to its end (default: every ray when no scene of the call
        has more than 64 triangles, else the last 2^20; 1 turns it off);
        splat: 1 direct 128-bit atomics per ramp point, 2 records binned by
        time tile and accumulated in LDS (same histograms); 0 the tiled one
        when the call's (row, tile) keys fit in 8192 and its queues hold at
        least 2^22 rays, else the atomics."""
        t = F.ear_tuning(block_extend, block_shade, block_shadow, ray_chunk, waves_per_cu, stack_lds, traversal,
                         graph_iters, refill, fetch_chunk, fetch_segments, tail_rays, splat)
        _check(self.lib, self.lib.ear_ctx_set_tuning(self.h, C.byref(t)), self.h)

    def get_tuning(self) -> dict:
        t = F.ear_tuning()
        _check(self.lib, self.lib.ear_ctx_get_tuning(self.h, C.byref(t)), self.h)
        return {k: getattr(t, k) for k, _ in F.ear_tuning._fields_}


def make_materials(specs):
    """specs: [(refl[3], refr[3] | None, spec[3] | None)] as in .EAR MAT blocks."""
    arr = (F.ear_material * max(1, len(specs)))()
    for i, (rl, rr, sp) in enumerate(specs):
        arr[i].refl[:] = [float(x) for x in rl]
        flags = 0
        if rr is not None:
            arr[i].refr[:] = [float(x) for x in rr]
            flags |= F.EAR_MAT_HAS_REFR
        if sp is not None:
            arr[i].spec[:] = [float(x) for x in sp]
            flags |= F.EAR_MAT_HAS_SPEC
        arr[i].flags = flags
    return arr


class Scene:
    """ear_scene: the merged triangle soup on the device with its LBVH."""

    def __init__(self, ctx: Context, tris, mat_ids, materials, leaf_size: int = 0, width: int = 0, builder: int = 0):
        self.ctx = ctx
        lib = ctx.lib
        t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 3, 3)
        self.n_tris = len(t)
        self.n_mat = len(materials)
        v0 = np.ascontiguousarray(t[:, 0, :])
        v1 = np.ascontiguousarray(t[:, 1, :])
        v2 = np.ascontiguousarray(t[:, 2, :])
        m = np.ascontiguousarray(mat_ids, dtype=np.uint16).reshape(-1)
        mats = make_materials(materials)
        opts = F.ear_bvh_opts(leaf_size, width, builder)
        h = C.c_void_p()
        _check(lib, lib.ear_scene_upload(ctx.h, _ptr(v0), _ptr(v1), _ptr(v2), _ptr(m), self.n_tris, mats,
                                         len(materials), C.byref(opts), C.byref(h)), ctx.h)
        self.h = h

    def set_materials(self, materials):
        """Replaces the material table (same count) without rebuilding the BVH."""
        mats = make_materials(materials)
        _check(self.ctx.lib, self.ctx.lib.ear_scene_set_materials(self.h, mats, len(materials)), self.ctx.h)

    def close(self):
        if self.h:
            self.ctx.lib.ear_scene_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self) -> dict:
        i = F.ear_scene_info()
        _check(self.ctx.lib, self.ctx.lib.ear_scene_get_info(self.h, C.byref(i)), self.ctx.h)
        return dict(n_tris=i.n_tris, n_nodes=i.n_nodes, max_depth=i.max_depth, leaf_size=i.leaf_size,
                    bounds=list(i.bounds), build_ms=i.build_ms, width=i.width, node_bytes=i.node_bytes,
                    collapse_ms=i.collapse_ms, builder=i.builder)

    def query_closest(self, o, d):
        o = np.ascontiguousarray(o, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
        n = len(o)
        t = np.zeros(n, np.float32)
        idx = np.zeros(n, np.int32)
        _check(self.ctx.lib, self.ctx.lib.ear_query_closest(self.h, n, _ptr(o), _ptr(d), _ptr(t), _ptr(idx)),
               self.ctx.h)
        return t, idx

    def query_occluded(self, p, q):
        p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        n = len(p)
        occ = np.zeros(n, np.uint8)
        _check(self.ctx.lib, self.ctx.lib.ear_query_occluded(self.h, n, _ptr(p), _ptr(q), _ptr(occ)), self.ctx.h)
        return occ.astype(bool)


def make_recorders(recs):
    arr = (F.ear_recorder * max(1, len(recs)))()
    for i, r in enumerate(recs):
        arr[i].kind = r.get("kind", F.EAR_REC_MONO)
        arr[i].pos[:] = [float(x) for x in r["pos"]]
        arr[i].right_ear[:] = [float(x) for x in r.get("right_ear", (1, 0, 0))]
        arr[i].head_size = float(r.get("head_size", 0.0))
        arr[i].head_abs[:] = [float(x) for x in r.get("head_abs", (1, 1, 1))]
        arr[i].radius = float(r.get("radius", 0.0))
    return arr


def n_channels(recs):
    return sum(2 if r.get("kind", F.EAR_REC_MONO) == F.EAR_REC_STEREO else 1 for r in recs)


def raw_to_float(raw: np.ndarray) -> np.ndarray:
    """{lo, hi} int64 pairs (128-bit two's complement, 2^-EAR_FRAC_BITS units) -> float64."""
    lo = raw[..., 0].view(np.uint64).astype(np.float64)
    hi = raw[..., 1].astype(np.float64)
    return (hi * 18446744073709551616.0 + lo) * FRAC_SCALE


def raw_to_int(raw: np.ndarray):
    """Exact Python integers (object array) of the 128-bit raw sums."""
    lo = raw[..., 0].view(np.uint64)
    hi = raw[..., 1]
    out = np.empty(lo.shape, dtype=object)
    for idx in np.ndindex(lo.shape):
        out[idx] = (int(hi[idx]) << 64) + int(lo[idx])
    return out


@dataclass
class RenderResult:
    value: np.ndarray            # [3 band slots][channels][bins] float64
    abs_value: np.ndarray | None
    raw: np.ndarray | None       # [3][ch][bins][2] int64
    abs_raw: np.ndarray | None
    stats: dict
    bands: tuple
    trace_hits: np.ndarray | None = None  # [3][trace_rays][trace_segs]
    trace_nseg: np.ndarray | None = None
    wall_s: float = 0.0
    extra: dict = field(default_factory=dict)


class _Prepared:
    """ctypes descriptors of one render and the numpy arrays they point at."""

    def __init__(self, *, recorders, n_rays, bands=(0, 1, 2), air=(0.99, 0.98, 0.95), dry=1.0, src_pos=(0, 0, 0),
                 src_tris=None, gain=1.0, max_bounces=999, cutoff=1e-8, seed=0xEA2, instance=0, keyframe=-1,
                 sound=0, n_bins=88200, abs_hist=True, instrument=False, trace_rays=0, trace_segs=0, want_raw=True,
                 want_value=True, estimator=0):
        self.recs = make_recorders(recorders)
        nch = n_channels(recorders)
        d = F.ear_render_desc()
        d.band_mask = sum(1 << b for b in bands)
        d.air[:] = [float(x) for x in air]
        d.dry = dry
        self.src_arr = None
        if src_tris is not None:
            self.src_arr = np.ascontiguousarray(src_tris, np.float32).reshape(-1, 3, 3)
            d.source.kind = F.EAR_SRC_MESH
            d.source.mesh_v = self.src_arr.ctypes.data
            d.source.mesh_n_tris = len(self.src_arr)
        else:
            d.source.kind = F.EAR_SRC_POINT
        d.source.pos[:] = [float(x) for x in src_pos]
        d.source.gain = gain
        d.recorders = self.recs
        d.n_recorders = len(recorders)
        d.n_rays = n_rays
        d.max_bounces = max_bounces
        d.energy_cutoff = cutoff
        d.seed = seed
        d.instance = instance
        d.keyframe = keyframe
        d.sound = sound
        d.hist_bins = n_bins
        flags = 0
        if abs_hist:
            flags |= F.EAR_FLAG_ABS_HIST
        if instrument:
            flags |= F.EAR_FLAG_INSTRUMENT
        if trace_rays and trace_segs:
            flags |= F.EAR_FLAG_TRACE_LOG
        d.flags = flags
        d.trace_rays = trace_rays
        d.trace_segs = trace_segs
        d.estimator = estimator  # F.EAR_EST_NEE / F.EAR_EST_SPHERE (recorder "radius")
        self.desc = d
        self.bands = tuple(bands)
        self.value = np.zeros((3, nch, n_bins), np.float64) if want_value else None
        self.abs_value = np.zeros((3, nch, n_bins), np.float64) if (abs_hist and want_value) else None
        self.raw = np.zeros((3, nch, n_bins, 2), np.int64) if want_raw else None
        self.abs_raw = np.zeros((3, nch, n_bins, 2), np.int64) if (abs_hist and want_raw) else None
        self.th = self.tn = None
        if flags & F.EAR_FLAG_TRACE_LOG:
            self.th = np.full((3, trace_rays, trace_segs), -2, np.int32)
            self.tn = np.zeros((3, trace_rays), np.uint32)
        self.hist = F.ear_hist(_ptr(self.raw), _ptr(self.value), _ptr(self.abs_raw), _ptr(self.abs_value),
                               _ptr(self.th), _ptr(self.tn))

    def result(self, stats: dict, wall: float) -> "RenderResult":
        return RenderResult(self.value, self.abs_value, self.raw, self.abs_raw, stats, self.bands, self.th, self.tn,
                            wall)


def render(ctx: Context, scene: Scene, **kw) -> RenderResult:
    """One render (ear_render): the band contexts of one (sound, keyframe)."""
    import time

    p = _Prepared(**kw)
    st = F.ear_stats()
    t0 = time.perf_counter()
    _check(ctx.lib, ctx.lib.ear_render(ctx.h, scene.h, C.byref(p.desc), C.byref(p.hist), C.byref(st)), ctx.h)
    wall = time.perf_counter() - t0
    return p.result({k: getattr(st, k) for k, _ in F.ear_stats._fields_}, wall)


def render_noalloc(ctx: Context, scene: Scene, **kw) -> RenderResult:
    """One render without device allocation (ear_render_noalloc): the same
    arguments and results as render, every device buffer carved from the
    library's fixed per-GPU scratch region."""
    import time

    p = _Prepared(**kw)
    st = F.ear_stats()
    t0 = time.perf_counter()
    _check(ctx.lib, ctx.lib.ear_render_noalloc(ctx.h, scene.h, C.byref(p.desc), C.byref(p.hist), C.byref(st)),
           ctx.h)
    wall = time.perf_counter() - t0
    return p.result({k: getattr(st, k) for k, _ in F.ear_stats._fields_}, wall)


def render_batch(ctx: Context, items) -> tuple[list, dict, float]:
    """Several renders in one wavefront (ear_render_batch).  items: [(scene,
    kwargs of render)].  Returns (results, aggregated stats, wall seconds)."""
    import time

    preps = [_Prepared(**kw) for _, kw in items]
    arr = (F.ear_batch_item * len(items))()
    for i, ((scene, _), p) in enumerate(zip(items, preps)):
        arr[i].scene = scene.h
        arr[i].desc = C.pointer(p.desc)
        arr[i].out = C.pointer(p.hist)
    st = F.ear_stats()
    t0 = time.perf_counter()
    _check(ctx.lib, ctx.lib.ear_render_batch(ctx.h, arr, len(items), C.byref(st)), ctx.h)
    wall = time.perf_counter() - t0
    stats = {k: getattr(st, k) for k, _ in F.ear_stats._fields_}
    return [p.result({}, wall) for p in preps], stats, wall


def render_backward(ctx: Context, scene: Scene, *, adj_value=None, adj_abs=None, local=False, score=False, **kw):
    """Gradient of sum(adj_value * value) + sum(adj_abs * abs_value) of the
    render `kw` describes with respect to each material's reflection
    coefficient and specularity (ear_render_backward).  The adjoints have the
    shape of RenderResult.value, (3, n_channels, n_bins).  Returns
    (d/drefl, d/dspec), each (n_mat, 3); bands outside `bands` are zero.
    local: EAR_FLAG_ADJ_LOCAL (direction tangents end at their vertex);
    score: EAR_FLAG_ADJ_SCORE (the unbiased score form of d/dspec)."""
    kw = {k: v for k, v in kw.items() if k not in ("want_raw", "want_value")}  # no histograms come back
    p = _Prepared(want_raw=False, want_value=False, **kw)
    if local:
        p.desc.flags |= F.EAR_FLAG_ADJ_LOCAL
    if score:
        p.desc.flags |= F.EAR_FLAG_ADJ_SCORE
    nch = n_channels(kw["recorders"])
    shape = (3, nch, p.desc.hist_bins)
    a_v = None if adj_value is None else np.ascontiguousarray(np.broadcast_to(adj_value, shape), np.float64)
    a_a = None if adj_abs is None else np.ascontiguousarray(np.broadcast_to(adj_abs, shape), np.float64)
    g_r = np.zeros((scene.n_mat, 3), np.float64)
    g_s = np.zeros((scene.n_mat, 3), np.float64)
    _check(ctx.lib, ctx.lib.ear_render_backward(ctx.h, scene.h, C.byref(p.desc), _ptr(a_v), _ptr(a_a), _ptr(g_r),
                                                _ptr(g_s)), ctx.h)
    return g_r, g_s


def render_backward_batch(ctx: Context, items):
    """The adjoint of many renders at once (ear_render_backward_batch).
    items: sequence of (scene, kw) with kw as for render_backward (adj_value,
    adj_abs, local, score and the render keywords).  Returns [(d/drefl, d/dspec)]."""
    keep, arr = [], (F.ear_backward_item * max(1, len(items)))()
    outs = []
    for i, (scene, kw) in enumerate(items):
        kw = {k: v for k, v in kw.items() if k not in ("want_raw", "want_value")}
        adj_value, adj_abs, local = kw.pop("adj_value", None), kw.pop("adj_abs", None), kw.pop("local", False)
        score = kw.pop("score", False)
        p = _Prepared(want_raw=False, want_value=False, **kw)
        if local:
            p.desc.flags |= F.EAR_FLAG_ADJ_LOCAL
        if score:
            p.desc.flags |= F.EAR_FLAG_ADJ_SCORE
        shape = (3, n_channels(kw["recorders"]), p.desc.hist_bins)
        a_v = None if adj_value is None else np.ascontiguousarray(np.broadcast_to(adj_value, shape), np.float64)
        a_a = None if adj_abs is None else np.ascontiguousarray(np.broadcast_to(adj_abs, shape), np.float64)
        g_r = np.zeros((scene.n_mat, 3), np.float64)
        g_s = np.zeros((scene.n_mat, 3), np.float64)
        keep.append((p, a_v, a_a))
        outs.append((g_r, g_s))
        arr[i] = F.ear_backward_item(scene.h, C.pointer(p.desc), _ptr(a_v), _ptr(a_a), _ptr(g_r), _ptr(g_s))
    _check(ctx.lib, ctx.lib.ear_render_backward_batch(ctx.h, arr, len(items)), ctx.h)
    return outs


def convolve(ctx: Context, ir, signal, ir2=None) -> np.ndarray:
    """RecorderTrack::Process on the device (ear_convolve): the convolution of
    the track `ir` with `signal`, or with a second track `ir2` the keyframe
    cross-fade (signal faded out for ir, faded in for ir2).  Returns
    max(len(ir), len(ir2)) + len(signal) + 1 samples."""
    a = np.ascontiguousarray(ir, np.float32)
    s = np.ascontiguousarray(signal, np.float32)
    b = None if ir2 is None else np.ascontiguousarray(ir2, np.float32)
    n_out = max(len(a), 0 if b is None else len(b)) + len(s) + 1
    out = np.zeros(n_out, np.float32)
    _check(ctx.lib, ctx.lib.ear_convolve(ctx.h, _ptr(a), len(a), _ptr(b), 0 if b is None else len(b), _ptr(s), len(s),
                                         _ptr(out)), ctx.h)
    return out
