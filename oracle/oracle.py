"""ctypes front end of the CPU oracle (libear_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, as the checker.  The product never imports it.

The C library restates the reference's Scene::Render (src/Scene.cpp:111-318)
and everything it calls; see oracle/ear_oracle.c for the per-function
citations.  This module only marshals numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libear_oracle.so")

REC_MONO, REC_STEREO = 0, 1
SRC_POINT, SRC_MESH = 0, 1
MAT_HAS_REFR, MAT_HAS_SPEC = 1, 2


class Material(C.Structure):
    _fields_ = [("refl", C.c_float * 3), ("refr", C.c_float * 3), ("spec", C.c_float * 3), ("flags", C.c_uint32)]


class Recorder(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pos", C.c_float * 3), ("right_ear", C.c_float * 3),
                ("head_size", C.c_float), ("head_abs", C.c_float * 3), ("radius", C.c_float)]


class Scene(C.Structure):
    _fields_ = [("n_tris", C.c_uint32), ("v", C.c_void_p), ("mat", C.c_void_p),
                ("n_mat", C.c_uint32), ("mats", C.POINTER(Material))]


class Desc(C.Structure):
    _fields_ = [("band", C.c_int32), ("air", C.c_float), ("dry", C.c_float), ("gain", C.c_float),
                ("src_kind", C.c_int32), ("src_pos", C.c_float * 3), ("src_n_tris", C.c_uint32),
                ("src_v", C.c_void_p), ("n_rec", C.c_uint32), ("rec", C.POINTER(Recorder)),
                ("ray_begin", C.c_uint32), ("ray_end", C.c_uint32), ("n_rays", C.c_uint32),
                ("max_bounces", C.c_uint32), ("cutoff", C.c_double), ("seed", C.c_uint64),
                ("ctx", C.c_uint32), ("instance", C.c_uint32), ("n_bins", C.c_uint32),
                ("n_threads", C.c_uint32), ("trace_rays", C.c_uint32), ("trace_segs", C.c_uint32),
                ("estimator", C.c_uint32)]


class Out(C.Structure):
    _fields_ = [("hist_signed", C.c_void_p), ("hist_abs", C.c_void_p), ("hist_f32", C.c_void_p),
                ("trace_hits", C.c_void_p), ("trace_nseg", C.c_void_p), ("trace_ncon", C.c_void_p),
                ("segments", C.c_uint64), ("shadow_rays", C.c_uint64), ("contributions", C.c_uint64),
                ("direct_visible", C.c_void_p), ("direct_s", C.c_void_p), ("direct_w", C.c_void_p),
                ("direct_a", C.c_void_p), ("tally", C.c_void_p), ("tally_abs", C.c_void_p),
                ("path_length", C.c_double), ("trace_rays_od", C.c_void_p)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.or_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.or_uniform.restype = C.c_float
        L.or_pow.argtypes = [C.c_float, C.c_float]
        L.or_pow.restype = C.c_float
        L.or_pow_d.argtypes = [C.c_double, C.c_double]
        L.or_pow_d.restype = C.c_double
        fp = C.POINTER(C.c_float)
        L.or_mt_intersect.argtypes = [fp, fp, fp, fp, fp, fp]
        L.or_closest_hit.argtypes = [C.POINTER(Scene), fp, fp, fp, C.POINTER(C.c_int32)]
        L.or_any_hit.argtypes = [C.POINTER(Scene), fp, fp]
        L.or_tri_normal.argtypes = [fp, fp]
        L.or_material_retention.argtypes = [C.POINTER(Material), fp]
        L.or_sample_dirs.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, fp, fp,
                                     C.POINTER(C.c_uint32)]
        L.or_record_f32.argtypes = [fp, C.c_uint32, C.c_float, C.c_float, C.c_float]
        L.or_render.argtypes = [C.POINTER(Scene), C.POINTER(Desc), C.POINTER(Out)]
        L.or_num_channels.argtypes = [C.POINTER(Desc)]
        L.or_gradient.argtypes = [C.POINTER(Scene), C.POINTER(Desc), C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                  C.c_void_p]
        L.or_gradient_k.argtypes = [C.POINTER(Scene), C.POINTER(Desc), C.c_void_p, C.c_void_p, C.c_int, C.c_float,
                                    C.c_void_p, C.c_void_p]
        L.or_now.restype = C.c_double
        L.or_score_dlogp.restype = C.c_double
        L.or_score_dlogp.argtypes = [C.c_double, C.c_double]
        L.or_score_flux.restype = C.c_double
        L.or_score_flux.argtypes = [C.c_double, C.c_double]
        L.or_score_rim_dirs.argtypes = [C.c_uint64, C.c_uint32, fp, fp, C.c_float, fp]
        _lib = L
    return _lib


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def uniform(seed, ray, ctx, instance, draw):
    return lib().or_uniform(seed, ray, ctx, instance, draw)


def powf(x, y):
    return lib().or_pow(x, y)


def pow_d(x, y):
    return lib().or_pow_d(x, y)


def mt_intersect(v0, e1, e2, o, d):
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (v0, e1, e2, o, d)]
    t = np.zeros(1, np.float32)
    hit = lib().or_mt_intersect(*[_fptr(a) for a in arrs], _fptr(t))
    return bool(hit), float(t[0])


def tri_normal(v9):
    v = np.ascontiguousarray(v9, dtype=np.float32).reshape(9)
    n = np.zeros(3, np.float32)
    lib().or_tri_normal(_fptr(v), _fptr(n))
    return n


def material_retention(refl, refr=None, spec=None):
    m = make_materials([(refl, refr, spec)])[0]
    r = np.zeros(3, np.float32)
    ok = lib().or_material_retention(C.byref(m), _fptr(r))
    return (r if ok else None)


def sample_dirs(seed, n, mode=0, normal=(0, 0, 1), ctx=0, instance=0):
    out = np.zeros((n, 3), np.float32)
    nd = np.zeros(n, np.uint32)
    nrm = np.asarray(normal, np.float32)
    lib().or_sample_dirs(seed, ctx, instance, n, mode, _fptr(nrm), _fptr(out),
                         nd.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out, nd


def record_f32(n_bins, a, t, dist, buf=None):
    if buf is None:
        buf = np.zeros(n_bins, np.float32)
    lib().or_record_f32(_fptr(buf), n_bins, a, t, dist)
    return buf


def make_materials(specs):
    """specs: list of (refl[3], refr[3]|None, spec[3]|None) as in a .EAR MAT block."""
    arr = (Material * max(1, len(specs)))()
    for i, (rl, rr, sp) in enumerate(specs):
        arr[i].refl[:] = [float(x) for x in rl]
        flags = 0
        if rr is not None:
            arr[i].refr[:] = [float(x) for x in rr]
            flags |= MAT_HAS_REFR
        if sp is not None:
            arr[i].spec[:] = [float(x) for x in sp]
            flags |= MAT_HAS_SPEC
        arr[i].flags = flags
    return arr


class OracleScene:
    """Holds numpy buffers alive for an or_scene."""

    def __init__(self, tris, mat_ids, materials):
        self.tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 3, 3)
        self.mat = np.ascontiguousarray(mat_ids, dtype=np.uint16).reshape(-1)
        self.mats = make_materials(materials)
        self.n_mat = len(materials)
        self.s = Scene(len(self.tris), self.tris.ctypes.data if len(self.tris) else None,
                       self.mat.ctypes.data if len(self.mat) else None, self.n_mat, self.mats)

    def closest_hit(self, o, d):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        t = np.zeros(1, np.float32)
        idx = C.c_int32(-1)
        hit = lib().or_closest_hit(C.byref(self.s), _fptr(o), _fptr(d), _fptr(t), C.byref(idx))
        return (float(t[0]), idx.value) if hit == 1 else (None, -1)

    def any_hit(self, p, q):
        p = np.ascontiguousarray(p, np.float32)
        q = np.ascontiguousarray(q, np.float32)
        return bool(lib().or_any_hit(C.byref(self.s), _fptr(p), _fptr(q)))


def make_recorders(recs):
    arr = (Recorder * max(1, len(recs)))()
    for i, r in enumerate(recs):
        arr[i].kind = r.get("kind", REC_MONO)
        arr[i].pos[:] = [float(x) for x in r["pos"]]
        arr[i].right_ear[:] = [float(x) for x in r.get("right_ear", (1, 0, 0))]
        arr[i].head_size = float(r.get("head_size", 0.0))
        arr[i].head_abs[:] = [float(x) for x in r.get("head_abs", (1, 1, 1))]
        arr[i].radius = float(r.get("radius", 0.0))
    return arr


def n_channels(recs):
    return sum(2 if r.get("kind", REC_MONO) == REC_STEREO else 1 for r in recs)


def render(scene: OracleScene, *, band, air, recorders, n_rays, src_pos=(0, 0, 0), src_tris=None,
           dry=1.0, gain=1.0, ray_begin=0, ray_end=None, max_bounces=999, cutoff=1e-8, seed=0xEA2,
           ctx=None, instance=0, n_bins=88200, n_threads=1, trace_rays=0, trace_segs=0,
           want_f32=False, want_abs=True, estimator=0):
    """One reference render context (sound, keyframe, band); returns a dict of
    raw sums.  estimator 1: receiver spheres (recorder "radius")."""
    if ray_end is None:
        ray_end = n_rays
    if ctx is None:
        ctx = band
    recs = make_recorders(recorders)
    nch = n_channels(recorders)
    d = Desc()
    d.band, d.air, d.dry, d.gain = band, air, dry, gain
    if src_tris is not None:
        src = np.ascontiguousarray(src_tris, dtype=np.float32).reshape(-1, 3, 3)
        d.src_kind, d.src_n_tris, d.src_v = SRC_MESH, len(src), src.ctypes.data
    else:
        src = None
        d.src_kind = SRC_POINT
    d.src_pos[:] = [float(x) for x in src_pos]
    d.n_rec, d.rec = len(recorders), recs
    d.ray_begin, d.ray_end, d.n_rays = ray_begin, ray_end, n_rays
    d.max_bounces, d.cutoff, d.seed = max_bounces, cutoff, seed
    d.ctx, d.instance, d.n_bins, d.n_threads = ctx, instance, n_bins, n_threads
    d.trace_rays, d.trace_segs = trace_rays, trace_segs
    d.estimator = estimator
    hs = np.zeros((nch, n_bins), np.float64)
    ha = np.zeros((nch, n_bins), np.float64) if want_abs else None
    hf = np.zeros((nch, n_bins), np.float32) if want_f32 else None
    th = np.zeros((max(trace_rays, 1), max(trace_segs, 1)), np.int32)
    tn = np.zeros(max(trace_rays, 1), np.uint32)
    tc = np.zeros(max(trace_rays, 1), np.uint32)
    dv = np.zeros(max(nch, 1), np.int32)
    ds = np.zeros(max(nch, 1), np.int32)
    dw = np.zeros(max(nch, 1), np.int32)
    da = np.zeros(max(nch, 1), np.float32)
    o = Out()
    o.hist_signed = hs.ctypes.data
    o.hist_abs = ha.ctypes.data if ha is not None else None
    o.hist_f32 = hf.ctypes.data if hf is not None else None
    tod = np.zeros((max(trace_rays, 1), max(trace_segs, 1), 6), np.float32)
    if trace_rays:
        o.trace_hits, o.trace_nseg, o.trace_ncon = th.ctypes.data, tn.ctypes.data, tc.ctypes.data
        o.trace_rays_od = tod.ctypes.data
    o.direct_visible, o.direct_s, o.direct_w, o.direct_a = dv.ctypes.data, ds.ctypes.data, dw.ctypes.data, da.ctypes.data
    tl = np.zeros(max(len(recorders), 1), np.float64)
    tla = np.zeros(max(len(recorders), 1), np.float64)
    o.tally, o.tally_abs = tl.ctypes.data, tla.ctypes.data
    t0 = lib().or_now()
    rc = lib().or_render(C.byref(scene.s), C.byref(d), C.byref(o))
    t1 = lib().or_now()
    if rc != 0:
        raise RuntimeError(f"or_render failed: {rc}")
    res = dict(hist_signed=hs, hist_abs=ha, hist_f32=hf, segments=o.segments, shadow_rays=o.shadow_rays,
               contributions=o.contributions, seconds=t1 - t0, path_length=o.path_length,
               tally=tl[:len(recorders)].copy(), tally_abs=tla[:len(recorders)].copy(),
               direct=dict(visible=dv[:nch].copy(), s=ds[:nch].copy(), w=dw[:nch].copy(), a=da[:nch].copy()))
    if trace_rays:
        res.update(trace_hits=th, trace_nseg=tn, trace_ncon=tc, trace_od=tod)
    del src
    return res


def gradient(scene: OracleScene, *, band, air, recorders, n_rays, adj_signed=None, adj_abs=None,
             src_pos=(0, 0, 0), src_tris=None, max_bounces=50, cutoff=0.0, seed=0xEA2, ctx=None, instance=0,
             n_bins=44100, n_threads=8, local=False, score=False, q_spawn=None):
    """FP64 reference gradient of sum(adj * H) (raw sums) with respect to the
    band's per-material reflection coefficient and specularity (or_gradient;
    local: direction tangents end at their vertex; score: the score form of
    the specularity gradient, a continuation path spawned at a vertex with
    probability min(1, q_spawn F), F its rim flux; None: the library's
    default)."""
    if ctx is None:
        ctx = band
    recs = make_recorders(recorders)
    d = Desc()
    d.band, d.air, d.dry, d.gain = band, air, 1.0, 1.0
    src = None
    if src_tris is not None:
        src = np.ascontiguousarray(src_tris, dtype=np.float32).reshape(-1, 3, 3)
        d.src_kind, d.src_n_tris, d.src_v = SRC_MESH, len(src), src.ctypes.data
    else:
        d.src_kind = SRC_POINT
    d.src_pos[:] = [float(x) for x in src_pos]
    d.n_rec, d.rec = len(recorders), recs
    d.ray_begin, d.ray_end, d.n_rays = 0, n_rays, n_rays
    d.max_bounces, d.cutoff, d.seed = max_bounces, cutoff, seed
    d.ctx, d.instance, d.n_bins, d.n_threads = ctx, instance, n_bins, n_threads
    a_s = None if adj_signed is None else np.ascontiguousarray(adj_signed, np.float64)
    a_a = None if adj_abs is None else np.ascontiguousarray(adj_abs, np.float64)
    gr = np.zeros(max(scene.n_mat, 1), np.float64)
    gs = np.zeros(max(scene.n_mat, 1), np.float64)
    mode = 2 if score else int(bool(local))
    if q_spawn is None:
        rc = lib().or_gradient(C.byref(scene.s), C.byref(d), None if a_s is None else a_s.ctypes.data,
                               None if a_a is None else a_a.ctypes.data, mode, gr.ctypes.data, gs.ctypes.data)
    else:
        rc = lib().or_gradient_k(C.byref(scene.s), C.byref(d), None if a_s is None else a_s.ctypes.data,
                                 None if a_a is None else a_a.ctypes.data, mode, float(q_spawn), gr.ctypes.data,
                                 gs.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"or_gradient failed: {rc}")
    del src
    return gr[:scene.n_mat].copy(), gs[:scene.n_mat].copy()


def direct_ramp(direct, n_bins):
    """Dense FP64 histogram of the direct-sound ramps (one per channel)."""
    nch = len(direct["visible"])
    h = np.zeros((nch, n_bins), np.float64)
    for c in range(nch):
        if not direct["visible"][c]:
            continue
        s, w, a = int(direct["s"][c]), int(direct["w"][c]), float(direct["a"][c])
        S = a / w
        for i in range(w):
            b = s + i
            if 0 <= b < n_bins:
                h[c, b] += (w - i) * S
    return h


def score_dlogp(c, s):
    """d log p(v) / ds of the specular blend's direction density at c = v.r."""
    return lib().or_score_dlogp(float(c), float(s))


def score_flux(cr, s):
    """The rate F at which the blend's direction probability leaves through
    the image of the hemisphere's rim, cr = r.nu."""
    return lib().or_score_flux(float(cr), float(s))


def score_rim_dirs(n, nu, r, s, seed=0xEA2):
    """n rim directions normalize((1-s) e(phi) + s r), phi drawn with the
    flux density, as a continuation draws them."""
    out = np.zeros((n, 3), np.float32)
    nu = np.ascontiguousarray(nu, np.float32)
    r = np.ascontiguousarray(r, np.float32)
    lib().or_score_rim_dirs(seed, n, _fptr(nu), _fptr(r), float(s), _fptr(out))
    return out
