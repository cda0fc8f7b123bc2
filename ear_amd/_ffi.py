"""ctypes mirror of include/ear.h (the C ABI of libear_hip.so).

Loading never falls back to anything: if the library is missing the import of
the high-level API raises, so no GPU test can pass on a non-native path.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libear_hip.so")

EAR_OK = 0
STATUS_NAMES = {
    0: "EAR_OK", 1: "EAR_ERR_INVALID_ARGUMENT", 2: "EAR_ERR_INVALID_MATERIAL", 3: "EAR_ERR_INVALID_GEOMETRY",
    4: "EAR_ERR_NO_DEVICE", 5: "EAR_ERR_OUT_OF_MEMORY", 6: "EAR_ERR_HIP", 7: "EAR_ERR_CAPACITY",
    8: "EAR_ERR_INTERNAL",
}
EAR_MAT_HAS_REFR, EAR_MAT_HAS_SPEC = 1, 2
EAR_REC_MONO, EAR_REC_STEREO = 0, 1
EAR_SRC_POINT, EAR_SRC_MESH = 0, 1
EAR_FLAG_ABS_HIST, EAR_FLAG_INSTRUMENT, EAR_FLAG_TRACE_LOG, EAR_FLAG_ADJ_LOCAL, EAR_FLAG_ADJ_SCORE = 1, 2, 4, 8, 16
EAR_FRAC_BITS = 80
EAR_EST_NEE, EAR_EST_SPHERE = 0, 1

# every symbol declared in include/ear.h
EXPORTS = [
    "ear_ctx_create", "ear_ctx_destroy", "ear_ctx_set_memory_cap", "ear_status_string", "ear_ctx_last_error",
    "ear_abi_version", "ear_scene_upload", "ear_scene_destroy", "ear_scene_get_info", "ear_query_closest",
    "ear_query_occluded", "ear_render", "ear_render_batch", "ear_ctx_set_tuning", "ear_ctx_get_tuning",
    "ear_render_backward", "ear_convolve", "ear_render_backward_batch", "ear_scene_set_materials",
    "ear_render_noalloc",
]


class ear_material(C.Structure):
    _fields_ = [("refl", C.c_float * 3), ("refr", C.c_float * 3), ("spec", C.c_float * 3), ("flags", C.c_uint32)]


class ear_bvh_opts(C.Structure):
    _fields_ = [("leaf_size", C.c_uint32), ("width", C.c_uint32), ("builder", C.c_uint32), ("reserved", C.c_uint32)]


EAR_BVH_DEFAULT, EAR_BVH_LBVH, EAR_BVH_SAH = 0, 1, 2


class ear_scene_info(C.Structure):
    _fields_ = [("n_tris", C.c_uint32), ("n_nodes", C.c_uint32), ("max_depth", C.c_uint32),
                ("leaf_size", C.c_uint32), ("bounds", C.c_float * 6), ("build_ms", C.c_double),
                ("width", C.c_uint32), ("node_bytes", C.c_uint32), ("collapse_ms", C.c_double),
                ("builder", C.c_uint32), ("reserved", C.c_uint32)]


class ear_recorder(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pos", C.c_float * 3), ("right_ear", C.c_float * 3),
                ("head_size", C.c_float), ("head_abs", C.c_float * 3), ("radius", C.c_float)]


class ear_source(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pos", C.c_float * 3), ("mesh_v", C.c_void_p),
                ("mesh_n_tris", C.c_uint32), ("gain", C.c_float)]


class ear_render_desc(C.Structure):
    _fields_ = [("band_mask", C.c_uint32), ("air", C.c_float * 3), ("dry", C.c_float), ("source", ear_source),
                ("recorders", C.POINTER(ear_recorder)), ("n_recorders", C.c_uint32), ("n_rays", C.c_uint32),
                ("max_bounces", C.c_uint32), ("energy_cutoff", C.c_double), ("seed", C.c_uint64),
                ("instance", C.c_uint32), ("keyframe", C.c_int32), ("sound", C.c_uint32),
                ("hist_bins", C.c_uint32), ("flags", C.c_uint32), ("trace_rays", C.c_uint32),
                ("trace_segs", C.c_uint32), ("estimator", C.c_uint32)]


class ear_hist(C.Structure):
    _fields_ = [("raw", C.c_void_p), ("value", C.c_void_p), ("abs_raw", C.c_void_p), ("abs_value", C.c_void_p),
                ("trace_hits", C.c_void_p), ("trace_nseg", C.c_void_p)]


class ear_stats(C.Structure):
    _fields_ = [("segments", C.c_uint64), ("shadow_rays", C.c_uint64), ("contributions", C.c_uint64),
                ("rays", C.c_uint64), ("node_visits_closest", C.c_uint64), ("tri_tests_closest", C.c_uint64),
                ("node_visits_shadow", C.c_uint64), ("tri_tests_shadow", C.c_uint64),
                ("bounces_launched", C.c_uint32), ("launches", C.c_uint32), ("trace_ms", C.c_double),
                ("extend_ms", C.c_double), ("shade_ms", C.c_double), ("shadow_ms", C.c_double),
                ("extend_launches", C.c_uint32), ("shadow_launches", C.c_uint32), ("extend_queries", C.c_uint64),
                ("shadow_queries", C.c_uint64), ("dropped_ramps", C.c_uint64), ("bins_needed", C.c_uint32),
                ("reserved", C.c_uint32), ("shadow_contributions", C.c_uint64), ("shadow_ramps", C.c_uint64),
                ("splat_ms", C.c_double)]


class ear_batch_item(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("desc", C.POINTER(ear_render_desc)), ("out", C.POINTER(ear_hist))]


class ear_backward_item(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("desc", C.POINTER(ear_render_desc)), ("adj_value", C.c_void_p),
                ("adj_abs", C.c_void_p), ("grad_refl", C.c_void_p), ("grad_spec", C.c_void_p)]


class ear_tuning(C.Structure):
    _fields_ = [("block_extend", C.c_uint32), ("block_shade", C.c_uint32), ("block_shadow", C.c_uint32),
                ("ray_chunk", C.c_uint32), ("waves_per_cu", C.c_uint32), ("stack_lds", C.c_uint32),
                ("traversal", C.c_uint32), ("graph_iters", C.c_uint32), ("refill", C.c_uint32),
                ("fetch_chunk", C.c_uint32), ("fetch_segments", C.c_uint32), ("tail_rays", C.c_uint32),
                ("splat", C.c_uint32)]


EAR_TRAVERSAL_DEFAULT, EAR_TRAVERSAL_STATIC, EAR_TRAVERSAL_FETCH, EAR_TRAVERSAL_FETCH_ALL = 0, 1, 2, 3
EAR_SPLAT_DEFAULT, EAR_SPLAT_ATOMIC, EAR_SPLAT_TILED = 0, 1, 2


_lib = None


def load(path: str | None = None):
    """Loads libear_hip.so (LIB_PATH unless a path is given) and declares the
    prototypes; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    L = C.CDLL(path)
    vp = C.c_void_p
    fp = C.c_void_p
    L.ear_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.ear_ctx_destroy.argtypes = [vp]
    L.ear_ctx_set_memory_cap.argtypes = [vp, C.c_uint64]
    L.ear_status_string.argtypes = [C.c_int]
    L.ear_status_string.restype = C.c_char_p
    L.ear_ctx_last_error.argtypes = [vp]
    L.ear_ctx_last_error.restype = C.c_char_p
    L.ear_abi_version.restype = C.c_int
    L.ear_scene_upload.argtypes = [vp, fp, fp, fp, fp, C.c_uint32, C.POINTER(ear_material), C.c_uint32,
                                   C.POINTER(ear_bvh_opts), C.POINTER(vp)]
    L.ear_scene_destroy.argtypes = [vp]
    L.ear_scene_get_info.argtypes = [vp, C.POINTER(ear_scene_info)]
    L.ear_query_closest.argtypes = [vp, C.c_uint32, fp, fp, fp, fp]
    L.ear_query_occluded.argtypes = [vp, C.c_uint32, fp, fp, fp]
    L.ear_render.argtypes = [vp, vp, C.POINTER(ear_render_desc), C.POINTER(ear_hist), C.POINTER(ear_stats)]
    L.ear_render_noalloc.argtypes = [vp, vp, C.POINTER(ear_render_desc), C.POINTER(ear_hist), C.POINTER(ear_stats)]
    L.ear_render_batch.argtypes = [vp, C.POINTER(ear_batch_item), C.c_uint32, C.POINTER(ear_stats)]
    L.ear_render_backward.argtypes = [vp, vp, C.POINTER(ear_render_desc), vp, vp, vp, vp]
    L.ear_convolve.argtypes = [vp, fp, C.c_uint32, fp, C.c_uint32, fp, C.c_uint32, fp]
    L.ear_render_backward_batch.argtypes = [vp, C.POINTER(ear_backward_item), C.c_uint32]
    L.ear_scene_set_materials.argtypes = [vp, C.POINTER(ear_material), C.c_uint32]
    L.ear_ctx_set_tuning.argtypes = [vp, C.POINTER(ear_tuning)]
    L.ear_ctx_get_tuning.argtypes = [vp, C.POINTER(ear_tuning)]
    for name in EXPORTS:
        f = getattr(L, name)
        if f.restype is C.c_int or name not in ("ear_status_string", "ear_ctx_last_error", "ear_abi_version"):
            if name not in ("ear_status_string", "ear_ctx_last_error"):
                f.restype = C.c_int
    _lib = L
    return L
