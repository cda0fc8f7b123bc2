/*
 * ear.h -- C ABI of the MI355X acoustic ray-tracing core (libear_hip.so).
 *
 * This is the drop-in boundary beneath the reference's
 *     void Scene::Render(int band, int sound, float absorbtion_factor,
 *                        int num_samples, float dry,
 *                        const std::vector<Recorder*>& rec, int keyframeID)
 * (src/Scene.h:77, defined at src/Scene.cpp:111-318).  The reference has no
 * plugin API: its caller is SceneContext::operator() (src/SceneContext.h:46-48),
 * one call per (sound, keyframe, band) context on its own thread
 * (src/EAR.cpp:170-207).  The host C++ Scene::Render of this project keeps
 * that signature and calls the functions below; every entry point takes plain
 * pointers and sizes, returns an ear_status and never throws.
 *
 * Memory: the caller owns every host buffer; the library owns device memory
 * behind the opaque handles.  Threading: one ear_ctx per host thread (each
 * has its own HIP stream); handles are not shared between threads.
 */
#ifndef EAR_H
#define EAR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EAR_ABI_VERSION 7

typedef enum {
    EAR_OK = 0,
    EAR_ERR_INVALID_ARGUMENT = 1, /* null pointer, size 0 where >0 needed, bad enum */
    EAR_ERR_INVALID_MATERIAL = 2, /* refl + refr > 1: the reference throws DatatypeException
                                     ("Invalid material settings", src/Material.cpp:56) */
    EAR_ERR_INVALID_GEOMETRY = 3, /* NaN / Inf vertex, material id out of range */
    EAR_ERR_NO_DEVICE = 4,        /* no HIP device / bad ordinal */
    EAR_ERR_OUT_OF_MEMORY = 5,    /* device allocation failed or exceeds ear_ctx memory cap */
    EAR_ERR_HIP = 6,              /* any other HIP runtime error */
    EAR_ERR_CAPACITY = 7,         /* a size exceeds a documented limit */
    EAR_ERR_INTERNAL = 8
} ear_status;

typedef struct ear_ctx ear_ctx;
typedef struct ear_scene ear_scene;

/* ------------------------------------------------------------------ */
/* context                                                              */
/* ------------------------------------------------------------------ */

/* Replaces the boost::thread per SceneContext (src/EAR.cpp:196-207): one
 * device + one HIP stream. */
ear_status ear_ctx_create(int device, ear_ctx** out);
ear_status ear_ctx_destroy(ear_ctx* ctx);
/* Cap on device bytes the context may hold (0 = unlimited); allocations past
 * it fail with EAR_ERR_OUT_OF_MEMORY (lets host tests force that path). */
ear_status ear_ctx_set_memory_cap(ear_ctx* ctx, uint64_t bytes);
const char* ear_status_string(ear_status s);
/* Last error message recorded on this context (never NULL). */
const char* ear_ctx_last_error(const ear_ctx* ctx);
int ear_abi_version(void);

/* ------------------------------------------------------------------ */
/* scene: Mesh / Triangle / Material                                    */
/* ------------------------------------------------------------------ */

#define EAR_MAT_HAS_REFR 1u /* .EAR MAT block carried the optional refraction triple */
#define EAR_MAT_HAS_SPEC 2u /* .EAR MAT block carried the optional specularity triple */

/* One .EAR MAT block (src/Material.cpp:26-72): reflection, optional
 * refraction, optional specularity per band.  Absent triples read as 0, as
 * the reference's std::map<int,float>::operator[] default does. */
typedef struct {
    float refl[3];
    float refr[3];
    float spec[3];
    uint32_t flags;
} ear_material;

typedef struct {
    uint32_t leaf_size;  /* max triangles per BVH leaf, 1..16 (0 = library default, 2) */
    uint32_t width;      /* children per node: 2 or 4 (0 = library default, 4) */
    uint32_t builder;    /* EAR_BVH_* (0 = library default, EAR_BVH_SAH_GPU when 4-wide) */
    uint32_t reserved;
} ear_bvh_opts;
#define EAR_BVH_DEFAULT 0u /* EAR_BVH_SAH_GPU for 4-wide trees, EAR_BVH_SAH for 2-wide ones */
#define EAR_BVH_LBVH 1u  /* device linear BVH: Morton order, Karras hierarchy */
#define EAR_BVH_SAH 2u   /* host binned surface-area-heuristic build, then uploaded */
#define EAR_BVH_PLOC 3u  /* device build: locally-ordered clustering of the Morton-sorted
                            triangles (Meister and Bittner 2018) and a device collapse into
                            the quantised 4-wide nodes (4-wide only) */
#define EAR_BVH_SAH_GPU 4u /* device build: EAR_BVH_SAH's binned-SAH splits made level by
                              level on the device (one workgroup per node), then the same
                              device collapse (4-wide only) */

/* Uploads the merged triangle soup of every MESH block (the reference merges
 * all meshes into meshes[0], src/Scene.cpp:103-106, src/Mesh.cpp:116-123)
 * and builds its BVH (by default the device binned-SAH build; or the device
 * PLOC, the host SAH or the device LBVH build).  Triangles are given as three vertex
 * arrays of n_tris xyz triples each, in the reference's triangle order; the
 * triangle index reported by queries is the position in that order.
 * n_tris may be 0 (the reference's Mesh::Empty(), src/EAR.cpp:160-163). */
ear_status ear_scene_upload(ear_ctx* ctx, const float* v0, const float* v1, const float* v2,
                            const uint16_t* mat_id, uint32_t n_tris, const ear_material* mats,
                            uint32_t n_mat, const ear_bvh_opts* opts, ear_scene** out);
ear_status ear_scene_destroy(ear_scene* scene);
/* Replaces the scene's material table (same count as at upload) without
 * rebuilding the BVH; no reference counterpart (its materials are fixed once
 * parsed): fitting changes them every step. */
ear_status ear_scene_set_materials(ear_scene* scene, const ear_material* mats, uint32_t n_mat);

typedef struct {
    uint32_t n_tris;
    uint32_t n_nodes;     /* internal BVH nodes */
    uint32_t max_depth;   /* longest root-to-leaf path */
    uint32_t leaf_size;
    float bounds[6];      /* scene AABB lo xyz, hi xyz */
    double build_ms;      /* build time: device time of the LBVH, host time of the SAH build */
    uint32_t width;       /* children per node */
    uint32_t node_bytes;  /* bytes fetched per node visit */
    double collapse_ms;   /* host time of the 2-wide -> 4-wide collapse */
    uint32_t builder;     /* EAR_BVH_LBVH or EAR_BVH_SAH */
    uint32_t reserved;
} ear_scene_info;
ear_status ear_scene_get_info(const ear_scene* scene, ear_scene_info* info);

/* Batched geometric queries (Mesh::RayIntersection, src/Mesh.cpp:33-56, and
 * Mesh::LineIntersection, src/Mesh.cpp:58-71), host arrays in and out.
 * closest: o[n*3], d[n*3] -> t[n] (hit distance) and tri[n] (original index,
 *          -1 on a miss); accepted t in (0.001, 1e6), ties to the lower index.
 * occluded: segment p -> q, any triangle with t in (1e-5, 1) -> 1, else 0. */
ear_status ear_query_closest(ear_scene* scene, uint32_t n, const float* o, const float* d, float* t,
                             int32_t* tri);
ear_status ear_query_occluded(ear_scene* scene, uint32_t n, const float* p, const float* q,
                              uint8_t* occluded);

/* ------------------------------------------------------------------ */
/* render: Scene::Render                                                */
/* ------------------------------------------------------------------ */

#define EAR_REC_MONO 0   /* MonoRecorder (src/MonoRecorder.cpp), 1 channel */
#define EAR_REC_STEREO 1 /* StereoRecorder (src/StereoRecorder.cpp), channels left, right */

typedef struct {
    int32_t kind;
    float pos[3];        /* Recorder::getLocation(keyframe) */
    float right_ear[3];  /* StereoRecorder::getRightEar(keyframe) */
    float head_size;
    float head_abs[3];   /* as stored by the StereoRecorder ctor: max(0,(1-h)^4) */
    float radius;        /* EAR_EST_SPHERE: receiver sphere radius (m); unused otherwise */
} ear_recorder;

#define EAR_SRC_POINT 0 /* AbstractSoundFile with a location */
#define EAR_SRC_MESH 1  /* AbstractSoundFile with an emitting mesh */

typedef struct {
    int32_t kind;
    float pos[3];          /* getLocation(keyframe) for point sources */
    const float* mesh_v;   /* mesh sources: [n][3][3] vertices (host) */
    uint32_t mesh_n_tris;
    float gain;            /* AbstractSoundFile::getGain(), applied squared */
} ear_source;

#define EAR_FLAG_ABS_HIST 1u      /* also accumulate the |c| histogram (EDC / T60) */
#define EAR_FLAG_INSTRUMENT 2u    /* count BVH node visits / triangle tests (slower) */
#define EAR_FLAG_TRACE_LOG 4u     /* log per-ray hit triangle indices */
#define EAR_FLAG_ADJ_LOCAL 8u     /* ear_render_backward: keep only each vertex's own
                                     dependence on spec (see there) */
#define EAR_FLAG_ADJ_SCORE 16u    /* ear_render_backward: the score form of d/dspec,
                                     unbiased at any path length (see there) */

/* Receiver estimators.  EAR_EST_NEE is the reference's: a connection from
 * every hit point to every recorder, the shadow ray, the lobe-weighted
 * contribution and the analytic direct sound (src/Scene.cpp:185-311).
 * EAR_EST_SPHERE has no reference counterpart: each recorder is a sphere of
 * `radius` (a stereo recorder two, centred head_size / 2 either side of pos
 * along right_ear, one per channel, without the ITD/ILD model).  Every traced
 * segment, including the first from the source and one that escapes, that
 * crosses a sphere deposits I(s) * chord / (4/3 pi r^3) at the entry distance
 * s (the track-length form of counting crossings; a full crossing averages
 * I / (pi r^2)), with the a9 ramp of its channel, the sign (-1)^(reflections
 * so far), and dry on the first segment of a point source.  No shadow rays,
 * no analytic direct sound.  The expected deposit is the volume average of
 * the intensity over the sphere, 1 / (4 pi d^2) in free field for r << d. */
#define EAR_EST_NEE 0u
#define EAR_EST_SPHERE 1u

typedef struct {
    uint32_t band_mask;     /* bit b renders band b (0 low, 1 mid, 2 high) as its own context */
    float air[3];           /* absorbtion_factor per band = 1 - SET absorption[b] */
    float dry;              /* SET drylevel */
    ear_source source;
    const ear_recorder* recorders;
    uint32_t n_recorders;
    uint32_t n_rays;        /* num_samples */
    uint32_t max_bounces;   /* closest-hit queries per ray; the reference loops to 999;
                               at most 65535 (EAR_ERR_CAPACITY above) */
    double energy_cutoff;   /* path stops when intensity < cutoff; reference 1e-8 */
    uint64_t seed;          /* Philox key */
    uint32_t instance;      /* Philox counter word 3 (batch instance) */
    int32_t keyframe;       /* -1 static */
    uint32_t sound;         /* sound index (part of the Philox context word) */
    uint32_t hist_bins;     /* histogram length per channel (44.1 kHz bins) */
    uint32_t flags;         /* EAR_FLAG_* */
    uint32_t trace_rays;    /* EAR_FLAG_TRACE_LOG: rays [0, trace_rays) of each band */
    uint32_t trace_segs;    /* segments logged per ray */
    uint32_t estimator;     /* EAR_EST_NEE (default) or EAR_EST_SPHERE */
} ear_render_desc;

/* Histogram output, all host memory provided by the caller, laid out
 * [band slot b in 0..2][channel][bin] for the bands in band_mask (slots of
 * bands not rendered are left untouched).  Channels are the recorders'
 * channels in order (mono 1, stereo 2).
 *   raw:       int64 pairs {lo, hi} = the exact two's-complement 128-bit sum of
 *              the splatted ramps before 1/N, in units of 2^-EAR_FRAC_BITS.
 *   value:     (raw * 2^-EAR_FRAC_BITS / n_rays + direct) * gain^2 as double,
 *              i.e. what the reference's recorder track holds after
 *              Scene::Render (src/Scene.cpp:286-316).
 *   abs_raw / abs_value: the same for |ramp| (EAR_FLAG_ABS_HIST).
 * Any pointer may be NULL to skip that output. */
#define EAR_FRAC_BITS 80
typedef struct {
    int64_t* raw;
    double* value;
    int64_t* abs_raw;
    double* abs_value;
    int32_t* trace_hits;   /* [band slot][trace_rays][trace_segs]: tri index, -1 miss, -2 none */
    uint32_t* trace_nseg;  /* [band slot][trace_rays] */
} ear_hist;

typedef struct {
    uint64_t segments;        /* closest-hit queries (ray-segments) */
    uint64_t shadow_rays;     /* recorder connections traced */
    uint64_t contributions;   /* connections recorded */
    uint64_t rays;            /* rays emitted */
    uint64_t node_visits_closest, tri_tests_closest;   /* EAR_FLAG_INSTRUMENT */
    uint64_t node_visits_shadow, tri_tests_shadow;     /* EAR_FLAG_INSTRUMENT */
    uint32_t bounces_launched; /* wavefront iterations */
    uint32_t launches;         /* kernel launches */
    double trace_ms;           /* device time of the whole render (HIP events) */
    double extend_ms;          /* device time in the closest-hit kernel */
    double shade_ms;           /* device time in the shading kernel */
    double shadow_ms;          /* device time in the shadow + splat kernel */
    uint32_t extend_launches;
    uint32_t shadow_launches;
    uint64_t extend_queries;   /* rays handled by closest-hit launches */
    uint64_t shadow_queries;
    /* Ramps cut at the end of the histogram: a contribution (or the direct
     * sound) whose ramp runs past hist_bins loses its late bins.  The
     * reference grows its buffer on such a write instead (FloatBuffer::
     * operator[], src/Recorder.cpp:52-59); a caller that wants every bin
     * renders again with hist_bins >= bins_needed (same seed: same paths). */
    uint64_t dropped_ramps;
    uint32_t bins_needed;      /* the largest ramp end among them; 0 if none was cut */
    uint32_t reserved;
    /* The shadow kernels' share (the connections traced by the tail launch
     * excluded): connections recorded and the ramps they made (one per
     * channel: mono 1, stereo 2). */
    uint64_t shadow_contributions;
    uint64_t shadow_ramps;
    double splat_ms;           /* device time of the time-tiled splat's kernels (EAR_SPLAT_TILED) */
} ear_stats;

ear_status ear_render(ear_ctx* ctx, ear_scene* scene, const ear_render_desc* desc, ear_hist* out,
                      ear_stats* stats);

/* ear_render without any device allocation.  No counterpart in the
 * reference (its Scene::Render allocates freely): it is for callers that
 * cannot allocate device memory, such as code under HIP graph capture, tight
 * serving loops and harnesses that forbid hipMalloc.  Same arguments and
 * same results as ear_render (bitwise identical histograms, paths and
 * counts); the caller still passes the scene, the description and the host
 * output buffers.  Every device buffer of the render (ray and shadow queues,
 * hit records, traversal stacks, histogram accumulators, tables, counters)
 * is carved from one fixed region of EAR_NOALLOC_SCRATCH_BYTES per GPU,
 * held from the first ear_ctx_create on that GPU until the process ends; no
 * hipMalloc, hipFree, pinned-memory, event or graph call is made, and device
 * memory is cleared by a kernel, not a memset.  The ray queues shrink to
 * what the region leaves after the histograms and stacks (at most as many
 * rays in flight as ear_render keeps); a render whose histograms, stacks and
 * tables leave room for fewer than 4096 rays in flight fails with
 * EAR_ERR_CAPACITY before any launch (ear_ctx_last_error gives the sizes).
 * The traversal grids are capped at 16 waves per CU.  All calls on one GPU
 * share its region and its contents persist between calls; the library
 * serialises them within a process (a per-GPU lock held for the whole
 * call), so renders that must overlap use ear_render.  Timings in ear_stats
 * (trace_ms, extend_ms, shade_ms, shadow_ms) are not measured and read 0. */
#define EAR_NOALLOC_SCRATCH_BYTES (256u * 1024u * 1024u)
ear_status ear_render_noalloc(ear_ctx* ctx, ear_scene* scene, const ear_render_desc* desc, ear_hist* out,
                              ear_stats* stats);

/* A batch of renders in one persistent wavefront (the reference runs its
 * SceneContexts on a thread pool, src/EAR.cpp:193-207): every item is a
 * (scene, description, output) triple; the rays of all items share the
 * device queues.  Each item's histograms are bitwise identical to rendering it
 * alone.  All scenes must have the same BVH width. */
typedef struct {
    ear_scene* scene;
    const ear_render_desc* desc;
    ear_hist* out;
} ear_batch_item;
ear_status ear_render_batch(ear_ctx* ctx, const ear_batch_item* items, uint32_t n_items, ear_stats* stats);

/* Adjoint of ear_render (no counterpart in the reference, which has no
 * gradients; it serves inverse design: fitting materials to a measured
 * decay).  For the loss L = sum(adj_value * value) + sum(adj_abs * abs_value)
 * over the histograms ear_render would write for `desc` (adjoints laid out
 * as ear_hist.value, [3][n_channels][hist_bins], either may be NULL), writes
 * dL/drefl and dL/dspec as [n_mat][3] for every band of desc->band_mask
 * (entries of other bands are left untouched).  The rays replay the render's
 * paths exactly (same Philox stream); tangents follow the specular lobe's
 * reparameterised directions with the hit triangles and hemisphere samples
 * held fixed.  Arrival-time shifts and visibility changes are not
 * differentiated.  Point sources, 1..8 materials, EAR_EST_NEE.
 * With EAR_FLAG_ADJ_LOCAL the direction tangents are not carried to later
 * bounces: d/dspec keeps only the explicit blend term (lobe - diffuse) of the
 * connection at each vertex.  That estimate is biased but of bounded variance;
 * the full path derivative's variance grows with the bounce count (the
 * 1000th-power lobe amplifies the accumulated tangents).
 * With EAR_FLAG_ADJ_SCORE d/dspec is the score (likelihood-ratio) form: each
 * connection is weighted by the derivative of the log density of the
 * directions drawn before it, plus the rim term of the blend's moving
 * support, estimated by continuation paths spawned at a vertex with
 * probability F / 20, F the vertex's rim flux (at most 0.3 below s = 0.5).
 * Paths never move, so wall switches, visibility changes
 * and arrival-time shifts are all accounted for: the estimate is unbiased at
 * any bounce count.  It needs every specularity of the rendered bands below
 * 0.5 (EAR_ERR_INVALID_ARGUMENT otherwise); d/drefl is the same as without
 * the flag.  EAR_FLAG_ADJ_SCORE and EAR_FLAG_ADJ_LOCAL exclude each other, and
 * a batch uses the score form for all its items or for none. */
ear_status ear_render_backward(ear_ctx* ctx, ear_scene* scene, const ear_render_desc* desc,
                               const double* adj_value, const double* adj_abs, double* grad_refl,
                               double* grad_spec);

/* The adjoint of a batch (no reference counterpart): every item's gradients
 * as ear_render_backward would write them, with the rays of all items sharing
 * the device queues band by band (fitting many rooms at once).  All scenes
 * must have the same BVH width. */
typedef struct {
    ear_scene* scene;
    const ear_render_desc* desc;
    const double* adj_value;  /* [3][n_channels][hist_bins] or NULL */
    const double* adj_abs;    /* [3][n_channels][hist_bins] or NULL */
    double* grad_refl;        /* [n_mat][3] */
    double* grad_spec;        /* [n_mat][3] */
} ear_backward_item;
ear_status ear_render_backward_batch(ear_ctx* ctx, const ear_backward_item* items, uint32_t n_items);

/* Auralization: the convolution of an impulse-response track with a sound
 * signal, RecorderTrack::Process (src/Recorder.cpp:145-243).  out[i], for i
 * in [0, n_out), n_out = max(n_ir, n_ir2) + n_sig + 1, is the linear
 * convolution of ir with the signal; when ir2 is given (the keyframe
 * cross-fade, src/Recorder.cpp:233-243) the signal is faded out linearly
 * over its length for ir and faded in for ir2, and the two are added.  The
 * fades follow the reference's running-sum factors.  Single precision
 * (float FFTs): results agree with an exact convolution to about 1e-6 of the
 * largest output sample. */
ear_status ear_convolve(ear_ctx* ctx, const float* ir, uint32_t n_ir, const float* ir2, uint32_t n_ir2,
                        const float* signal, uint32_t n_sig, float* out);

/* Launch configuration of the tracer (block sizes, ray chunk, stack split);
 * values of 0 keep the library default.  Results do not depend on them
 * (integer histogram sums are order-independent). */
typedef struct {
    uint32_t block_extend;
    uint32_t block_shade;
    uint32_t block_shadow;
    uint32_t ray_chunk;      /* queue capacity = 3 x ray_chunk rays in flight (default
                                2^23); finished rays are replaced every iteration */
    uint32_t waves_per_cu;   /* persistent grid size */
    uint32_t stack_lds;      /* traversal stack entries kept in LDS (16 or 32) */
    uint32_t traversal;      /* EAR_TRAVERSAL_*: how BVH walks are assigned to lanes */
    uint32_t graph_iters;    /* > 0: bounce-loop iterations captured in one HIP graph and
                                replayed (per-kernel times are then not measured); after a
                                replay the last live rays go to the tail launch as in the
                                eager loop */
    uint32_t refill;         /* refilling traversal: idle lanes (of 64) that trigger a refill
                                (default 24) */
    uint32_t fetch_chunk;    /* refilling traversal: queue entries a wave claims per atomic
                                (0: one per idle lane) */
    uint32_t fetch_segments; /* refilling traversal: queue segments with a claim counter each
                                (1..64, default 64) */
    uint32_t tail_rays;      /* once every ray is emitted and at most this many are live,
                                one launch traces each through all its remaining segments
                                (same histograms); 0: default, 2^20, or every ray when no
                                scene has more than 64 triangles; 1: never (the bounce loop
                                to the end) */
    uint32_t splat;          /* EAR_SPLAT_*: how the shadow kernels add ramps to the
                                histograms (same histograms either way) */
} ear_tuning;
#define EAR_SPLAT_DEFAULT 0u /* EAR_SPLAT_TILED when the call's (row, tile) keys number at most
                                8192 (3 bands x channels x kinds x tiles: a 3 s stereo render with
                                the |c| histogram has 6204) and its queues hold at least 2^22
                                rays, else EAR_SPLAT_ATOMIC */
#define EAR_SPLAT_ATOMIC 1u  /* each ramp point is a 128-bit add to the histogram in HBM */
#define EAR_SPLAT_TILED 2u   /* visible connections write records keyed by (row, time tile of
                                256 bins); per bounce iteration a counting sort groups them, and
                                workgroups of 8192 sorted records accumulate each tile's ramps
                                in LDS (64-bit LDS atomics), then add the tile to HBM once
                                (ramps wider than 62 bins are added directly) */
#define EAR_TRAVERSAL_DEFAULT 0u   /* = EAR_TRAVERSAL_FETCH_ALL */
#define EAR_TRAVERSAL_STATIC 1u    /* one queue entry per lane per grid-stride step */
#define EAR_TRAVERSAL_FETCH 2u     /* closest hit: persistent lanes take the next entry when
                                      their walk ends; shadow rays one per lane */
#define EAR_TRAVERSAL_FETCH_ALL 3u /* both kernels refill lanes */
ear_status ear_ctx_set_tuning(ear_ctx* ctx, const ear_tuning* t);
ear_status ear_ctx_get_tuning(const ear_ctx* ctx, ear_tuning* t);

#ifdef __cplusplus
}
#endif
#endif
