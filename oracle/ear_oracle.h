/*
 * ear_oracle.h -- CPU oracle for the E.A.R acoustic ray tracer.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke
 * check and the CPU-baseline leg of the benchmark load it.  The product
 * (libear_hip.so and the host C++ layer) never includes, links or calls it.
 *
 * It restates, in plain C11, what the reference's Scene::Render does for one
 * (sound, keyframe, band) context (src/Scene.cpp:111-318) and everything that
 * function calls:
 *   - Mesh::RayIntersection / LineIntersection  (src/Mesh.cpp:33-71)
 *   - Scene::Bounce / Scene::Connect             (src/Scene.cpp:49-96)
 *   - Material ctor + Material::Bounce           (src/Material.cpp:26-83)
 *   - Sample_Sphere / Sample_Hemi                (src/Distributions.h:48-75)
 *   - AbstractSoundFile::SoundRay                (src/SoundFile.cpp:215-228)
 *   - Mesh::SamplePoint, Triangle::SamplePoint   (src/Mesh.cpp:143-154,
 *                                                 src/Triangle.cpp:44-53)
 *   - MonoRecorder::Record / StereoRecorder::Record
 *                                                (src/MonoRecorder.cpp:83-97,
 *                                                 src/StereoRecorder.cpp:97-130)
 * with the random stream replaced by counter-based Philox4x32-10 (the
 * reference seeds libc rand() with time(NULL), src/Scene.cpp:116, so its
 * output is not reproducible) and the un-vendored gmtl 0.6.1 arithmetic
 * restated from the textbook routines (Moeller-Trumbore 1997 non-culling
 * test, reflect, normalize).
 *
 * Histograms are accumulated two ways: in FP64 (exact ramp (w-i)*S per bin,
 * signed and |c|) and, single-threaded only, in the reference's own fp32
 * iterated-subtraction form.
 *
 * Parity unpinned (no reference output exists to check them against): the
 * exact bits of gmtl's intersectDoubleSided (its epsilon, here |det| < 1e-5,
 * and operation order), gmtl's unitRandom float mapping and the glibc rand()
 * sequence (both replaced by Philox, u = (word >> 8) * 2^-24), and the
 * operand order of gmtl's normal (here normalize(cross(v1 - v0, v2 - v0))).
 * Each is fixed here once, and the device path is held bit-exact to it; the
 * geometry itself is pinned by brute-force FP64 checks and closed forms in
 * the tests.
 */
#ifndef EAR_ORACLE_H
#define EAR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_REC_MONO 0
#define OR_REC_STEREO 1
#define OR_SRC_POINT 0
#define OR_SRC_MESH 1

#define OR_MAT_HAS_REFR 1u
#define OR_MAT_HAS_SPEC 2u

typedef struct {
    float refl[3];
    float refr[3];
    float spec[3];
    uint32_t flags; /* OR_MAT_HAS_REFR | OR_MAT_HAS_SPEC: optional .EAR fields present */
} or_material;

typedef struct {
    int32_t kind; /* OR_REC_MONO / OR_REC_STEREO */
    float pos[3];
    float right_ear[3];
    float head_size;
    float head_abs[3]; /* already max(0,(1-h)^4) as in the StereoRecorder ctor */
    float radius;      /* OR_EST_SPHERE: receiver sphere radius */
} or_recorder;

/* receiver estimators: the reference's connections (NEE) or spheres that
 * collect the traced segments crossing them (no reference counterpart) */
#define OR_EST_NEE 0u
#define OR_EST_SPHERE 1u

typedef struct {
    uint32_t n_tris;
    const float* v;       /* [n_tris][3 vertices][xyz] */
    const uint16_t* mat;  /* [n_tris] */
    uint32_t n_mat;
    const or_material* mats;
} or_scene;

typedef struct {
    int32_t band;
    float air; /* absorbtion_factor = 1 - absorption[band] */
    float dry;
    float gain;
    int32_t src_kind;
    float src_pos[3];
    uint32_t src_n_tris;
    const float* src_v; /* mesh source triangles [n][3][3] */
    uint32_t n_rec;
    const or_recorder* rec;
    uint32_t ray_begin, ray_end; /* traced sub-range of [0, n_rays) */
    uint32_t n_rays;             /* N used for 1/N */
    uint32_t max_bounces;        /* reference: 999 (num_bounces < 1000) */
    double cutoff;               /* reference: 1e-8 */
    uint64_t seed;
    uint32_t ctx;      /* context word of the Philox counter */
    uint32_t instance; /* instance word of the Philox counter */
    uint32_t n_bins;
    uint32_t n_threads;
    uint32_t trace_rays, trace_segs;
    uint32_t estimator; /* OR_EST_* */
} or_desc;

typedef struct {
    double* hist_signed;  /* [n_ch][n_bins] raw sums (before 1/N), may be NULL */
    double* hist_abs;     /* [n_ch][n_bins] raw sums of |ramp|, may be NULL */
    float* hist_f32;      /* [n_ch][n_bins] reference fp32 accumulation (n_threads==1) */
    int32_t* trace_hits;  /* [trace_rays][trace_segs] original tri index, -1 miss, -2 unused */
    uint32_t* trace_nseg; /* [trace_rays] closest-hit queries per ray */
    uint32_t* trace_ncon; /* [trace_rays] recorded contributions per ray */
    uint64_t segments, shadow_rays, contributions;
    /* direct sound per channel (src/Scene.cpp:299-311): */
    int32_t* direct_visible; /* [n_ch] */
    int32_t* direct_s;       /* [n_ch] first bin */
    int32_t* direct_w;       /* [n_ch] ramp width */
    float* direct_a;         /* [n_ch] ramp amplitude (fp32, 2a/sqrt(d) and ILD applied) */
    double* tally;           /* [n_rec] sum of recorded contributions c (before the splat) */
    double* tally_abs;       /* [n_rec] sum of |c| */
    double path_length;      /* sum of traced segment lengths (hits only) */
    float* trace_rays_od;    /* [trace_rays][trace_segs][6] origin + direction of each query, may be NULL */
} or_out;

/* ---- building blocks (exported for the tests) ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float or_uniform(uint64_t seed, uint32_t ray, uint32_t ctx, uint32_t instance, uint32_t draw);
float or_pow(float x, float y);
double or_pow_d(double x, double y);
int or_mt_intersect(const float v0[3], const float e1[3], const float e2[3],
                    const float o[3], const float d[3], float* t);
int or_closest_hit(const or_scene* s, const float o[3], const float d[3], float* t, int32_t* idx);
int or_any_hit(const or_scene* s, const float p[3], const float q[3]);
void or_tri_normal(const float* v9, float n[3]);
int or_material_retention(const or_material* m, float ret[3]);
/* draws n samples from ray stream `ray`; mode 0 sphere, 1 hemisphere about n;
 * writes xyz to out and the number of uniforms consumed per sample to ndraws */
void or_sample_dirs(uint64_t seed, uint32_t ctx, uint32_t instance, uint32_t n, int mode,
                    const float nrm[3], float* out, uint32_t* ndraws);
/* reference ramp splat, fp32 iterated form, into buf[n] (src/MonoRecorder.cpp:83-97) */
void or_record_f32(float* buf, uint32_t n, float a, float t, float dist);

/* ---- the per-context render ---- */
int or_render(const or_scene* s, const or_desc* d, or_out* out);
int or_num_channels(const or_desc* d);
/* FP64 reference gradient of sum(adj_signed * H_signed + adj_abs * H_abs)
 * (raw sums, before 1/N) with respect to the rendered band's per-material
 * reflection coefficient and specularity; adjoints are [n_ch][n_bins] or NULL.
 * local == 1: direction tangents are not carried past the vertex that
 * sampled them (only each connection's explicit dependence on spec).
 * local == 2: the score form of the specularity gradient (unbiased; point
 * sources, specularities below 1/2, else -1 / -3), a continuation path
 * spawned at a vertex with probability min(1, q_spawn F), F its rim flux
 * (or_gradient: q_spawn = OR_SCORE_SPAWN). */
#define OR_SCORE_SPAWN 0.05f
int or_gradient(const or_scene* s, const or_desc* d, const double* adj_signed, const double* adj_abs, int local,
                double* grad_refl, double* grad_spec);
int or_gradient_k(const or_scene* s, const or_desc* d, const double* adj_signed, const double* adj_abs, int local,
                  float q_spawn, double* grad_refl, double* grad_spec);
/* the score form's closed forms: d log p(v)/ds of the blend's direction
 * density at c = v.r, the rim flux F at cr = r.nu, and rim directions */
double or_score_dlogp(double c, double s);
double or_score_flux(double cr, double s);
void or_score_rim_dirs(uint64_t seed, uint32_t n, const float nu[3], const float r[3], float s, float* out);
double or_now(void);

/* ---- auralization: the three-band equalizer in single precision ---- */
/* coefficients of one fourth-order Linkwitz-Riley section: numerator a[5],
 * denominator b[5] (b[0] = 1) */
void or_lr4_coefficients(float fc, int high_pass, float a[5], float b[5]);
/* Equalizer::Split: low = LP(fc1), mid = HP(fc1) LP(fc2), high = HP(fc2) */
void or_equalizer_split(const float* x, uint32_t n, float* low, float* mid, float* high, float f1, float f2,
                        float f3);

#ifdef __cplusplus
}
#endif
#endif
