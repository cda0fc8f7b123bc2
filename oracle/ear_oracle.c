/*
 * ear_oracle.c -- CPU oracle for the E.A.R acoustic ray tracer.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the
 * CPU-baseline leg of bench.py as the checker.  The product never links it.
 *
 * Plain C11, built with -O3 -ffp-contract=off (no fused multiply-add, no
 * fast-math) so that every float operation below is one IEEE-754 binary32 or
 * binary64 operation in the order written.  The arithmetic contract it pins
 * (operation order, rounding points, random-stream discipline) is the one the
 * HIP kernels implement independently.
 *
 * Reference functions restated here (file:line in aothms/ear):
 *   Scene::Render             src/Scene.cpp:111-318   -> or_render / trace_ray
 *   Scene::Bounce             src/Scene.cpp:49-82     -> trace_ray (k > 0 branch)
 *   Scene::Connect            src/Scene.cpp:84-96     -> or_any_hit
 *   Mesh::RayIntersection     src/Mesh.cpp:33-56      -> or_closest_hit
 *   Mesh::LineIntersection    src/Mesh.cpp:58-71      -> or_any_hit
 *   Mesh::SamplePoint         src/Mesh.cpp:143-154    -> emit_mesh
 *   Triangle::SamplePoint     src/Triangle.cpp:44-53  -> emit_mesh
 *   Triangle::calcArea        src/Triangle.cpp:26-30  -> tri_area
 *   Material::Material        src/Material.cpp:26-72  -> or_material_retention
 *   Material::Bounce          src/Material.cpp:76-83  -> material_bounce
 *   Sample_Sphere/Sample_Hemi src/Distributions.h:48-75
 *   MonoRecorder::Record      src/MonoRecorder.cpp:83-97
 *   StereoRecorder::Record    src/StereoRecorder.cpp:97-130
 *   Equalizer::Pass / Split   lib/equalizer/Equalizer.cpp:27-95 -> or_lr4_coefficients,
 *                                                     or_equalizer_split
 * and, with no reference counterpart, the receiver-sphere estimator
 * (sphere_deposits) and the FP64 gradient of a loss on the histograms
 * (or_gradient).
 *   gmtl 0.6.1 (not vendored): intersectDoubleSided = Moeller-Trumbore 1997
 *   non-culling test with |det| < 1e-5 rejected; reflect(v,n) = v - 2(v.n)n;
 *   normalize(v) = v / sqrt(v.v); unitRandom -> Philox4x32-10 (Salmon et al.,
 *   SC'11) with u = (x >> 8) * 2^-24.
 */
#define _POSIX_C_SOURCE 200809L
#include "ear_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define OR_PI 3.14159265f /* src/Distributions.h:37 */
#define OR_EXP 1000.0f    /* src/Scene.cpp:46 */
#define OR_EXP_INT (OR_EXP + 1.0f)

/* ------------------------------------------------------------------------ */
/* Philox4x32-10                                                             */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

typedef struct {
    uint32_t key[2];
    uint32_t ray, ctx, instance;
    uint32_t buf[4];
    uint32_t draw;
} or_rng;

static void rng_init(or_rng* r, uint64_t seed, uint32_t ray, uint32_t ctx, uint32_t instance) {
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->ray = ray;
    r->ctx = ctx;
    r->instance = instance;
    r->draw = 0;
}

/* uniform number `draw` of the stream: word (draw & 3) of block (draw >> 2) */
static float rng_next(or_rng* r) {
    uint32_t w = r->draw & 3u;
    if (w == 0) {
        uint32_t ctr[4] = {r->draw >> 2, r->ray, r->ctx, r->instance};
        or_philox4x32_10(ctr, r->key, r->buf);
    }
    r->draw++;
    return (float)(r->buf[w] >> 8) * (1.0f / 16777216.0f);
}

float or_uniform(uint64_t seed, uint32_t ray, uint32_t ctx, uint32_t instance, uint32_t draw) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {draw >> 2, ray, ctx, instance}, out[4];
    or_philox4x32_10(ctr, key, out);
    return (float)(out[draw & 3u] >> 8) * (1.0f / 16777216.0f);
}

/* ------------------------------------------------------------------------ */
/* pow: the reference calls pow(float,float) (float overload) at            */
/* src/Scene.cpp:154,232,243,251,308 and powf at src/StereoRecorder.cpp:110. */
/* Restated as exp2(y*log2 x) evaluated in binary64 with +,-,*,/ only:      */
/* log via the atanh series ln m = 2 sum s^(2j+1)/(2j+1), s=(m-1)/(m+1),    */
/* m in [sqrt(1/2), sqrt 2); exp2 via n = nearest(z), 2^f = exp(f ln2) by a */
/* degree-17 Taylor polynomial; the binary32 result is the round-to-nearest */
/* of the binary64 value.                                                   */
/* ------------------------------------------------------------------------ */
static uint64_t dbits(double x) {
    uint64_t b;
    memcpy(&b, &x, 8);
    return b;
}
static double dfrom(uint64_t b) {
    double x;
    memcpy(&x, &b, 8);
    return x;
}

static double log2_pos(double x) { /* x > 0, finite, normal */
    uint64_t b = dbits(x);
    int e = (int)((b >> 52) & 0x7ffu) - 1023;
    double m = dfrom((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double p = 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    double ln_m = 2.0 * s * p;
    return (double)e + ln_m * 1.4426950408889634; /* 1/ln 2 */
}

static double exp2_d(double z) {
    if (z < -1060.0) return 0.0;
    if (z > 1060.0) return INFINITY;
    double zn = z + 0.5;
    long long n = (long long)zn; /* trunc */
    if ((double)n > zn) n -= 1;  /* floor(z + 0.5) */
    double f = z - (double)n;    /* exact, |f| <= 0.5 */
    double g = f * 0.6931471805599453;
    double p = 1.0 / 355687428096000.0; /* 1/17! */
    p = p * g + 1.0 / 20922789888000.0; /* 1/16! */
    p = p * g + 1.0 / 1307674368000.0;
    p = p * g + 1.0 / 87178291200.0;
    p = p * g + 1.0 / 6227020800.0;
    p = p * g + 1.0 / 479001600.0;
    p = p * g + 1.0 / 39916800.0;
    p = p * g + 1.0 / 3628800.0;
    p = p * g + 1.0 / 362880.0;
    p = p * g + 1.0 / 40320.0;
    p = p * g + 1.0 / 5040.0;
    p = p * g + 1.0 / 720.0;
    p = p * g + 1.0 / 120.0;
    p = p * g + 1.0 / 24.0;
    p = p * g + 1.0 / 6.0;
    p = p * g + 0.5;
    p = p * g + 1.0;
    p = p * g + 1.0;
    /* scale by 2^n in two exact steps so that 2^n factors stay normal */
    long long n1 = n / 2, n2 = n - n1;
    double s1 = dfrom((uint64_t)(n1 + 1023) << 52);
    double s2 = dfrom((uint64_t)(n2 + 1023) << 52);
    return (p * s1) * s2;
}

double or_pow_d(double x, double y) {
    if (isnan(x) || isnan(y)) return NAN;
    if (y == 0.0 || x == 1.0) return 1.0;
    if (x == 0.0) return y > 0.0 ? 0.0 : INFINITY;
    if (x < 0.0) return NAN;
    if (isinf(x)) return y > 0.0 ? INFINITY : 0.0;
    if (isinf(y)) return (x < 1.0) == (y > 0.0) ? 0.0 : INFINITY;
    return exp2_d(y * log2_pos(x));
}

float or_pow(float x, float y) { return (float)or_pow_d((double)x, (double)y); }

/* ------------------------------------------------------------------------ */
/* vector helpers: every component operation one rounding, left to right    */
/* ------------------------------------------------------------------------ */
static inline float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const float a[3], const float b[3], float r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void sub3(const float a[3], const float b[3], float r[3]) {
    r[0] = a[0] - b[0];
    r[1] = a[1] - b[1];
    r[2] = a[2] - b[2];
}
static inline float len3(const float a[3]) { return sqrtf(dot3(a, a)); }
/* gmtl::normalize: divide each component by the length unless it is 0 */
static inline void normalize3(float a[3]) {
    float l = len3(a);
    if (l != 0.0f) {
        a[0] = a[0] / l;
        a[1] = a[1] / l;
        a[2] = a[2] / l;
    }
}
/* gmtl::reflect(result, v, n) = v - 2 (v.n) n */
static inline void reflect3(const float v[3], const float n[3], float r[3]) {
    float t = dot3(v, n);
    r[0] = v[0] - 2.0f * (t * n[0]);
    r[1] = v[1] - 2.0f * (t * n[1]);
    r[2] = v[2] - 2.0f * (t * n[2]);
}
static inline int is_normal_f(float x) { return fpclassify(x) == FP_NORMAL; }

/* gmtl::intersectDoubleSided(Tri, Ray, u, v, t): Moeller-Trumbore (JGT 1997)
 * non-culling branch, EPSILON = 1e-5 on the determinant. */
int or_mt_intersect(const float v0[3], const float e1[3], const float e2[3], const float o[3], const float d[3],
                    float* t) {
    float pvec[3], tvec[3], qvec[3];
    cross3(d, e2, pvec);
    float det = dot3(e1, pvec);
    if (det > -0.00001f && det < 0.00001f) return 0;
    float inv_det = 1.0f / det;
    sub3(o, v0, tvec);
    float u = dot3(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return 0;
    cross3(tvec, e1, qvec);
    float v = dot3(d, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return 0;
    *t = dot3(e2, qvec) * inv_det;
    return 1;
}

/* triangle normal: normalize(cross(v1 - v0, v2 - v0)) (gmtl::normal(Tri),
 * src/Triangle.cpp:32) */
void or_tri_normal(const float* v9, float n[3]) {
    float e1[3], e2[3];
    sub3(v9 + 3, v9, e1);
    sub3(v9 + 6, v9, e2);
    cross3(e1, e2, n);
    normalize3(n);
}

/* Triangle::calcArea, src/Triangle.cpp:26-30: |edge(0) x edge(1)| / 2 with
 * gmtl Tri::edge(i) = v[(i+1)%3] - v[i] */
static float tri_area(const float* v9) {
    float a[3], b[3], c[3];
    sub3(v9 + 3, v9, a);
    sub3(v9 + 6, v9 + 3, b);
    cross3(a, b, c);
    return len3(c) / 2.0f;
}

/* Material ctor, src/Material.cpp:26-72: the field the reference calls
 * absorption_coefficient ends up holding retention = 1 - (1 - (fl-1e-9) - (fr-1e-9)).
 * Returns 0 (the reference throws DatatypeException) when refl + refr > 1. */
int or_material_retention(const or_material* m, float ret[3]) {
    for (int i = 0; i < 3; i++) {
        float a = 1.0f;
        a -= m->refl[i] - 1e-9f;
        if (m->flags & OR_MAT_HAS_REFR) a -= m->refr[i] - 1e-9f;
        if (a < 0.0f) return 0;
        ret[i] = 1.0f - a;
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* scene queries (brute force, as the reference)                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t n;
    float* v0; /* [n][3] */
    float* e1;
    float* e2;
    float* nrm;
    uint16_t* mat;
} prep_scene;

static int prep(const or_scene* s, prep_scene* p) {
    p->n = s->n_tris;
    size_t n = s->n_tris ? s->n_tris : 1;
    p->v0 = (float*)malloc(n * 3 * sizeof(float));
    p->e1 = (float*)malloc(n * 3 * sizeof(float));
    p->e2 = (float*)malloc(n * 3 * sizeof(float));
    p->nrm = (float*)malloc(n * 3 * sizeof(float));
    p->mat = (uint16_t*)malloc(n * sizeof(uint16_t));
    if (!p->v0 || !p->e1 || !p->e2 || !p->nrm || !p->mat) return 0;
    for (uint32_t i = 0; i < s->n_tris; i++) {
        const float* v = s->v + 9 * (size_t)i;
        memcpy(p->v0 + 3 * i, v, 3 * sizeof(float));
        sub3(v + 3, v, p->e1 + 3 * i);
        sub3(v + 6, v, p->e2 + 3 * i);
        or_tri_normal(v, p->nrm + 3 * i);
        p->mat[i] = s->mat ? s->mat[i] : 0;
    }
    return 1;
}
static void unprep(prep_scene* p) {
    free(p->v0);
    free(p->e1);
    free(p->e2);
    free(p->nrm);
    free(p->mat);
}

/* Mesh::RayIntersection: accept t in (0.001, d), d starts at 1e6; strict '<'
 * keeps the lowest index on ties. */
static int closest_hit_p(const prep_scene* s, const float o[3], const float d[3], float* t_out, int32_t* idx) {
    float best = 1000000.0f;
    int32_t bi = -1;
    for (uint32_t i = 0; i < s->n; i++) {
        float t;
        if (or_mt_intersect(s->v0 + 3 * i, s->e1 + 3 * i, s->e2 + 3 * i, o, d, &t) && t > 0.001f && t < best) {
            best = t;
            bi = (int32_t)i;
        }
    }
    if (bi < 0) return 0;
    *t_out = best;
    *idx = bi;
    return 1;
}

/* Mesh::LineIntersection on LineSeg(p, q): direction q - p, t in (1e-5, 1) */
static int any_hit_p(const prep_scene* s, const float p[3], const float dv[3]) {
    for (uint32_t i = 0; i < s->n; i++) {
        float t;
        if (or_mt_intersect(s->v0 + 3 * i, s->e1 + 3 * i, s->e2 + 3 * i, p, dv, &t) && t > 1e-5f && t < 1.0f) return 1;
    }
    return 0;
}

int or_closest_hit(const or_scene* s, const float o[3], const float d[3], float* t, int32_t* idx) {
    prep_scene p;
    if (!prep(s, &p)) return -1;
    int r = closest_hit_p(&p, o, d, t, idx);
    unprep(&p);
    return r;
}

int or_any_hit(const or_scene* s, const float p[3], const float q[3]) {
    prep_scene ps;
    if (!prep(s, &ps)) return -1;
    float dv[3];
    sub3(q, p, dv);
    int r = any_hit_p(&ps, p, dv);
    unprep(&ps);
    return r;
}

/* ------------------------------------------------------------------------ */
/* sampling, src/Distributions.h:48-75                                      */
/* ------------------------------------------------------------------------ */
static void sample_sphere(or_rng* r, float v[3]) {
    for (;;) {
        float f1 = rng_next(r) * 2.0f - 1.0f;
        float f2 = rng_next(r) * 2.0f - 1.0f;
        float f3 = rng_next(r) * 2.0f - 1.0f;
        float l = f1 * f1 + f2 * f2 + f3 * f3;
        if (l < 0.001f || l > 1.0f) continue;
        float s = sqrtf(l);
        v[0] = f1 / s;
        v[1] = f2 / s;
        v[2] = f3 / s;
        return;
    }
}
static void sample_hemi(or_rng* r, const float n[3], float v[3]) {
    for (;;) {
        sample_sphere(r, v);
        if (dot3(n, v) < 0.0f) continue;
        return;
    }
}
static void sample_hemi_spec(or_rng* r, const float n[3], const float refl[3], float f, float v[3]) {
    sample_hemi(r, n, v);
    float g = 1.0f - f;
    v[0] = v[0] * g + refl[0] * f;
    v[1] = v[1] * g + refl[1] * f;
    v[2] = v[2] * g + refl[2] * f;
    normalize3(v);
}

void or_sample_dirs(uint64_t seed, uint32_t ctx, uint32_t instance, uint32_t n, int mode, const float nrm[3],
                    float* out, uint32_t* ndraws) {
    for (uint32_t i = 0; i < n; i++) {
        or_rng r;
        rng_init(&r, seed, i, ctx, instance);
        if (mode == 0)
            sample_sphere(&r, out + 3 * i);
        else
            sample_hemi(&r, nrm, out + 3 * i);
        if (ndraws) ndraws[i] = r.draw;
    }
}

#define OR_REFLECT 0
#define OR_REFRACT 1

/* ------------------------------------------------------------------------ */
/* render                                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    const or_desc* d;
    const prep_scene* s;
    const float (*ret)[3];   /* retention per material */
    const or_material* mats;
    /* mesh source */
    const float* src_v;
    const float* src_area;
    float src_total_area;
    const float* src_nrm;
    int n_ch;
    const int* ch_of_rec; /* first channel of recorder r */
    /* OR_EST_SPHERE: per recorder radius^2, 3 / (4 pi r^3), sphere centres */
    const float* sph_r2;
    const float* sph_inv_vol;
    const float (*sph_cen)[2][3];
} rctx;

typedef struct {
    double* hs;
    double* ha;
    float* hf;
    double* tally;
    double* tally_abs;
    double path_length;
    uint64_t segments, shadow_rays, contributions;
} acc_t;

static void add_ramp(const rctx* C, acc_t* acc, int ch, int s, int w, float A) {
    uint32_t nb = C->d->n_bins;
    double S = (double)A / (double)w;
    for (int i = 0; i < w; i++) {
        long long b = (long long)s + i;
        if (b < 0 || b >= (long long)nb) continue;
        double v = (double)(w - i) * S;
        if (acc->hs) acc->hs[(size_t)ch * nb + (size_t)b] += v;
        if (acc->ha) acc->ha[(size_t)ch * nb + (size_t)b] += fabs(v);
    }
}

/* MonoRecorder::Record fp32 form into one track */
static void ramp_f32(float* buf, uint32_t nb, int s, int w, float ampl) {
    float step = ampl / (float)w;
    for (int i = 0; i < w; i++) {
        long long b = (long long)s + i;
        if (b >= 0 && b < (long long)nb) buf[b] += ampl;
        ampl -= step;
    }
}

void or_record_f32(float* buf, uint32_t n, float a, float t, float dist) {
    int s = (int)((double)t * 44100.0);
    float width = sqrtf(dist);
    float ampl = 2.0f * a / width;
    int w = (int)ceilf(width);
    ramp_f32(buf, n, s, w, ampl);
}

/* Recorder::Record for recorder r: mono (src/MonoRecorder.cpp:83-97) or
 * stereo (src/StereoRecorder.cpp:97-130); fills the FP64 ramps and, when
 * requested, the reference fp32 track. */
static void record(const rctx* C, acc_t* acc, int r, const float dir[3], float a, float t, float dist) {
    const or_recorder* rec = &C->d->rec[r];
    int ch = C->ch_of_rec[r];
    uint32_t nb = C->d->n_bins;
    acc->tally[r] += (double)a;
    acc->tally_abs[r] += fabs((double)a);
    if (rec->kind == OR_REC_MONO) {
        int s = (int)((double)t * 44100.0);
        float width = sqrtf(dist);
        float ampl = 2.0f * a / width;
        int w = (int)ceilf(width);
        add_ramp(C, acc, ch, s, w, ampl);
        if (acc->hf) ramp_f32(acc->hf + (size_t)ch * nb, nb, s, w, ampl);
    } else {
        float dt = dot3(dir, rec->right_ear);
        float td = rec->head_size / 343.0f;
        int s_right = (int)((double)(t - dt * td) * 44100.0);
        int s_left = (int)((double)(t + dt * td) * 44100.0);
        float width = sqrtf(dist);
        float ampl = 2.0f * a / width;
        float al = ampl, ar = ampl;
        float idiff = fabsf(dt);
        float factor = or_pow(rec->head_abs[C->d->band], idiff * rec->head_size);
        if (dt < 0.0f)
            ar *= factor * factor;
        else
            al *= factor * factor;
        int w = (int)ceilf(width);
        add_ramp(C, acc, ch, s_left, w, al);
        add_ramp(C, acc, ch + 1, s_right, w, ar);
        if (acc->hf) {
            ramp_f32(acc->hf + (size_t)ch * nb, nb, s_left, w, al);
            ramp_f32(acc->hf + (size_t)(ch + 1) * nb, nb, s_right, w, ar);
        }
    }
}

/* Receiver spheres (no reference counterpart): the segment o + s u, s in
 * [0, seg), deposits into each sphere it crosses I0 air^s0 (s1 - s0) *
 * 3 / (4 pi r^3) at distance L0 + s0, [s0, s1) the chord clipped to the
 * segment; sign (-1)^j for j reflections so far, dry on a point source's
 * first segment; each sphere is one channel with the mono ramp. */
static void sphere_deposits(const rctx* C, acc_t* acc, const float o[3], const float u[3], float seg, float I0,
                            float L0, uint32_t j) {
    const or_desc* D = C->d;
    uint32_t nb = D->n_bins;
    for (uint32_t r = 0; r < D->n_rec; r++) {
        int nc = D->rec[r].kind == OR_REC_STEREO ? 2 : 1;
        for (int e = 0; e < nc; e++) {
            float oc[3];
            sub3(o, C->sph_cen[r][e], oc);
            float b = dot3(oc, u);
            float disc = b * b - (dot3(oc, oc) - C->sph_r2[r]);
            if (!(disc > 0.0f)) continue;
            float q = sqrtf(disc);
            float s0 = fmaxf(-b - q, 0.0f), s1 = fminf(-b + q, seg);
            if (!(s1 > s0)) continue;
            float c = I0;
            c *= or_pow(D->air, s0);
            c *= s1 - s0;
            c *= C->sph_inv_vol[r];
            if (j == 0 && D->src_kind == OR_SRC_POINT) c *= D->dry;
            if (!is_normal_f(c)) continue;
            if (j & 1u) c *= -1.0f;
            float dist = L0 + s0;
            float t = dist / 343.0f;
            int ch = C->ch_of_rec[r] + e;
            int sb = (int)((double)t * 44100.0);
            float width = sqrtf(dist);
            float ampl = 2.0f * c / width;
            int w = (int)ceilf(width);
            acc->tally[r] += (double)c;
            acc->tally_abs[r] += fabs((double)c);
            add_ramp(C, acc, ch, sb, w, ampl);
            if (acc->hf) ramp_f32(acc->hf + (size_t)ch * nb, nb, sb, w, ampl);
            acc->contributions++;
        }
    }
}

/* emission from a mesh source: Mesh::SamplePoint + Triangle::SamplePoint +
 * 2-argument Sample_Hemi about the (unflipped) triangle normal. */
static void emit_mesh(const rctx* C, or_rng* r, float o[3], float d[3]) {
    const or_desc* D = C->d;
    float x = rng_next(r) * C->src_total_area; /* rangeRandom(0, total_area) */
    uint32_t pick = D->src_n_tris ? D->src_n_tris - 1 : 0;
    for (uint32_t i = 0; i < D->src_n_tris; i++) {
        x -= C->src_area[i];
        if (x < 0.0f) {
            pick = i;
            break;
        }
    }
    const float* v = C->src_v + 9 * (size_t)pick;
    float r1 = rng_next(r);
    float r2 = rng_next(r);
    float sr1 = sqrtf(r1);
    float a = 1.0f - sr1, b = sr1 * (1.0f - r2), c = sr1 * r2;
    for (int k = 0; k < 3; k++) o[k] = a * v[k] + b * v[3 + k] + c * v[6 + k];
    sample_hemi(r, C->src_nrm + 3 * pick, d);
}

static uint32_t material_bounce(const rctx* C, or_rng* r, int mat, int band) {
    float fl = C->mats[mat].refl[band];
    float fr = (C->mats[mat].flags & OR_MAT_HAS_REFR) ? C->mats[mat].refr[band] : 0.0f;
    if (fl < 0.0001 && fr < 0.0001) return OR_REFLECT;
    float fl2 = fl / (fl + fr);
    return (rng_next(r) <= fl2) ? OR_REFLECT : OR_REFRACT;
}

static void trace_ray(const rctx* C, acc_t* acc, uint32_t ray, int32_t* log_hits, uint32_t* log_nseg,
                      uint32_t* log_ncon, float* log_od) {
    const or_desc* D = C->d;
    const int band = D->band;
    const int mesh_src = D->src_kind == OR_SRC_MESH;
    or_rng R;
    rng_init(&R, D->seed, ray, D->ctx, D->instance);

    float I = 1.0f, L = 0.0f;
    float o[3], dd[3], prev[3] = {0, 0, 0}, n[3] = {0, 0, 0};
    int bt = OR_REFLECT, mat = -1;
    uint32_t nseg = 0, ncon = 0;

    for (uint32_t k = 0; k <= D->max_bounces; k++) {
        float spec = 0.0f;
        if (k == 0) {
            if (mesh_src) {
                emit_mesh(C, &R, o, dd);
            } else {
                o[0] = D->src_pos[0];
                o[1] = D->src_pos[1];
                o[2] = D->src_pos[2];
                sample_sphere(&R, dd);
            }
        } else {
            float t = 0.0f;  /* set by closest_hit_p when it hits */
            int32_t idx = -1;
            acc->segments++;
            int hit = closest_hit_p(C->s, o, dd, &t, &idx);
            if (log_hits && nseg < D->trace_segs) log_hits[nseg] = hit ? idx : -1;
            if (log_od && nseg < D->trace_segs) {
                float* q = log_od + 6 * (size_t)nseg;
                q[0] = o[0], q[1] = o[1], q[2] = o[2], q[3] = dd[0], q[4] = dd[1], q[5] = dd[2];
            }
            nseg++;
            if (D->estimator == OR_EST_SPHERE) {
                float u[3] = {dd[0], dd[1], dd[2]};
                normalize3(u);
                float seg = INFINITY;
                if (hit) {
                    float q[3], dq[3];
                    q[0] = o[0] + dd[0] * t;
                    q[1] = o[1] + dd[1] * t;
                    q[2] = o[2] + dd[2] * t;
                    sub3(q, o, dq);
                    seg = len3(dq);
                }
                sphere_deposits(C, acc, o, u, seg, I, L, k - 1);
            }
            if (!hit) break;
            float p[3];
            p[0] = o[0] + dd[0] * t;
            p[1] = o[1] + dd[1] * t;
            p[2] = o[2] + dd[2] * t;
            const float* N = C->s->nrm + 3 * idx;
            if (dot3(N, dd) > 0.0f) {
                n[0] = N[0] * -1.0f;
                n[1] = N[1] * -1.0f;
                n[2] = N[2] * -1.0f;
            } else {
                n[0] = N[0];
                n[1] = N[1];
                n[2] = N[2];
            }
            mat = C->s->mat[idx];
            bt = (int)material_bounce(C, &R, mat, band);
            spec = (C->mats[mat].flags & OR_MAT_HAS_SPEC) ? C->mats[mat].spec[band] : 0.0f;
            float v[3];
            if (bt == OR_REFRACT) {
                n[0] = -n[0];
                n[1] = -n[1];
                n[2] = -n[2];
                sample_hemi_spec(&R, n, dd, spec, v);
            } else {
                float rf[3];
                reflect3(dd, n, rf);
                sample_hemi_spec(&R, n, rf, spec, v);
            }
            float dist[3];
            sub3(p, o, dist);
            float l = len3(dist);
            acc->path_length += (double)l;
            I *= or_pow(D->air, l);
            L += l;
            o[0] = p[0];
            o[1] = p[1];
            o[2] = p[2];
            dd[0] = v[0];
            dd[1] = v[1];
            dd[2] = v[2];
            I *= C->ret[mat][band];
        }
        if (!is_normal_f(I)) break;
        if ((k > 0 || mesh_src) && D->estimator == OR_EST_NEE) {
            for (uint32_t r = 0; r < D->n_rec; r++) {
                const float* RP = D->rec[r].pos;
                float dv[3];
                sub3(RP, o, dv);
                acc->shadow_rays++;
                if (any_hit_p(C->s, o, dv)) continue;
                float lsdir[3] = {dv[0], dv[1], dv[2]};
                normalize3(lsdir);
                float dt = k ? dot3(lsdir, n) : 1.0f;
                if (!(dt > 0.0f)) continue;
                float c = I;
                if (k) {
                    float factor;
                    if (bt == OR_REFLECT) {
                        float rv[3];
                        reflect3(prev, n, rv);
                        float diff = -dot3(n, prev);
                        float sf = fmaxf(0.0f, dot3(rv, lsdir));
                        factor = spec * OR_EXP_INT * or_pow(sf, OR_EXP) + (1.0f - spec) * diff;
                    } else {
                        float diff = dot3(n, prev);
                        float sf = fmaxf(0.0f, dot3(prev, lsdir));
                        factor = spec * OR_EXP_INT * or_pow(sf, OR_EXP) + (1.0f - spec) * diff;
                    }
                    c *= factor;
                }
                float l = len3(dv);
                c *= or_pow(D->air, l);
                c *= 2.0f / (4.0f * OR_PI * l * l);
                if (is_normal_f(c)) {
                    if (k % 2) c *= -1.0f;
                    record(C, acc, (int)r, lsdir, c, (L + l) / 343.0f, L + l);
                    acc->contributions++;
                    ncon++;
                }
            }
        }
        if ((double)I < D->cutoff) break;
        prev[0] = dd[0];
        prev[1] = dd[1];
        prev[2] = dd[2];
        normalize3(prev);
    }
    if (log_nseg) *log_nseg = nseg;
    if (log_ncon) *log_ncon = ncon;
}

typedef struct {
    const rctx* C;
    or_out* out;
    uint32_t begin, end;
    acc_t acc;
    int ok;
} worker_t;

static void* worker(void* arg) {
    worker_t* W = (worker_t*)arg;
    const rctx* C = W->C;
    const or_desc* D = C->d;
    for (uint32_t ray = W->begin; ray < W->end; ray++) {
        uint32_t li = ray - D->ray_begin;
        int logit = W->out->trace_hits && li < D->trace_rays;
        int32_t* lh = logit ? W->out->trace_hits + (size_t)li * D->trace_segs : NULL;
        if (lh)
            for (uint32_t j = 0; j < D->trace_segs; j++) lh[j] = -2;
        trace_ray(C, &W->acc, ray, lh, (logit && W->out->trace_nseg) ? W->out->trace_nseg + li : NULL,
                  (logit && W->out->trace_ncon) ? W->out->trace_ncon + li : NULL,
                  (logit && W->out->trace_rays_od) ? W->out->trace_rays_od + (size_t)li * D->trace_segs * 6 : NULL);
    }
    W->ok = 1;
    return NULL;
}

int or_num_channels(const or_desc* d) {
    int n = 0;
    for (uint32_t r = 0; r < d->n_rec; r++) n += d->rec[r].kind == OR_REC_STEREO ? 2 : 1;
    return n;
}

double or_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int or_render(const or_scene* s, const or_desc* D, or_out* out) {
    if (D->band < 0 || D->band > 2 || D->n_bins == 0) return -1;
    if (D->ray_end < D->ray_begin || D->ray_end > D->n_rays) return -1;
    prep_scene ps;
    if (!prep(s, &ps)) return -2;
    int rc = 0;
    float (*ret)[3] = (float (*)[3])calloc(s->n_mat ? s->n_mat : 1, sizeof(float[3]));
    for (uint32_t m = 0; m < s->n_mat; m++)
        if (!or_material_retention(&s->mats[m], ret[m])) rc = -3;
    for (uint32_t i = 0; i < s->n_tris && rc == 0; i++)
        if (ps.mat[i] >= s->n_mat) rc = -4;
    int n_ch = or_num_channels(D);
    if (D->estimator > OR_EST_SPHERE) rc = -6;
    uint32_t nr_alloc = D->n_rec ? D->n_rec : 1;
    float* sph_r2 = (float*)calloc(nr_alloc, sizeof(float));
    float* sph_iv = (float*)calloc(nr_alloc, sizeof(float));
    float (*sph_cen)[2][3] = (float (*)[2][3])calloc(nr_alloc, sizeof(float[2][3]));
    for (uint32_t r = 0; r < D->n_rec && D->estimator == OR_EST_SPHERE; r++) {
        const or_recorder* R = &D->rec[r];
        float rad = R->radius;
        float rn = sqrtf(R->right_ear[0] * R->right_ear[0] + R->right_ear[1] * R->right_ear[1] +
                         R->right_ear[2] * R->right_ear[2]);
        if (!(rad > 0.0f) || (R->kind == OR_REC_STEREO && !(rn > 0.0f))) rc = -6;
        sph_r2[r] = rad * rad;
        sph_iv[r] = 3.0f / (4.0f * OR_PI * rad * rad * rad);
        for (int k = 0; k < 3; k++) sph_cen[r][0][k] = sph_cen[r][1][k] = R->pos[k];
        if (R->kind == OR_REC_STEREO && rn > 0.0f) {
            float h = (0.5f * R->head_size) / rn;
            for (int k = 0; k < 3; k++) {
                sph_cen[r][0][k] = R->pos[k] - R->right_ear[k] * h;
                sph_cen[r][1][k] = R->pos[k] + R->right_ear[k] * h;
            }
        }
    }
    int* ch_of = (int*)calloc(D->n_rec ? D->n_rec : 1, sizeof(int));
    {
        int c = 0;
        for (uint32_t r = 0; r < D->n_rec; r++) {
            ch_of[r] = c;
            c += D->rec[r].kind == OR_REC_STEREO ? 2 : 1;
        }
    }
    float* src_area = NULL;
    float* src_nrm = NULL;
    float src_total = 0.0f;
    if (D->src_kind == OR_SRC_MESH) {
        src_area = (float*)calloc(D->src_n_tris ? D->src_n_tris : 1, sizeof(float));
        src_nrm = (float*)calloc(3 * (D->src_n_tris ? D->src_n_tris : 1), sizeof(float));
        for (uint32_t i = 0; i < D->src_n_tris; i++) {
            src_area[i] = tri_area(D->src_v + 9 * (size_t)i);
            src_total += src_area[i];
            or_tri_normal(D->src_v + 9 * (size_t)i, src_nrm + 3 * i);
        }
        if (D->src_n_tris == 0) rc = -5;
    }
    if (rc != 0) {
        free(ret);
        free(ch_of);
        free(src_area);
        free(src_nrm);
        free(sph_r2);
        free(sph_iv);
        free(sph_cen);
        unprep(&ps);
        return rc;
    }

    rctx C;
    C.d = D;
    C.s = &ps;
    C.ret = (const float (*)[3])ret;
    C.mats = s->mats;
    C.src_v = D->src_v;
    C.src_area = src_area;
    C.src_total_area = src_total;
    C.src_nrm = src_nrm;
    C.n_ch = n_ch;
    C.ch_of_rec = ch_of;
    C.sph_r2 = sph_r2;
    C.sph_inv_vol = sph_iv;
    C.sph_cen = (const float (*)[2][3])sph_cen;

    uint32_t nt = D->n_threads ? D->n_threads : 1;
    uint32_t nr = D->ray_end - D->ray_begin;
    if (nt > nr && nr > 0) nt = nr;
    if (nt == 0) nt = 1;
    size_t hsz = (size_t)n_ch * D->n_bins;
    worker_t* W = (worker_t*)calloc(nt, sizeof(worker_t));
    pthread_t* th = (pthread_t*)calloc(nt, sizeof(pthread_t));
    for (uint32_t i = 0; i < nt; i++) {
        W[i].C = &C;
        W[i].out = out;
        W[i].begin = D->ray_begin + (uint32_t)(((uint64_t)nr * i) / nt);
        W[i].end = D->ray_begin + (uint32_t)(((uint64_t)nr * (i + 1)) / nt);
        W[i].acc.hs = out->hist_signed ? (double*)calloc(hsz, sizeof(double)) : NULL;
        W[i].acc.ha = out->hist_abs ? (double*)calloc(hsz, sizeof(double)) : NULL;
        W[i].acc.hf = (out->hist_f32 && nt == 1) ? (float*)calloc(hsz, sizeof(float)) : NULL;
        W[i].acc.tally = (double*)calloc(D->n_rec ? D->n_rec : 1, sizeof(double));
        W[i].acc.tally_abs = (double*)calloc(D->n_rec ? D->n_rec : 1, sizeof(double));
    }
    if (nt == 1) {
        worker(&W[0]);
    } else {
        for (uint32_t i = 0; i < nt; i++) pthread_create(&th[i], NULL, worker, &W[i]);
        for (uint32_t i = 0; i < nt; i++) pthread_join(th[i], NULL);
    }
    if (out->hist_signed) memset(out->hist_signed, 0, hsz * sizeof(double));
    if (out->hist_abs) memset(out->hist_abs, 0, hsz * sizeof(double));
    if (out->hist_f32) memset(out->hist_f32, 0, hsz * sizeof(float));
    out->segments = out->shadow_rays = out->contributions = 0;
    out->path_length = 0.0;
    for (uint32_t r = 0; r < D->n_rec; r++) {
        if (out->tally) out->tally[r] = 0.0;
        if (out->tally_abs) out->tally_abs[r] = 0.0;
    }
    for (uint32_t i = 0; i < nt; i++) {
        for (uint32_t r = 0; r < D->n_rec; r++) {
            if (out->tally) out->tally[r] += W[i].acc.tally[r];
            if (out->tally_abs) out->tally_abs[r] += W[i].acc.tally_abs[r];
        }
        out->path_length += W[i].acc.path_length;
        free(W[i].acc.tally);
        free(W[i].acc.tally_abs);
        for (size_t j = 0; j < hsz; j++) {
            if (W[i].acc.hs) out->hist_signed[j] += W[i].acc.hs[j];
            if (W[i].acc.ha) out->hist_abs[j] += W[i].acc.ha[j];
        }
        if (W[i].acc.hf) memcpy(out->hist_f32, W[i].acc.hf, hsz * sizeof(float));
        out->segments += W[i].acc.segments;
        out->shadow_rays += W[i].acc.shadow_rays;
        out->contributions += W[i].acc.contributions;
        free(W[i].acc.hs);
        free(W[i].acc.ha);
        free(W[i].acc.hf);
    }
    free(W);
    free(th);

    /* direct sound, src/Scene.cpp:299-311: connect listener -> source
     * (the sphere estimator collects the first segments instead) */
    if (D->src_kind == OR_SRC_POINT && D->estimator == OR_EST_NEE) {
        for (uint32_t r = 0; r < D->n_rec; r++) {
            const or_recorder* rec = &D->rec[r];
            int ch = ch_of[r];
            int nchr = rec->kind == OR_REC_STEREO ? 2 : 1;
            for (int c = 0; c < nchr; c++) {
                if (out->direct_visible) out->direct_visible[ch + c] = 0;
            }
            float dv[3];
            sub3(D->src_pos, rec->pos, dv);
            if (any_hit_p(&ps, rec->pos, dv)) continue;
            float dist[3];
            sub3(rec->pos, D->src_pos, dist);
            float len = len3(dist);
            float dir[3] = {dist[0], dist[1], dist[2]};
            normalize3(dir);
            float a = 1.0f / (4.0f * OR_PI * len * len) * or_pow(D->air, len) * D->dry;
            float t = len / 343.0f;
            float width = sqrtf(len);
            float ampl = 2.0f * a / width;
            int w = (int)ceilf(width);
            if (rec->kind == OR_REC_MONO) {
                if (out->direct_visible) {
                    out->direct_visible[ch] = 1;
                    out->direct_s[ch] = (int)((double)t * 44100.0);
                    out->direct_w[ch] = w;
                    out->direct_a[ch] = ampl;
                }
            } else {
                float dt = dot3(dir, rec->right_ear);
                float td = rec->head_size / 343.0f;
                float al = ampl, ar = ampl;
                float factor = or_pow(rec->head_abs[D->band], fabsf(dt) * rec->head_size);
                if (dt < 0.0f)
                    ar *= factor * factor;
                else
                    al *= factor * factor;
                if (out->direct_visible) {
                    out->direct_visible[ch] = 1;
                    out->direct_visible[ch + 1] = 1;
                    out->direct_s[ch] = (int)((double)(t + dt * td) * 44100.0);
                    out->direct_s[ch + 1] = (int)((double)(t - dt * td) * 44100.0);
                    out->direct_w[ch] = w;
                    out->direct_w[ch + 1] = w;
                    out->direct_a[ch] = al;
                    out->direct_a[ch + 1] = ar;
                }
            }
        }
    } else if (out->direct_visible) {
        for (int c = 0; c < n_ch; c++) out->direct_visible[c] = 0;
    }

    free(ret);
    free(ch_of);
    free(src_area);
    free(src_nrm);
    free(sph_r2);
    free(sph_iv);
    free(sph_cen);
    unprep(&ps);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 reference gradient of a render with respect to the per-material    */
/* reflection coefficient and specularity of the rendered band.            */
/*                                                                          */
/* The primal path is exactly trace_ray's (binary32, same Philox stream).  */
/* Alongside it, forward-mode tangents in binary64 follow each specularity */
/* parameter through the reparameterised scattering direction             */
/* v = normalize(h (1 - s) + r s) (Sample_Hemi, src/Distributions.h:71-75): */
/* hemisphere samples h and the hit triangles are held fixed, so the        */
/* tangent of a hit point follows the hit triangle's plane.  The reflection */
/* coefficient enters through the retention factor (src/Material.cpp:59)   */
/* and d retention / d refl = 1.  Each recorded contribution c contributes  */
/* (dL/dc) (dc/dtheta), dL/dc = sum over its ramp bins of adj[bin] times   */
/* the ramp weight (w - i)/w * 2/sqrt(d) (and ILD gains, stereo);          */
/* adj_abs weighs |c| (sign(c) dc/dtheta).  Time-of-arrival shifts of a     */
/* contribution are not differentiated (binning is piecewise constant).     */
/* ------------------------------------------------------------------------ */
typedef struct {
    double o[3], d[3], dL, dlogI;
} tan_t;

typedef struct {
    const rctx* C;
    const double* adj_s; /* [n_ch][n_bins] or NULL */
    const double* adj_a;
    uint32_t begin, end;
    uint32_t n_mat;
    int local;      /* 1: direction tangents are not carried past their vertex; 2: score form */
    float q_spawn;  /* score form: continuations per vertex per unit rim flux F */
    double* g_refl; /* [n_mat] */
    double* g_spec;
} gworker_t;

static double ramp_adj(const double* adj, uint32_t nb, int s, int w, double A_per_c) {
    double acc = 0.0;
    for (int i = 0; i < w; i++) {
        long long b = (long long)s + i;
        if (b < 0 || b >= (long long)nb) continue;
        acc += adj[b] * ((double)(w - i) / (double)w) * A_per_c;
    }
    return acc;
}

/* dL/dc for a contribution c of recorder r (signed and |c| adjoints) */
static double contrib_adjoint(const gworker_t* G, int r, const float dir[3], float c, float t, float dist) {
    const rctx* C = G->C;
    const or_recorder* rec = &C->d->rec[r];
    int ch = C->ch_of_rec[r];
    uint32_t nb = C->d->n_bins;
    float width = sqrtf(dist);
    int w = (int)ceilf(width);
    double per_c = 2.0 / (double)width; /* A = 2 c / sqrt(d) */
    double sgn = c < 0.0f ? -1.0 : 1.0;
    double a = 0.0;
    if (rec->kind == OR_REC_MONO) {
        int s = (int)((double)t * 44100.0);
        if (G->adj_s) a += ramp_adj(G->adj_s + (size_t)ch * nb, nb, s, w, per_c);
        if (G->adj_a) a += sgn * ramp_adj(G->adj_a + (size_t)ch * nb, nb, s, w, per_c);
    } else {
        float dt = dot3(dir, rec->right_ear);
        float td = rec->head_size / 343.0f;
        int s_right = (int)((double)(t - dt * td) * 44100.0);
        int s_left = (int)((double)(t + dt * td) * 44100.0);
        float factor = or_pow(rec->head_abs[C->d->band], fabsf(dt) * rec->head_size);
        double gl = 1.0, gr = 1.0;
        if (dt < 0.0f)
            gr = (double)(factor * factor);
        else
            gl = (double)(factor * factor);
        if (G->adj_s) {
            a += gl * ramp_adj(G->adj_s + (size_t)ch * nb, nb, s_left, w, per_c);
            a += gr * ramp_adj(G->adj_s + (size_t)(ch + 1) * nb, nb, s_right, w, per_c);
        }
        if (G->adj_a) {
            a += sgn * gl * ramp_adj(G->adj_a + (size_t)ch * nb, nb, s_left, w, per_c);
            a += sgn * gr * ramp_adj(G->adj_a + (size_t)(ch + 1) * nb, nb, s_right, w, per_c);
        }
    }
    return a;
}

static inline double ddot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double dfdot(const double a[3], const float b[3]) {
    return a[0] * (double)b[0] + a[1] * (double)b[1] + a[2] * (double)b[2];
}

static void grad_ray(const gworker_t* G, uint32_t ray, tan_t* T, double* dlogI_refl) {
    const rctx* C = G->C;
    const or_desc* D = C->d;
    const int band = D->band;
    const uint32_t nm = G->n_mat;
    const double ln_air = log((double)D->air);
    or_rng R;
    rng_init(&R, D->seed, ray, D->ctx, D->instance);
    float I = 1.0f, L = 0.0f;
    float o[3], dd[3], prev[3] = {0, 0, 0}, n[3] = {0, 0, 0};
    int bt = OR_REFLECT, mat = -1;
    double dprev[16][3];
    for (uint32_t q = 0; q < nm; q++) {
        memset(&T[q], 0, sizeof(tan_t));
        dlogI_refl[q] = 0.0;
    }
    for (uint32_t k = 0; k <= D->max_bounces; k++) {
        float spec = 0.0f;
        float dnorm = 1.0f;
        if (k == 0) {
            if (D->src_kind == OR_SRC_MESH) {
                emit_mesh(C, &R, o, dd);
            } else {
                o[0] = D->src_pos[0];
                o[1] = D->src_pos[1];
                o[2] = D->src_pos[2];
                sample_sphere(&R, dd);
            }
        } else {
            float t;
            int32_t idx;
            int hit = closest_hit_p(C->s, o, dd, &t, &idx);
            if (!hit) break;
            float p[3] = {o[0] + dd[0] * t, o[1] + dd[1] * t, o[2] + dd[2] * t};
            const float* N = C->s->nrm + 3 * idx;
            if (dot3(N, dd) > 0.0f) {
                n[0] = N[0] * -1.0f;
                n[1] = N[1] * -1.0f;
                n[2] = N[2] * -1.0f;
            } else {
                n[0] = N[0];
                n[1] = N[1];
                n[2] = N[2];
            }
            mat = C->s->mat[idx];
            bt = (int)material_bounce(C, &R, mat, band);
            spec = (C->mats[mat].flags & OR_MAT_HAS_SPEC) ? C->mats[mat].spec[band] : 0.0f;
            /* hemisphere sample h (rejection against the final normal) and the blend */
            float nn[3] = {n[0], n[1], n[2]};
            if (bt == OR_REFRACT) {
                nn[0] = -n[0];
                nn[1] = -n[1];
                nn[2] = -n[2];
            }
            float h[3], rdir[3];
            sample_hemi(&R, nn, h);
            if (bt == OR_REFRACT)
                memcpy(rdir, dd, sizeof(rdir));
            else
                reflect3(dd, nn, rdir);
            float g = 1.0f - spec;
            float wv[3] = {h[0] * g + rdir[0] * spec, h[1] * g + rdir[1] * spec, h[2] * g + rdir[2] * spec};
            float wl = len3(wv);
            float v[3] = {wv[0], wv[1], wv[2]};
            normalize3(v);
            /* tangents of the hit point (plane of the hit triangle) and the new direction */
            double Nd = dfdot((const double[3]){N[0], N[1], N[2]}, dd);
            float dist[3];
            sub3(p, o, dist);
            float l = len3(dist);
            for (uint32_t q = 0; q < nm; q++) {
                tan_t* tq = &T[q];
                double Nv[3] = {N[0], N[1], N[2]};
                double dt_ = -(ddot(Nv, tq->o) + (double)t * ddot(Nv, tq->d)) / Nd;
                double dp[3];
                for (int a = 0; a < 3; a++) dp[a] = tq->o[a] + tq->d[a] * (double)t + (double)dd[a] * dt_;
                double dl = 0.0;
                for (int a = 0; a < 3; a++) dl += (double)dist[a] * (dp[a] - tq->o[a]);
                dl /= (double)l;
                tq->dL += dl;
                tq->dlogI += ln_air * dl;
                /* prev = d / |d| */
                double dn = (double)len3(dd);
                double pd = 0.0;
                for (int a = 0; a < 3; a++) pd += (double)dd[a] / dn * tq->d[a];
                for (int a = 0; a < 3; a++) dprev[q][a] = (tq->d[a] - (double)dd[a] / dn * pd) / dn;
                /* new direction */
                double dr[3];
                if (bt == OR_REFRACT) {
                    for (int a = 0; a < 3; a++) dr[a] = tq->d[a];
                } else {
                    double dn2 = dfdot(tq->d, nn);
                    for (int a = 0; a < 3; a++) dr[a] = tq->d[a] - 2.0 * dn2 * (double)nn[a];
                }
                double dw[3];
                for (int a = 0; a < 3; a++) {
                    dw[a] = dr[a] * (double)spec;
                    if ((int)q == mat) dw[a] += (double)rdir[a] - (double)h[a];
                }
                double vdw = dfdot(dw, v);
                for (int a = 0; a < 3; a++) {
                    tq->o[a] = dp[a];
                    tq->d[a] = G->local ? 0.0 : (dw[a] - (double)v[a] * vdw) / (double)wl;
                }
            }
            (void)dnorm;
            dlogI_refl[mat] += 1.0 / (double)C->ret[mat][band];
            I *= or_pow(D->air, l);
            L += l;
            o[0] = p[0];
            o[1] = p[1];
            o[2] = p[2];
            dd[0] = v[0];
            dd[1] = v[1];
            dd[2] = v[2];
            I *= C->ret[mat][band];
            n[0] = nn[0];
            n[1] = nn[1];
            n[2] = nn[2];
        }
        if (!is_normal_f(I)) break;
        if (k > 0 || D->src_kind == OR_SRC_MESH) {
            for (uint32_t r = 0; r < D->n_rec; r++) {
                const float* RP = D->rec[r].pos;
                float dv[3];
                sub3(RP, o, dv);
                if (any_hit_p(C->s, o, dv)) continue;
                float lsdir[3] = {dv[0], dv[1], dv[2]};
                normalize3(lsdir);
                float dtt = k ? dot3(lsdir, n) : 1.0f;
                if (!(dtt > 0.0f)) continue;
                float c = I, factor = 1.0f, lobe = 0.0f, diff = 0.0f, sf = 0.0f;
                float rv[3] = {0, 0, 0};
                if (k) {
                    if (bt == OR_REFLECT) {
                        reflect3(prev, n, rv);
                        diff = -dot3(n, prev);
                        sf = fmaxf(0.0f, dot3(rv, lsdir));
                    } else {
                        diff = dot3(n, prev);
                        sf = fmaxf(0.0f, dot3(prev, lsdir));
                    }
                    lobe = OR_EXP_INT * or_pow(sf, OR_EXP);
                    factor = spec * OR_EXP_INT * or_pow(sf, OR_EXP) + (1.0f - spec) * diff;
                    c *= factor;
                }
                float l = len3(dv);
                float geo = or_pow(D->air, l) * (2.0f / (4.0f * OR_PI * l * l));
                c *= or_pow(D->air, l);
                c *= 2.0f / (4.0f * OR_PI * l * l);
                if (!is_normal_f(c)) continue;
                if (k % 2) c *= -1.0f;
                double a_c = contrib_adjoint(G, (int)r, lsdir, c, (L + l) / 343.0f, L + l);
                if (a_c == 0.0) continue;
                double sgn_k = (k % 2) ? -1.0 : 1.0;
                /* c = sgn_k * I * factor * geo(l) */
                double cnof = sgn_k * (double)I * (double)geo; /* c without the factor */
                for (uint32_t q = 0; q < nm; q++) {
                    const tan_t* tq = &T[q];
                    double dvv[3] = {-tq->o[0], -tq->o[1], -tq->o[2]};
                    double dls = (dvv[0] * dv[0] + dvv[1] * dv[1] + dvv[2] * dv[2]) / (double)l;
                    double dlsdir[3];
                    double ld = 0.0;
                    for (int a = 0; a < 3; a++) ld += (double)lsdir[a] * dvv[a];
                    for (int a = 0; a < 3; a++) dlsdir[a] = (dvv[a] - (double)lsdir[a] * ld) / (double)l;
                    double dfactor = 0.0;
                    if (k) {
                        double ddiff, dsf = 0.0;
                        if (bt == OR_REFLECT) {
                            double drv[3];
                            double pn = ddot(dprev[q], (const double[3]){n[0], n[1], n[2]});
                            for (int a = 0; a < 3; a++) drv[a] = dprev[q][a] - 2.0 * pn * (double)n[a];
                            ddiff = -pn;
                            if (sf > 0.0f) dsf = dfdot(drv, lsdir) + dfdot(dlsdir, rv);
                        } else {
                            ddiff = ddot(dprev[q], (const double[3]){n[0], n[1], n[2]});
                            if (sf > 0.0f) dsf = dfdot(dprev[q], lsdir) + dfdot(dlsdir, prev);
                        }
                        double dlobe = sf > 0.0f ? (double)OR_EXP_INT * 1000.0 * pow((double)sf, 999.0) * dsf : 0.0;
                        dfactor = (double)spec * dlobe + (1.0 - (double)spec) * ddiff;
                        if ((int)q == mat) dfactor += (double)lobe - (double)diff;
                    }
                    double dgeo_rel = log((double)D->air) * dls - 2.0 * dls / (double)l;
                    double dc = cnof * dfactor + (double)c * (tq->dlogI + dgeo_rel);
                    G->g_spec[q] += a_c * dc;
                }
                for (uint32_t m = 0; m < nm; m++) G->g_refl[m] += a_c * (double)c * dlogI_refl[m];
            }
        }
        if ((double)I < D->cutoff) break;
        prev[0] = dd[0];
        prev[1] = dd[1];
        prev[2] = dd[2];
        normalize3(prev);
    }
}

/* ------------------------------------------------------------------------ */
/* The score form of the specularity gradient (local == 2): unbiased at any */
/* path length.                                                             */
/*                                                                          */
/* Sample_Hemi (src/Distributions.h:71-75) draws h uniformly on the         */
/* hemisphere of nu and returns v = normalize((1-s) h + s r), r.nu > 0.    */
/* For s < 1/2 the map is one-to-one and v has the density                  */
/*   p(v) = t^2 / (2 pi (1-s) sqrt(D)),  c = v.r,  D = (1-s)^2 - s^2 (1-c^2),*/
/*   t = s c + sqrt(D)                                                      */
/* on the image Omega(s) of the hemisphere.  With G(v) the expected         */
/* downstream value of a path leaving a vertex in direction v,              */
/*   d/ds int_Omega(s) G p = E[G(v) dlog p/ds] - int_0^2pi G(v_b(phi)) w(phi)*/
/* where v_b(phi) = normalize((1-s) e(phi) + s r) is the image of the       */
/* hemisphere's rim e(phi) and w(phi) = (r.nu) / (2 pi (1-s)((1-s) + s e.r))*/
/* the rate at which probability leaves through it; its integral is         */
/*   F = (r.nu) / ((1-s) sqrt((1-s)^2 - s^2 (1 - (r.nu)^2))) = E[dlog p/ds].*/
/* The rim integral is estimated by a continuation path: with probability  */
/* q = min(1, rate F) at a vertex (proportional to the rim term's weight,  */
/* which minimises the variance for a given number of continuations), a    */
/* second path leaves the vertex in direction v_b(phi), phi drawn with     */
/* density w / F (a wrapped                                                 */
/* Cauchy law, drawn as the Moebius image of a uniform point of the circle), */
/* and adds -(F / q) times its connections to the gradient.  The main path  */
/* adds G (dlog p/ds - F + [spawned] F / q): the last two terms have mean   */
/* zero and cancel most of the continuation's weight.  Summed over the      */
/* vertices of the material, each connection of the main path adds its     */
/* dL/dc times S_m, the running sum of those coefficients, plus the         */
/* explicit dependence of its lobe on s.  Nothing of a path moves, so       */
/* wall switches, visibility and arrival times need no derivative.          */
/* ------------------------------------------------------------------------ */
static double score_dlogp(double c, double s) {
    double D = (1.0 - s) * (1.0 - s) - s * s * (1.0 - c * c);
    double sq = sqrt(D);
    double Ds = -2.0 * (1.0 - s) - 2.0 * s * (1.0 - c * c);
    double t = s * c + sq;
    double ts = c + Ds / (2.0 * sq);
    return 2.0 * ts / t + 1.0 / (1.0 - s) - Ds / (2.0 * D);
}

static double score_flux(double cr, double s) {
    double sin2 = 1.0 - cr * cr;
    if (sin2 < 0.0) sin2 = 0.0;
    return cr / ((1.0 - s) * sqrt((1.0 - s) * (1.0 - s) - s * s * sin2));
}

/* the stream of the spawn decision and the continuation of vertex k */
static void score_rng_init(or_rng* r, uint64_t seed, uint32_t ray, uint32_t ctx, uint32_t instance, uint32_t k) {
    rng_init(r, seed, ray, ctx, instance);
    r->key[0] ^= 0x85EBCA6Bu;
    r->key[1] ^= 0xC2B2AE35u * (k + 1u);
}

/* the rim direction v_b (phi with density w / F) for hemisphere normal nu,
 * blend target r and specularity s */
static void score_rim_dir(or_rng* R, const float nu[3], const float r[3], float s, float out[3]) {
    float zx, zy;
    for (;;) {
        float x = rng_next(R) * 2.0f - 1.0f;
        float y = rng_next(R) * 2.0f - 1.0f;
        float l = x * x + y * y;
        if (l < 0.001f || l > 1.0f) continue;
        float sl = sqrtf(l);
        zx = x / sl;
        zy = y / sl;
        break;
    }
    float cr = dot3(r, nu);
    float rp[3] = {r[0] - nu[0] * cr, r[1] - nu[1] * cr, r[2] - nu[2] * cr};
    float sr = len3(rp);
    float a[3];
    if (sr > 1e-6f) {
        a[0] = rp[0] / sr;
        a[1] = rp[1] / sr;
        a[2] = rp[2] / sr;
    } else {
        /* r along nu: any direction of the rim plane (the law is uniform) */
        float ax = fabsf(nu[0]), ay = fabsf(nu[1]), az = fabsf(nu[2]);
        float t[3] = {0.0f, 0.0f, 0.0f};
        if (ax <= ay && ax <= az) t[0] = 1.0f;
        else if (ay <= az) t[1] = 1.0f;
        else t[2] = 1.0f;
        float td = dot3(t, nu);
        a[0] = t[0] - nu[0] * td;
        a[1] = t[1] - nu[1] * td;
        a[2] = t[2] - nu[2] * td;
        normalize3(a);
        sr = 0.0f;
    }
    float b[3];
    cross3(nu, a, b);
    float A = 1.0f - s, B = s * sr;
    float rho = B / (A + sqrtf(A * A - B * B));
    /* w = (z + rho) / (1 + rho z): arg w is wrapped Cauchy about 0 */
    float wx = (zx + rho) * (1.0f + rho * zx) + rho * zy * zy;
    float wy = zy * (1.0f - rho * rho);
    float wl = sqrtf(wx * wx + wy * wy);
    wx = wx / wl;
    wy = wy / wl;
    /* phi = pi + arg w: the rim point e = -wx a - wy b */
    float e[3] = {-wx * a[0] - wy * b[0], -wx * a[1] - wy * b[1], -wx * a[2] - wy * b[2]};
    float g = 1.0f - s;
    out[0] = e[0] * g + r[0] * s;
    out[1] = e[1] * g + r[1] * s;
    out[2] = e[2] * g + r[2] * s;
    normalize3(out);
}

double or_score_dlogp(double c, double s) { return score_dlogp(c, s); }
double or_score_flux(double cr, double s) { return score_flux(cr, s); }
/* n rim directions for hemisphere normal nu, blend target r, specularity s,
 * from the stream of vertex 1 of rays 0..n-1 (after its spawn draw) */
void or_score_rim_dirs(uint64_t seed, uint32_t n, const float nu[3], const float r[3], float s, float* out) {
    for (uint32_t i = 0; i < n; i++) {
        or_rng R;
        score_rng_init(&R, seed, i, 0u, 0u, 1u);
        (void)rng_next(&R);
        score_rim_dir(&R, nu, r, s, out + 3 * (size_t)i);
    }
}

/* the continuation of a rim direction: trace_ray's transport from vertex
 * k0 (point o, direction d, energy I, length L) with the vertex's own
 * stream; each connection adds weight * dL/dc * c to g_spec[m] */
static void score_continue(const gworker_t* G, or_rng* R, uint32_t k0, const float o_in[3], const float d_in[3],
                           float I, float L, double weight, int m) {
    const rctx* C = G->C;
    const or_desc* D = C->d;
    const int band = D->band;
    float o[3] = {o_in[0], o_in[1], o_in[2]}, dd[3] = {d_in[0], d_in[1], d_in[2]};
    float prev[3] = {dd[0], dd[1], dd[2]}, n[3];
    normalize3(prev);
    for (uint32_t k = k0 + 1; k <= D->max_bounces; k++) {
        float t;
        int32_t idx;
        if (!closest_hit_p(C->s, o, dd, &t, &idx)) break;
        float p[3] = {o[0] + dd[0] * t, o[1] + dd[1] * t, o[2] + dd[2] * t};
        const float* N = C->s->nrm + 3 * idx;
        if (dot3(N, dd) > 0.0f) {
            n[0] = N[0] * -1.0f;
            n[1] = N[1] * -1.0f;
            n[2] = N[2] * -1.0f;
        } else {
            n[0] = N[0];
            n[1] = N[1];
            n[2] = N[2];
        }
        int mat = C->s->mat[idx];
        int bt = (int)material_bounce(C, R, mat, band);
        float spec = (C->mats[mat].flags & OR_MAT_HAS_SPEC) ? C->mats[mat].spec[band] : 0.0f;
        float v[3];
        if (bt == OR_REFRACT) {
            n[0] = -n[0];
            n[1] = -n[1];
            n[2] = -n[2];
            sample_hemi_spec(R, n, dd, spec, v);
        } else {
            float rf[3];
            reflect3(dd, n, rf);
            sample_hemi_spec(R, n, rf, spec, v);
        }
        float dist[3];
        sub3(p, o, dist);
        float l = len3(dist);
        I *= or_pow(D->air, l);
        L += l;
        o[0] = p[0];
        o[1] = p[1];
        o[2] = p[2];
        dd[0] = v[0];
        dd[1] = v[1];
        dd[2] = v[2];
        I *= C->ret[mat][band];
        if (!is_normal_f(I)) break;
        for (uint32_t r = 0; r < D->n_rec; r++) {
            const float* RP = D->rec[r].pos;
            float dv[3];
            sub3(RP, o, dv);
            if (any_hit_p(C->s, o, dv)) continue;
            float lsdir[3] = {dv[0], dv[1], dv[2]};
            normalize3(lsdir);
            if (!(dot3(lsdir, n) > 0.0f)) continue;
            float diff, sf;
            if (bt == OR_REFLECT) {
                float rv[3];
                reflect3(prev, n, rv);
                diff = -dot3(n, prev);
                sf = fmaxf(0.0f, dot3(rv, lsdir));
            } else {
                diff = dot3(n, prev);
                sf = fmaxf(0.0f, dot3(prev, lsdir));
            }
            float c = I;
            c *= spec * OR_EXP_INT * or_pow(sf, OR_EXP) + (1.0f - spec) * diff;
            float ls = len3(dv);
            c *= or_pow(D->air, ls);
            c *= 2.0f / (4.0f * OR_PI * ls * ls);
            if (!is_normal_f(c)) continue;
            if (k % 2) c *= -1.0f;
            double a_c = contrib_adjoint(G, (int)r, lsdir, c, (L + ls) / 343.0f, L + ls);
            G->g_spec[m] += weight * a_c * (double)c;
        }
        if ((double)I < D->cutoff) break;
        prev[0] = dd[0];
        prev[1] = dd[1];
        prev[2] = dd[2];
        normalize3(prev);
    }
}

static void grad_ray_score(const gworker_t* G, uint32_t ray, double* S, double* dlogI_refl) {
    const rctx* C = G->C;
    const or_desc* D = C->d;
    const int band = D->band;
    const uint32_t nm = G->n_mat;
    or_rng R;
    rng_init(&R, D->seed, ray, D->ctx, D->instance);
    float I = 1.0f, L = 0.0f;
    float o[3], dd[3], prev[3] = {0, 0, 0}, n[3] = {0, 0, 0};
    int bt = OR_REFLECT, mat = -1;
    for (uint32_t q = 0; q < nm; q++) S[q] = dlogI_refl[q] = 0.0;
    for (uint32_t k = 0; k <= D->max_bounces; k++) {
        float spec = 0.0f;
        float h_nu[3] = {0, 0, 0}, h_r[3] = {0, 0, 0};
        if (k == 0) {
            o[0] = D->src_pos[0];
            o[1] = D->src_pos[1];
            o[2] = D->src_pos[2];
            sample_sphere(&R, dd);
        } else {
            float t;
            int32_t idx;
            if (!closest_hit_p(C->s, o, dd, &t, &idx)) break;
            float p[3] = {o[0] + dd[0] * t, o[1] + dd[1] * t, o[2] + dd[2] * t};
            const float* N = C->s->nrm + 3 * idx;
            if (dot3(N, dd) > 0.0f) {
                n[0] = N[0] * -1.0f;
                n[1] = N[1] * -1.0f;
                n[2] = N[2] * -1.0f;
            } else {
                n[0] = N[0];
                n[1] = N[1];
                n[2] = N[2];
            }
            mat = C->s->mat[idx];
            bt = (int)material_bounce(C, &R, mat, band);
            spec = (C->mats[mat].flags & OR_MAT_HAS_SPEC) ? C->mats[mat].spec[band] : 0.0f;
            float v[3];
            if (bt == OR_REFRACT) {
                n[0] = -n[0];
                n[1] = -n[1];
                n[2] = -n[2];
                memcpy(h_r, dd, sizeof(h_r));
            } else {
                reflect3(dd, n, h_r);
            }
            memcpy(h_nu, n, sizeof(h_nu));
            sample_hemi_spec(&R, n, h_r, spec, v);
            float dist[3];
            sub3(p, o, dist);
            float l = len3(dist);
            dlogI_refl[mat] += 1.0 / (double)C->ret[mat][band];
            I *= or_pow(D->air, l);
            L += l;
            o[0] = p[0];
            o[1] = p[1];
            o[2] = p[2];
            dd[0] = v[0];
            dd[1] = v[1];
            dd[2] = v[2];
            I *= C->ret[mat][band];
        }
        if (!is_normal_f(I)) break;
        if (k > 0) {
            for (uint32_t r = 0; r < D->n_rec; r++) {
                const float* RP = D->rec[r].pos;
                float dv[3];
                sub3(RP, o, dv);
                if (any_hit_p(C->s, o, dv)) continue;
                float lsdir[3] = {dv[0], dv[1], dv[2]};
                normalize3(lsdir);
                if (!(dot3(lsdir, n) > 0.0f)) continue;
                float diff, sf;
                if (bt == OR_REFLECT) {
                    float rv[3];
                    reflect3(prev, n, rv);
                    diff = -dot3(n, prev);
                    sf = fmaxf(0.0f, dot3(rv, lsdir));
                } else {
                    diff = dot3(n, prev);
                    sf = fmaxf(0.0f, dot3(prev, lsdir));
                }
                float lobe = OR_EXP_INT * or_pow(sf, OR_EXP);
                float c = I;
                c *= spec * OR_EXP_INT * or_pow(sf, OR_EXP) + (1.0f - spec) * diff;
                float ls = len3(dv);
                float geo = or_pow(D->air, ls) * (2.0f / (4.0f * OR_PI * ls * ls));
                c *= or_pow(D->air, ls);
                c *= 2.0f / (4.0f * OR_PI * ls * ls);
                if (!is_normal_f(c)) continue;
                if (k % 2) c *= -1.0f;
                double a_c = contrib_adjoint(G, (int)r, lsdir, c, (L + ls) / 343.0f, L + ls);
                if (a_c == 0.0) continue;
                double cnof = ((k % 2) ? -1.0 : 1.0) * (double)I * (double)geo;
                for (uint32_t q = 0; q < nm; q++) {
                    G->g_refl[q] += a_c * (double)c * dlogI_refl[q];
                    G->g_spec[q] += a_c * (double)c * S[q];
                }
                G->g_spec[mat] += a_c * cnof * ((double)lobe - (double)diff);
            }
        }
        if ((double)I < D->cutoff) break;
        if (k > 0 && k < D->max_bounces) {
            /* the direction just drawn at this vertex leaves it: its score,
             * the rim flux and perhaps a continuation */
            double cdir = (double)dd[0] * (double)h_r[0] + (double)dd[1] * (double)h_r[1] +
                          (double)dd[2] * (double)h_r[2];
            double cr = (double)h_r[0] * (double)h_nu[0] + (double)h_r[1] * (double)h_nu[1] +
                        (double)h_r[2] * (double)h_nu[2];
            double F = score_flux(cr, (double)spec);
            double coef = score_dlogp(cdir, (double)spec) - F;
            or_rng Rs;
            score_rng_init(&Rs, D->seed, ray, D->ctx, D->instance, k);
            const float qk = fminf(1.0f, (float)(F * (double)G->q_spawn));
            if (rng_next(&Rs) < qk) {
                float vb[3];
                score_rim_dir(&Rs, h_nu, h_r, spec, vb);
                double wgt = F / (double)qk;
                coef += wgt;
                score_continue(G, &Rs, k, o, vb, I, L, -wgt, mat);
            }
            S[mat] += coef;
        }
        prev[0] = dd[0];
        prev[1] = dd[1];
        prev[2] = dd[2];
        normalize3(prev);
    }
}

static void* gworker(void* arg) {
    gworker_t* G = (gworker_t*)arg;
    tan_t T[16];
    double dr[16];
    double S[16];
    for (uint32_t ray = G->begin; ray < G->end; ray++) {
        if (G->local == 2)
            grad_ray_score(G, ray, S, dr);
        else
            grad_ray(G, ray, T, dr);
    }
    return NULL;
}

int or_gradient(const or_scene* s, const or_desc* D, const double* adj_signed, const double* adj_abs, int local,
                double* grad_refl, double* grad_spec) {
    return or_gradient_k(s, D, adj_signed, adj_abs, local, OR_SCORE_SPAWN, grad_refl, grad_spec);
}

int or_gradient_k(const or_scene* s, const or_desc* D, const double* adj_signed, const double* adj_abs, int local,
                  float q_spawn, double* grad_refl, double* grad_spec) {
    if (s->n_mat > 16 || D->band < 0 || D->band > 2) return -1;
    if (local == 2) {
        if (D->src_kind != OR_SRC_POINT || !(q_spawn > 0.0f)) return -1;
        for (uint32_t m = 0; m < s->n_mat; m++)
            if ((s->mats[m].flags & OR_MAT_HAS_SPEC) && !(s->mats[m].spec[D->band] < 0.5f)) return -3;
    }
    prep_scene ps;
    if (!prep(s, &ps)) return -2;
    float (*ret)[3] = (float (*)[3])calloc(s->n_mat ? s->n_mat : 1, sizeof(float[3]));
    for (uint32_t m = 0; m < s->n_mat; m++) or_material_retention(&s->mats[m], ret[m]);
    int* ch_of = (int*)calloc(D->n_rec ? D->n_rec : 1, sizeof(int));
    int c = 0;
    for (uint32_t r = 0; r < D->n_rec; r++) {
        ch_of[r] = c;
        c += D->rec[r].kind == OR_REC_STEREO ? 2 : 1;
    }
    rctx C;
    memset(&C, 0, sizeof(C));
    C.d = D;
    C.s = &ps;
    C.ret = (const float (*)[3])ret;
    C.mats = s->mats;
    C.n_ch = c;
    C.ch_of_rec = ch_of;
    uint32_t nt = D->n_threads ? D->n_threads : 1;
    uint32_t nr = D->ray_end - D->ray_begin;
    if (nt > nr && nr > 0) nt = nr;
    gworker_t* W = (gworker_t*)calloc(nt, sizeof(gworker_t));
    pthread_t* th = (pthread_t*)calloc(nt, sizeof(pthread_t));
    for (uint32_t i = 0; i < nt; i++) {
        W[i].C = &C;
        W[i].adj_s = adj_signed;
        W[i].adj_a = adj_abs;
        W[i].n_mat = s->n_mat;
        W[i].local = local;
        W[i].q_spawn = q_spawn;
        W[i].begin = D->ray_begin + (uint32_t)(((uint64_t)nr * i) / nt);
        W[i].end = D->ray_begin + (uint32_t)(((uint64_t)nr * (i + 1)) / nt);
        W[i].g_refl = (double*)calloc(16, sizeof(double));
        W[i].g_spec = (double*)calloc(16, sizeof(double));
        pthread_create(&th[i], NULL, gworker, &W[i]);
    }
    for (uint32_t m = 0; m < s->n_mat; m++) grad_refl[m] = grad_spec[m] = 0.0;
    for (uint32_t i = 0; i < nt; i++) {
        pthread_join(th[i], NULL);
        for (uint32_t m = 0; m < s->n_mat; m++) {
            grad_refl[m] += W[i].g_refl[m];
            grad_spec[m] += W[i].g_spec[m];
        }
        free(W[i].g_refl);
        free(W[i].g_spec);
    }
    free(W);
    free(th);
    free(ret);
    free(ch_of);
    unprep(&ps);
    return 0;
}

/* ======================================================================
 * Auralization: Equalizer::Pass / Split (lib/equalizer/Equalizer.cpp:27-95).
 * Single precision, each operation rounded in the order the reference
 * writes it; tanf because the reference's C++ tan(float) is the float
 * overload.  The coefficient expressions are the expansion of the squared
 * Butterworth quadratic under the bilinear transform (checked in double
 * against scipy's butter in the tests).
 * ====================================================================== */
void or_lr4_coefficients(float fc, int high_pass, float a[5], float b[5]) {
    const float srate = 44100.0f, pi = 3.14285714285714f;
    float wc = 2.0f * pi * srate;
    float wc2 = wc * wc, wc3 = wc2 * wc, wc4 = wc2 * wc2;
    float k = wc / tanf(pi * fc / srate);
    float k2 = k * k, k3 = k2 * k, k4 = k2 * k2;
    float r2 = sqrtf(2.0f);
    float t1 = r2 * wc3 * k, t2 = r2 * wc * k3;
    float den = 4.0f * wc2 * k2 + 2.0f * t1 + k4 + 2.0f * t2 + wc4;
    b[0] = 1.0f;
    b[1] = (4.0f * (wc4 + t1 - k4 - t2)) / den;
    b[2] = (6.0f * wc4 - 8.0f * wc2 * k2 + 6.0f * k4) / den;
    b[3] = (4.0f * (wc4 - t1 + t2 - k4)) / den;
    b[4] = (k4 - 2.0f * t1 + wc4 - 2.0f * t2 + 4.0f * wc2 * k2) / den;
    float g = high_pass ? k4 : wc4;
    a[0] = g / den;
    a[1] = (high_pass ? -4.0f : 4.0f) * g / den;
    a[2] = 6.0f * g / den;
    a[3] = a[1];
    a[4] = a[0];
}

typedef struct {
    float a[5], b[5], x[4], y[4];
} or_lr4;

static void lr4_init(or_lr4* f, float fc, int hp) {
    memset(f, 0, sizeof(*f));
    or_lr4_coefficients(fc, hp, f->a, f->b);
}

static float lr4_step(or_lr4* f, float in) {
    float out = f->a[0] * in + f->a[1] * f->x[0] + f->a[2] * f->x[1] + f->a[3] * f->x[2] + f->a[4] * f->x[3] -
                f->b[1] * f->y[0] - f->b[2] * f->y[1] - f->b[3] * f->y[2] - f->b[4] * f->y[3];
    f->x[3] = f->x[2]; f->x[2] = f->x[1]; f->x[1] = f->x[0]; f->x[0] = in;
    f->y[3] = f->y[2]; f->y[2] = f->y[1]; f->y[1] = f->y[0]; f->y[0] = out;
    return out;
}

void or_equalizer_split(const float* x, uint32_t n, float* low, float* mid, float* high, float f1, float f2,
                        float f3) {
    const float fc1 = (f1 + f2) / 2.0f, fc2 = (f2 + f3) / 2.0f;
    or_lr4 hp1, hp2, lp1, lp2;
    lr4_init(&hp1, fc1, 1);
    lr4_init(&hp2, fc2, 1);
    lr4_init(&lp1, fc1, 0);
    lr4_init(&lp2, fc2, 0);
    for (uint32_t i = 0; i < n; ++i) {
        high[i] = lr4_step(&hp2, x[i]);
        low[i] = lr4_step(&lp1, x[i]);
        mid[i] = lr4_step(&hp1, lr4_step(&lp2, x[i]));
    }
}
