// ear_math.hpp -- device arithmetic of the tracer (gfx950).
//
// Every float expression here is evaluated as written, one IEEE rounding per
// operation: the library is compiled with -ffp-contract=off and without
// fast-math, so the CPU oracle and these kernels produce the same bits for the
// same Philox stream.  Where the reference relies on un-vendored gmtl 0.6.1 the
// textbook routine is used: Moeller-Trumbore 1997 (non-culling, |det| < 1e-5
// rejected) for gmtl::intersectDoubleSided, v - 2(v.n)n for gmtl::reflect,
// v / |v| for gmtl::normalize.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ear.h"

namespace ear {

#define EAR_HD __host__ __device__ __forceinline__

constexpr float kPi = 3.14159265f;  // src/Distributions.h:37
constexpr float kExp = 1000.0f;     // src/Scene.cpp:46
constexpr float kExpInt = kExp + 1.0f;
constexpr int kReflect = 0;
constexpr int kRefract = 1;

struct V3 {
    float x, y, z;
};
EAR_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
EAR_HD V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
EAR_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
EAR_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
EAR_HD V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
EAR_HD float length(V3 a) { return sqrtf(dot(a, a)); }
EAR_HD V3 normalized(V3 a) {
    float l = length(a);
    if (l != 0.0f) return V3{a.x / l, a.y / l, a.z / l};
    return a;
}
// gmtl::reflect(r, v, n): r = v - 2 (v.n) n
EAR_HD V3 reflect(V3 v, V3 n) {
    float t = dot(v, n);
    return V3{v.x - 2.0f * (t * n.x), v.y - 2.0f * (t * n.y), v.z - 2.0f * (t * n.z)};
}
EAR_HD bool is_normal(float x) {
    float a = fabsf(x);
    return a >= 1.17549435e-38f && a <= 3.40282347e+38f;  // false for 0, subnormal, inf, NaN
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Stream of one ray:
// key = seed, counter = (draw >> 2, ray, ctx, instance), word draw & 3,
// u = (word >> 8) * 2^-24.
// ---------------------------------------------------------------------------
EAR_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

struct Rng {
    uint32_t k0, k1, ray, ctx, inst, draw;
    uint32_t buf[4];
    EAR_HD void init(uint64_t seed, uint32_t r, uint32_t cx, uint32_t in, uint32_t d) {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        ray = r;
        ctx = cx;
        inst = in;
        draw = d;
        if (d & 3u) refill();
    }
    EAR_HD void refill() {
        buf[0] = draw >> 2;
        buf[1] = ray;
        buf[2] = ctx;
        buf[3] = inst;
        philox4x32_10(buf, k0, k1);
    }
    EAR_HD float next() {
        uint32_t w = draw & 3u;
        if (w == 0) refill();
        ++draw;
        uint32_t x = w == 0 ? buf[0] : w == 1 ? buf[1] : w == 2 ? buf[2] : buf[3];
        return to_unit(x);
    }
    EAR_HD static float to_unit(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }
    // The next three uniforms, exactly as three next() calls, with at most one
    // Philox block computed (a wave then runs one block evaluation per call
    // instead of one per uniform whose lanes sit at different phases).
    EAR_HD void next3(float& u1, float& u2, float& u3) {
        const uint32_t ph = draw & 3u;
        uint32_t nb[4] = {buf[0], buf[1], buf[2], buf[3]};
        if (ph != 1u) {
            // phase 0: the block of `draw` itself; phases 2, 3: the next one
            nb[0] = (draw >> 2) + (ph == 0u ? 0u : 1u);
            nb[1] = ray;
            nb[2] = ctx;
            nb[3] = inst;
            philox4x32_10(nb, k0, k1);
        }
        uint32_t a, b, c;
        if (ph == 0u) {
            a = nb[0];
            b = nb[1];
            c = nb[2];
        } else if (ph == 1u) {
            a = buf[1];
            b = buf[2];
            c = buf[3];
        } else if (ph == 2u) {
            a = buf[2];
            b = buf[3];
            c = nb[0];
        } else {
            a = buf[3];
            b = nb[0];
            c = nb[1];
        }
        if (ph != 1u) {
            buf[0] = nb[0];
            buf[1] = nb[1];
            buf[2] = nb[2];
            buf[3] = nb[3];
        }
        draw += 3u;
        u1 = to_unit(a);
        u2 = to_unit(b);
        u3 = to_unit(c);
    }
};

// ---------------------------------------------------------------------------
// pow(x, y) for the reference's pow(float,float)/powf call sites
// (src/Scene.cpp:154,232,243,251,308; src/StereoRecorder.cpp:110):
// exp2(y * log2 x) in binary64 with +,-,*,/ only, rounded once to binary32.
// log: m in [sqrt(1/2), sqrt 2), ln m = 2 atanh(s) = 2 sum s^(2j+1)/(2j+1),
// s = (m-1)/(m+1), j = 0..11.  exp2: n = floor(z + 1/2), f = z - n,
// 2^f = exp(f ln 2) by Taylor to degree 17, scaled by 2^n in two exact steps.
// ---------------------------------------------------------------------------
EAR_HD double log2_pos(double x) {
    uint64_t b = __builtin_bit_cast(uint64_t, x);
    int e = (int)((b >> 52) & 0x7ffu) - 1023;
    double m = __builtin_bit_cast(double, (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
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
    return (double)e + ln_m * 1.4426950408889634;
}

EAR_HD double exp2_d(double z) {
    if (z < -1060.0) return 0.0;
    if (z > 1060.0) return __builtin_bit_cast(double, 0x7FF0000000000000ull);
    double zn = z + 0.5;
    long long n = (long long)zn;
    if ((double)n > zn) n -= 1;
    double f = z - (double)n;
    double g = f * 0.6931471805599453;
    double p = 1.0 / 355687428096000.0;
    p = p * g + 1.0 / 20922789888000.0;
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
    long long n1 = n / 2, n2 = n - n1;
    double s1 = __builtin_bit_cast(double, (uint64_t)(n1 + 1023) << 52);
    double s2 = __builtin_bit_cast(double, (uint64_t)(n2 + 1023) << 52);
    return (p * s1) * s2;
}

EAR_HD float powf_ear(float xf, float yf) {
    double x = (double)xf, y = (double)yf;
    if (x != x || y != y) return __builtin_bit_cast(float, 0x7FC00000u);
    if (y == 0.0 || x == 1.0) return 1.0f;
    if (x == 0.0) return y > 0.0 ? 0.0f : __builtin_bit_cast(float, 0x7F800000u);
    if (x < 0.0) return __builtin_bit_cast(float, 0x7FC00000u);
    const double inf = __builtin_bit_cast(double, 0x7FF0000000000000ull);
    if (x == inf) return y > 0.0 ? __builtin_bit_cast(float, 0x7F800000u) : 0.0f;
    if (y == inf || y == -inf) return ((x < 1.0) == (y > 0.0)) ? 0.0f : __builtin_bit_cast(float, 0x7F800000u);
    return (float)exp2_d(y * log2_pos(x));
}

// powf_ear(xf, yf) for a base whose log2_pos((double)xf) is known (l2x):
// the same special cases and the same binary64 steps, without the logarithm
// (air absorption and head shadow raise a per-band constant to many powers).
EAR_HD float powf_ear_l2(float xf, double l2x, float yf) {
    double x = (double)xf, y = (double)yf;
    if (x != x || y != y) return __builtin_bit_cast(float, 0x7FC00000u);
    if (y == 0.0 || x == 1.0) return 1.0f;
    if (x == 0.0) return y > 0.0 ? 0.0f : __builtin_bit_cast(float, 0x7F800000u);
    if (x < 0.0) return __builtin_bit_cast(float, 0x7FC00000u);
    const double inf = __builtin_bit_cast(double, 0x7FF0000000000000ull);
    if (x == inf) return y > 0.0 ? __builtin_bit_cast(float, 0x7F800000u) : 0.0f;
    if (y == inf || y == -inf) return ((x < 1.0) == (y > 0.0)) ? 0.0f : __builtin_bit_cast(float, 0x7F800000u);
    return (float)exp2_d(y * l2x);
}

// log2_pos of a base for powf_ear_l2 (0 where powf_ear_l2 does not use it)
EAR_HD double pow_base_l2(float xf) {
    const double x = (double)xf;
    const double inf = __builtin_bit_cast(double, 0x7FF0000000000000ull);
    return (x > 0.0 && x < inf && x != 1.0) ? log2_pos(x) : 0.0;
}

// ---------------------------------------------------------------------------
// Moeller-Trumbore, double sided (gmtl::intersectDoubleSided), on a triangle
// stored as (v0, e1 = v1 - v0, e2 = v2 - v0).
// ---------------------------------------------------------------------------
EAR_HD bool mt_intersect(V3 v0, V3 e1, V3 e2, V3 o, V3 d, float& t) {
    V3 pvec = cross(d, e2);
    float det = dot(e1, pvec);
    if (det > -0.00001f && det < 0.00001f) return false;
    float inv_det = 1.0f / det;
    V3 tvec = sub(o, v0);
    float u = dot(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    V3 qvec = cross(tvec, e1);
    float v = dot(d, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, qvec) * inv_det;
    return true;
}

// ---------------------------------------------------------------------------
// Sampling (src/Distributions.h:48-75): rejection from the cube [-1,1]^3,
// then from the hemisphere of n; the specular blend normalizes once.
// ---------------------------------------------------------------------------
EAR_HD V3 sample_sphere(Rng& r) {
    for (;;) {
        float u1, u2, u3;
        r.next3(u1, u2, u3);
        float f1 = u1 * 2.0f - 1.0f;
        float f2 = u2 * 2.0f - 1.0f;
        float f3 = u3 * 2.0f - 1.0f;
        float l = f1 * f1 + f2 * f2 + f3 * f3;
        if (l < 0.001f || l > 1.0f) continue;
        float s = sqrtf(l);
        return V3{f1 / s, f2 / s, f3 / s};
    }
}
EAR_HD V3 sample_hemi(Rng& r, V3 n) {
    for (;;) {
        V3 v = sample_sphere(r);
        if (dot(n, v) < 0.0f) continue;
        return v;
    }
}
// The specular blend of Sample_Hemi (src/Distributions.h:71-75).
EAR_HD V3 blend_spec(V3 v, V3 refl, float f) {
    float g = 1.0f - f;
    V3 w{v.x * g + refl.x * f, v.y * g + refl.y * f, v.z * g + refl.z * f};
    return normalized(w);
}
EAR_HD V3 sample_hemi_spec(Rng& r, V3 n, V3 refl, float f) { return blend_spec(sample_hemi(r, n), refl, f); }

// ---------------------------------------------------------------------------
// The score form of the specularity gradient.  v = normalize((1-s) h + s r),
// h uniform on the hemisphere of nu, r.nu > 0, s < 1/2, has the density
// p(v) = t^2 / (2 pi (1-s) sqrt(D)) with c = v.r, D = (1-s)^2 - s^2 (1-c^2),
// t = s c + sqrt(D) on the image of the hemisphere; probability leaves that
// image through the image of the hemisphere's rim at the total rate
// F = (r.nu) / ((1-s) sqrt((1-s)^2 - s^2 (1 - (r.nu)^2))), with density
// w(phi) proportional to 1 / ((1-s) + s e(phi).r) along the rim e(phi).
// ---------------------------------------------------------------------------
EAR_HD double score_dlogp(double c, double s) {  // d log p(v) / ds at fixed v
    double D = (1.0 - s) * (1.0 - s) - s * s * (1.0 - c * c);
    double sq = sqrt(D);
    double Ds = -2.0 * (1.0 - s) - 2.0 * s * (1.0 - c * c);
    double t = s * c + sq;
    double ts = c + Ds / (2.0 * sq);
    return 2.0 * ts / t + 1.0 / (1.0 - s) - Ds / (2.0 * D);
}
EAR_HD double score_flux(double cr, double s) {  // F, cr = r.nu
    double sin2 = 1.0 - cr * cr;
    if (sin2 < 0.0) sin2 = 0.0;
    return cr / ((1.0 - s) * sqrt((1.0 - s) * (1.0 - s) - s * s * sin2));
}
// The stream of vertex k's spawn decision and continuation: the ray's
// counters under a key that depends on k.
EAR_HD void score_rng_init(Rng& r, uint64_t seed, uint32_t ray, uint32_t ctx, uint32_t inst, uint32_t k,
                           uint32_t draw) {
    r.k0 = (uint32_t)seed ^ 0x85EBCA6Bu;
    r.k1 = (uint32_t)(seed >> 32) ^ (0xC2B2AE35u * (k + 1u));
    r.ray = ray;
    r.ctx = ctx;
    r.inst = inst;
    r.draw = draw;
    if (draw & 3u) r.refill();
}
// A rim direction normalize((1-s) e(phi) + s r) with phi of density w / F:
// a uniform point z of the circle (rejection from the square) moved by the
// disc automorphism (z + rho) / (1 + rho z) has a wrapped Cauchy angle
// about 0; phi = pi + that angle, measured from r's component along the rim.
EAR_HD V3 score_rim_dir(Rng& R, V3 nu, V3 r, float s) {
    float zx, zy;
    for (;;) {
        float x = R.next() * 2.0f - 1.0f;
        float y = R.next() * 2.0f - 1.0f;
        float l = x * x + y * y;
        if (l < 0.001f || l > 1.0f) continue;
        float sl = sqrtf(l);
        zx = x / sl;
        zy = y / sl;
        break;
    }
    float cr = dot(r, nu);
    V3 rp{r.x - nu.x * cr, r.y - nu.y * cr, r.z - nu.z * cr};
    float sr = length(rp);
    V3 a;
    if (sr > 1e-6f) {
        a = V3{rp.x / sr, rp.y / sr, rp.z / sr};
    } else {
        // r along nu: any direction of the rim plane (the law is uniform)
        float ax = fabsf(nu.x), ay = fabsf(nu.y), az = fabsf(nu.z);
        V3 t{0.0f, 0.0f, 0.0f};
        if (ax <= ay && ax <= az)
            t.x = 1.0f;
        else if (ay <= az)
            t.y = 1.0f;
        else
            t.z = 1.0f;
        float td = dot(t, nu);
        a = normalized(V3{t.x - nu.x * td, t.y - nu.y * td, t.z - nu.z * td});
        sr = 0.0f;
    }
    V3 b = cross(nu, a);
    float A = 1.0f - s, B = s * sr;
    float rho = B / (A + sqrtf(A * A - B * B));
    float wx = (zx + rho) * (1.0f + rho * zx) + rho * zy * zy;
    float wy = zy * (1.0f - rho * rho);
    float wl = sqrtf(wx * wx + wy * wy);
    wx = wx / wl;
    wy = wy / wl;
    V3 e{-wx * a.x - wy * b.x, -wx * a.y - wy * b.y, -wx * a.z - wy * b.z};
    float g = 1.0f - s;
    return normalized(V3{e.x * g + r.x * s, e.y * g + r.y * s, e.z * g + r.z * s});
}

// ---------------------------------------------------------------------------
// 128-bit two's-complement fixed point (value * 2^EAR_FRAC_BITS) for the histograms.
// ---------------------------------------------------------------------------
struct I128 {
    uint64_t lo;
    int64_t hi;
};
EAR_HD I128 i128_add(I128 a, I128 b) {
    I128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1 : 0);
    return r;
}
EAR_HD I128 i128_neg(I128 a) {
    I128 r;
    r.lo = ~a.lo + 1;
    r.hi = ~a.hi + (r.lo == 0 ? 1 : 0);
    return r;
}
// a * k for 0 <= k < 2^31
EAR_HD I128 i128_mul_small(I128 a, uint32_t k) {
    uint64_t a0 = a.lo & 0xFFFFFFFFull, a1 = a.lo >> 32;
    uint64_t p0 = a0 * k;
    uint64_t p1 = a1 * k + (p0 >> 32);
    I128 r;
    r.lo = (p0 & 0xFFFFFFFFull) | (p1 << 32);
    r.hi = (int64_t)((uint64_t)a.hi * k + (p1 >> 32));
    return r;
}
// round-to-nearest-even of x * 2^EAR_FRAC_BITS as a 128-bit integer (|x| < 2^46)
EAR_HD I128 fixed_from_double(double x) {
    I128 r{0, 0};
    if (x == 0.0) return r;
    static_assert(EAR_FRAC_BITS == 80, "the scale below is 2^80");
    double y = x * 0x1p80;  // exact
    double a = y < 0 ? -y : y;
    // round to integer (ties to even) for a < 2^52; larger values are integers
    if (a < 4503599627370496.0) {
        double t = a + 4503599627370496.0;  // 2^52: forces rounding to an integer
        a = t - 4503599627370496.0;
    }
    uint64_t b = __builtin_bit_cast(uint64_t, a);
    int e = (int)((b >> 52) & 0x7ff) - 1075;  // a = m * 2^e
    uint64_t m = (b & 0x000FFFFFFFFFFFFFull) | 0x0010000000000000ull;
    if (a == 0.0) return r;
    if (e <= 0) {
        r.lo = m >> (-e);
        r.hi = 0;
    } else if (e < 64) {
        r.lo = m << e;
        r.hi = (int64_t)(e > 11 ? (m >> (64 - e)) : 0);
    } else {
        r.lo = 0;
        r.hi = (int64_t)(m << (e - 64));
    }
    return y < 0 ? i128_neg(r) : r;
}

}  // namespace ear
