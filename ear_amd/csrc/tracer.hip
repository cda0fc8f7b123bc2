// tracer.hip -- wavefront Monte-Carlo acoustic tracer (gfx950).
//
// One render replaces Scene::Render (src/Scene.cpp:111-318) for every band of
// ear_render_desc.band_mask at once; each band is its own reference context
// (SceneContext, src/SceneContext.h:26-54) with its own Philox streams.  The
// reference's per-ray loop over num_bounces (src/Scene.cpp:143-281) becomes a
// loop over wavefront iterations, each a chain of kernels over a compacted
// queue of live rays:
//   k_generate  emission, k = 0 (SoundRay, src/SoundFile.cpp:215-228; mesh
//               sources also connect to the recorders at k = 0)
//   k_extend    closest hit, BVH traversal (Mesh::RayIntersection)
//   k_shade     Scene::Bounce + Material::Bounce + energy bookkeeping + the
//               NEE weight of src/Scene.cpp:197-261; emits shadow requests and
//               the next ray (Sample_Hemi) into the next queue
//   k_shadow    any hit on the segment to the recorder (Mesh::LineIntersection)
//               and, when visible, Recorder::Record as a fixed-point ramp splat
//   k_finalize  exact prefix sums of the second-difference histograms,
//               1/N, direct sound (src/Scene.cpp:286-316), gain^2
// Histograms are 128-bit two's-complement fixed point (2^-80 units).  A ramp of
// width w and slope S (MonoRecorder::Record, src/MonoRecorder.cpp:83-97) is
// stored as three second-difference points (+wS at s, -(w+1)S at s+1, +S at
// s+w+1); integer sums are order independent, so results are bitwise
// deterministic whatever the launch shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "ear_internal.hpp"

namespace ear {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMaxBounces = 0xFFFFu;  // the bounce number's field in a queue entry (bits 8..23)
constexpr int kLdsMin = 16;  // smallest LDS part of the traversal stack; the overflow area covers the rest
constexpr int kShadeBlock = 256;  // k_shade: its block-wide appends take one counter atomic per block step
// no-allocation render: traversal grids of at most this many waves per CU
// (their overflow stacks come out of the fixed region), and the fewest rays
// in flight it accepts
constexpr uint32_t kNoallocWaves = 16;
constexpr size_t kNoallocMinRays = 4096;

enum StatIdx {
    ST_SEGMENTS = 0,
    ST_SHADOW = 1,
    ST_CONTRIB = 2,
    ST_NV_CL = 3,
    ST_NT_CL = 4,
    ST_NV_SH = 5,
    ST_NT_SH = 6,
    ST_RAYS = 7,
    ST_TAIL_SEG = 8,  // segments and shadow rays traced by k_tail (not by the
    ST_TAIL_SH = 9,   // wavefront kernels whose launch times are measured)
    ST_DROPPED = 10,  // ramps cut at the end of the histogram
    ST_BINS = 11,     // the largest ramp end among them (a maximum, not a sum)
    ST_SH_CONTRIB = 12,  // connections recorded by the shadow kernels (not k_tail)
    ST_SH_RAMPS = 13,    // and the ramps they made (one per channel)
    ST_N = 14
};

struct Queue {
    float4* o;  // {ox, oy, oz, I}
    float4* d;  // {dx, dy, dz, L}
    uint4* m;   // {ray id, draw counter, band | k << 8, -}
};

struct ShadowQ {
    float4* p;  // {px, py, pz, c}
    float4* m;  // {L, rec | band << 8, -, -}
};

struct Traversal {
    const BvhNode* __restrict__ nodes;    // 2-wide tree
    const Bvh4Node* __restrict__ nodes4;  // 4-wide tree (null when the scene is 2-wide)
    const float4* __restrict__ tris;
    uint32_t n_tris;
    int* gstack;        // overflow stack [(kStackDepth - LDS)][gthreads]
    uint32_t gthreads;  // persistent threads sharing gstack
};

// Lane refilling for the persistent traversal kernels: once at least `refill`
// lanes of a wave are idle, the wave hands them queue entries from its claimed
// run, claiming a new run with one atomic on `counter` (zero at launch) when
// its run is used up: `chunk` entries, or (chunk = 0) one per idle lane.
struct Fetch {
    uint32_t* counter;  // n_seg counters, one per 128-byte line (stride kSegStride)
    uint32_t refill;
    uint32_t chunk;
    uint32_t n_seg;
};
constexpr uint32_t kSegStride = 32;  // uint32 words per counter line
constexpr uint32_t kMaxSeg = 64;

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Appends the lanes with pred set; returns each such lane's slot.  All lanes
// of the wave must call it.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    uint64_t mask = __ballot(pred);
    uint32_t base = 0;
    if (mask) {
        int leader = __ffsll((unsigned long long)mask) - 1;
        if ((int)lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
        base = __shfl(base, leader);
    }
    uint64_t below = mask & ((1ull << lane_id()) - 1ull);
    return base + (uint32_t)__popcll(below);
}

// Block-wide append: like wave_append with one atomic per block instead of
// per wave (a queue counter is one address; per-wave atomics from every wave
// of the GPU serialise there).  Every thread of the block must call it;
// `sc` is block-shared scratch of BLOCK / 64 + 1 words that consecutive calls
// must not share (the caller alternates two).
template <int BLOCK>
__device__ __forceinline__ uint32_t block_append(uint32_t* counter, bool pred, uint32_t* sc) {
    constexpr uint32_t kWaves = BLOCK / 64;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint64_t mask = __ballot(pred);
    if (lane == 0) sc[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < kWaves; ++w) {
            const uint32_t cw = sc[w];
            sc[w] = tot;
            tot += cw;
        }
        sc[kWaves] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    return sc[kWaves] + sc[wave] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// Orders one wave's LDS accesses (lanes read what other lanes of the same
// wave wrote; the LDS executes a wave's instructions in order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sample_Hemi's rejection loops (src/Distributions.h:48-70) for the lanes of
// one wave together.  A lane's candidates are its consecutive triples of
// uniforms, so which one it accepts and how many draws it uses do not depend
// on which lane computes its Philox blocks.  In each round the wave's lanes
// compute blocks for the A lanes still sampling (64 / A blocks for each),
// and every sampling lane scans its own words in order with sample_hemi's
// tests and arithmetic: same vector, same draw count.  A wave then runs
// until its sampling lanes' candidates are used up together, instead of one
// block per candidate until its slowest lane accepts.  All lanes of the wave
// call it; `need` marks those that sample: around normal n if `hemi`
// (sample_hemi), else the whole sphere (sample_sphere).
struct HemiLds {
    uint4 own[64];   // the rank-th sampling lane's ray, ctx, inst and first block to compute
    uint2 key[64];   // its Philox key
    uint4 blk[128];  // per sampling lane (m + 1 each): its current block, then its m new blocks
};
__device__ __forceinline__ V3 sample_hemi_wave(bool need, bool hemi, Rng& rng, V3 n, HemiLds& S) {
    const uint32_t lane = lane_id();
    V3 v{0.0f, 0.0f, 0.0f};
    uint64_t act = __ballot(need);
    while (act) {
        const uint32_t A = (uint32_t)__popcll(act);
        const uint32_t m = 64u / A;  // A * (m + 1) <= 64 + A block slots
        const bool mine = (act >> lane) & 1ull;
        const uint32_t rank = (uint32_t)__popcll(act & ((1ull << lane) - 1ull));
        uint32_t ph = 0;
        if (mine) {
            ph = rng.draw & 3u;
            S.own[rank] = make_uint4(rng.ray, rng.ctx, rng.inst, (rng.draw >> 2) + (ph ? 1u : 0u));
            S.key[rank] = make_uint2(rng.k0, rng.k1);
            S.blk[rank * (m + 1)] = make_uint4(rng.buf[0], rng.buf[1], rng.buf[2], rng.buf[3]);
        }
        wave_sync();
        const uint32_t a = lane / m, o = lane - a * m;
        if (a < A) {
            const uint4 w = S.own[a];
            const uint2 k = S.key[a];
            uint32_t c[4] = {w.w + o, w.x, w.y, w.z};
            philox4x32_10(c, k.x, k.y);
            S.blk[a * (m + 1) + 1 + o] = make_uint4(c[0], c[1], c[2], c[3]);
        }
        wave_sync();
        bool left = false;
        if (mine) {
            // word `off` of this lane's run; the first unused one is its draw
            const uint32_t* words = (const uint32_t*)&S.blk[rank * (m + 1)];
            const uint32_t off0 = ph ? ph : 4u, end = 4u * (m + 1u);
            uint32_t off = off0;
            left = true;
            while (off + 3u <= end) {
                const float f1 = Rng::to_unit(words[off]) * 2.0f - 1.0f;
                const float f2 = Rng::to_unit(words[off + 1u]) * 2.0f - 1.0f;
                const float f3 = Rng::to_unit(words[off + 2u]) * 2.0f - 1.0f;
                off += 3u;
                const float l = f1 * f1 + f2 * f2 + f3 * f3;
                if (l < 0.001f || l > 1.0f) continue;
                const float s = sqrtf(l);
                const V3 c{f1 / s, f2 / s, f3 / s};
                if (hemi && dot(n, c) < 0.0f) continue;
                v = c;
                left = false;
                break;
            }
            rng.draw += off - off0;
            if (off & 3u) {
                const uint4 b = S.blk[rank * (m + 1) + (off >> 2)];
                rng.buf[0] = b.x;
                rng.buf[1] = b.y;
                rng.buf[2] = b.z;
                rng.buf[3] = b.w;
            }
        }
        wave_sync();  // this round's reads before the next round's writes
        act = __ballot(left);
    }
    return v;
}

// Statistics counters: [kStatCopies][ST_N], a block adding into copy
// blockIdx % kStatCopies (the host sums the copies), so that the one atomic
// per wave at the end of every launch does not pile up on one address.
constexpr uint32_t kStatCopies = 64;
__device__ __forceinline__ void stat_add(unsigned long long* stats, int idx, uint64_t v) {
    // reduce within the wave, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (lane_id() == 0 && v) atomicAdd(stats + (size_t)(blockIdx.x % kStatCopies) * ST_N + idx, (unsigned long long)v);
}

// Ramps that run past the last bin of the histogram.  The reference grows its
// buffer on such a write (FloatBuffer::operator[], src/Recorder.cpp:52-59);
// here the late bins are not recorded, and each such ramp is counted with the
// bin count that would have kept it whole, so that a caller can size the
// histogram and render again.
struct Drops {
    uint32_t n = 0, end = 0;
    __device__ __forceinline__ void add(uint32_t e) {
        if (e) {
            ++n;
            end = max(end, e);
        }
    }
};
// All lanes of the wave must call it.
__device__ __forceinline__ void stat_drops(unsigned long long* stats, const Drops& d) {
    stat_add(stats, ST_DROPPED, d.n);
    uint32_t e = d.end;
    for (int off = 32; off > 0; off >>= 1) e = max(e, (uint32_t)__shfl_down(e, off));
    if (lane_id() == 0 && e)
        atomicMax(stats + (size_t)(blockIdx.x % kStatCopies) * ST_N + ST_BINS, (unsigned long long)e);
}

// Lane refilling (see Fetch): call with the whole wave when !exhausted and at
// least f.refill lanes are idle (or all are); `act` is the ballot of busy
// lanes.  Returns the queue entry handed to this lane, or ~0u.  The queue
// [0, n) is cut into f.n_seg segments with a counter each, so that claims of
// many waves do not serialise on one address: a wave claims from its current
// segment `seg` and moves on to the next one when that is used up; after
// f.n_seg used-up segments in a row everything has been taken.
__device__ __forceinline__ uint32_t take_entries(const Fetch& f, uint32_t n, uint64_t act, uint32_t lane,
                                                 uint32_t& c_next, uint32_t& c_end, bool& exhausted,
                                                 uint32_t& seg) {
    const uint32_t n_idle = 64u - (uint32_t)__popcll(act);
    if (c_next == c_end) {
        // claim f.chunk entries, or exactly one per idle lane
        const uint32_t want = f.chunk ? f.chunk : n_idle;
        for (uint32_t fails = 0;;) {
            const uint32_t lo = (uint32_t)(((uint64_t)n * seg) / f.n_seg);
            const uint32_t hi = (uint32_t)(((uint64_t)n * (seg + 1)) / f.n_seg);
            uint32_t old = 0;
            if (lo < hi && lane == 0) old = atomicAdd(f.counter + seg * kSegStride, want);
            old = __shfl(old, 0);
            if (lo < hi && old < hi - lo) {
                c_next = lo + old;
                c_end = min(lo + old + want, hi);
                break;
            }
            if (++fails == f.n_seg) {
                exhausted = true;
                c_next = c_end = 0;
                break;
            }
            seg = seg + 1 == f.n_seg ? 0 : seg + 1;
        }
    }
    const uint64_t idle = ~act;
    const uint32_t avail = c_end - c_next;
    const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
    const uint32_t mine = c_next + rank;
    c_next += min(avail, n_idle);
    return (((act >> lane) & 1ull) == 0 && rank < avail) ? mine : 0xFFFFFFFFu;
}

// ---------------------------------------------------------------------------
// BVH traversal
// ---------------------------------------------------------------------------
// Loads through a pointer known to address global memory (the scene arrays
// reach the kernels inside structs, where the compiler would otherwise emit
// flat loads that also wait on the LDS counter).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ldg(const float4* p) {
    const f32x4 v = *(const __attribute__((address_space(1))) f32x4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 ldg(const int4* p) {
    const i32x4 v = *(const __attribute__((address_space(1))) i32x4*)p;
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 ldg(const uint4* p) {
    const i32x4 v = *(const __attribute__((address_space(1))) i32x4*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ float safe_inv(float x) {
    float y = fabsf(x) > 1e-20f ? x : copysignf(1e-20f, x);
    return 1.0f / y;
}

struct SlabRay {
    float ox, oy, oz, ix, iy, iz;
};

// entry/exit of the ray against box {lo, hi}; robust form (lo - o) * inv
__device__ __forceinline__ void slab(const SlabRay& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                     float& tn, float& tf) {
    float ax = (lx - r.ox) * r.ix, bx = (hx - r.ox) * r.ix;
    float ay = (ly - r.oy) * r.iy, by = (hy - r.oy) * r.iy;
    float az = (lz - r.oz) * r.iz, bz = (hz - r.oz) * r.iz;
    tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
}

// Per-thread traversal stack: the first LDS entries in shared memory
// (column-major over the block so a wave's accesses are conflict free), the
// rest in a global overflow area indexed [entry][persistent thread].
template <int LDS, int BLOCK>
struct Stack {
    int* lds;
    int* gstack;
    uint32_t gtid, gthreads;
    int sp = 0;
    __device__ __forceinline__ int* slot(int i) const {
        return i < LDS ? lds + i * BLOCK : gstack + (size_t)(i - LDS) * gthreads + gtid;
    }
    __device__ __forceinline__ void push(int v) {
        *slot(sp) = v;
        ++sp;
    }
    __device__ __forceinline__ int pop() {
        --sp;
        return *slot(sp);
    }
};

// Tests the triangles of leaf reference `cref`.  Closest hit (ANY = false):
// t in (0.001, best_t) with the lower original index winning ties, as the
// reference's strict '<' scan does (src/Mesh.cpp:39-45).  Any hit (ANY =
// true): segment parameter t in (1e-5, 1) (src/Mesh.cpp:58-71); returns true
// as soon as one is found.
template <bool ANY, bool INSTR>
__device__ __forceinline__ bool test_leaf(const Traversal& T, int32_t cref, V3 o, V3 d, float& best_t, int& best_s,
                                          uint32_t& best_o, bool& found, uint32_t& ntri) {
    uint32_t u = (uint32_t)cref;
    uint32_t first = u & 0x00FFFFFFu;
    uint32_t cnt = ((u >> 24) & 0x7Fu) + 1;
    for (uint32_t i = first; i < first + cnt; ++i) {
        float4 t0 = ldg(T.tris + 3 * i), t1 = ldg(T.tris + 3 * i + 1), t2 = ldg(T.tris + 3 * i + 2);
        V3 v0{t0.x, t0.y, t0.z}, e1{t0.w, t1.x, t1.y}, e2{t1.z, t1.w, t2.x};
        float t;
        if (INSTR) ++ntri;
        if (mt_intersect(v0, e1, e2, o, d, t)) {
            if (ANY) {
                if (t > 1e-5f && t < 1.0f) return true;
            } else {
                uint32_t orig = __float_as_uint(t2.y);
                if (t > 0.001f && (t < best_t || (t == best_t && orig < best_o))) {
                    best_t = t;
                    best_s = (int)i;
                    best_o = orig;
                    found = true;
                }
            }
        }
    }
    return false;
}

// a_k by masks (a chain of selects on k would be turned into an indexed
// array in scratch memory)
__device__ __forceinline__ uint32_t sel4(int k, uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3) {
    return (a0 & (0u - (uint32_t)(k == 0))) | (a1 & (0u - (uint32_t)(k == 1))) | (a2 & (0u - (uint32_t)(k == 2))) |
           (a3 & (0u - (uint32_t)(k == 3)));
}

__device__ __forceinline__ void cswap(float& ka, int& va, float& kb, int& vb) {
    if (kb < ka) {
        float tk = ka;
        ka = kb;
        kb = tk;
        int tv = va;
        va = vb;
        vb = tv;
    }
}

// One ray's BVH walk as a resumable state machine, so that a persistent lane
// can interleave the walks of rays it fetches one after another.  4-wide
// tree: a step either tests one pending leaf (nearest first; for the closest
// hit a leaf whose box starts beyond the best hit so far is dropped) or
// visits one internal node, whose hit children are sorted by entry distance:
// leaves become pending (a bit mask over the sorted children), the nearest
// internal child is visited next and the others are pushed.  One leaf per step keeps a wave's triangle tests in one
// code path instead of one per child slot.  2-wide tree: a step visits one
// node and tests its leaf children at once.  Closest hit (ANY = false):
// best_* hold the nearest hit so far (order independent: ties go to the lower
// original index).  Any hit: `found` is set when a blocker is found.
template <int LDS, int BLOCK, int WIDTH, bool ANY, bool INSTR>
struct Walker {
    SlabRay sr;
    V3 o, d;
    float best_t;
    int best_s;
    uint32_t best_o;
    int node;  // next internal node to visit, -1: pop one
    bool found;
    uint32_t pend;             // 4-wide tree: bit k set = child k of the last node is a leaf to test
    int pr0, pr1, pr2, pr3;    // that node's children, sorted by entry distance
    float pk0, pk1, pk2, pk3;  // and their entry distances
    Stack<LDS, BLOCK> stk;

    __device__ __forceinline__ Walker(int* lds, int* gstack, uint32_t gtid, uint32_t gthreads)
        : stk{lds, gstack, gtid, gthreads} {}

    // returns false when the walk is already over (empty scene)
    __device__ __forceinline__ bool start(const Traversal& T, V3 o_, V3 d_, float t_max) {
        o = o_;
        d = d_;
        sr = SlabRay{o.x, o.y, o.z, safe_inv(d.x), safe_inv(d.y), safe_inv(d.z)};
        best_t = t_max;
        best_s = -1;
        best_o = 0;
        node = 0;
        found = false;
        pend = 0;
        stk.sp = 0;
        return T.n_tris != 0;
    }

    __device__ __forceinline__ bool exhausted() const { return pend == 0 && node < 0 && stk.sp == 0; }

    // One step; returns true when the walk has ended (found / best_* final,
    // or for ANY a blocker found: found = true).
    __device__ __forceinline__ bool step(const Traversal& T, uint32_t& nvis, uint32_t& ntri) {
        if (WIDTH == 2) return step2(T, nvis, ntri);
        if (pend) {
            const int k = __builtin_ctz(pend);  // nearest pending leaf
            pend &= pend - 1u;
            const int ref = (int)sel4(k, (uint32_t)pr0, (uint32_t)pr1, (uint32_t)pr2, (uint32_t)pr3);
            const float tn = __uint_as_float(sel4(k, __float_as_uint(pk0), __float_as_uint(pk1),
                                                  __float_as_uint(pk2), __float_as_uint(pk3)));
            if (ANY || tn <= best_t) {
                if (test_leaf<ANY, INSTR>(T, ref, o, d, best_t, best_s, best_o, found, ntri)) {
                    found = true;
                    return true;
                }
            }
            return exhausted();
        }
        if (node < 0) node = stk.pop();  // not exhausted: the stack holds a node
        if (INSTR) ++nvis;
        const Bvh4Node* N = T.nodes4 + node;
        const uint4 og = ldg(&N->org), qa = ldg(&N->qa), qb = ldg(&N->qb);
        int4 rf = ldg(&N->ref);
        // decode the quantised child boxes: org + q * 2^(e - 127).  The
        // product is exact (q < 2^8, the scale a power of two, q * s normal or
        // 0), so one fused multiply-add rounds to the same float as the
        // host's multiply then add.
        const float ox = __uint_as_float(og.x), oy = __uint_as_float(og.y), oz = __uint_as_float(og.z);
        const float sx = __uint_as_float((og.w & 0xFFu) << 23), sy = __uint_as_float(((og.w >> 8) & 0xFFu) << 23),
                    sz = __uint_as_float(((og.w >> 16) & 0xFFu) << 23);
        float lxa[4], hxa[4], lya[4], hya[4], lza[4], hza[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            lxa[k] = __builtin_fmaf((float)((qa.x >> (8 * k)) & 0xFFu), sx, ox);
            hxa[k] = __builtin_fmaf((float)((qa.y >> (8 * k)) & 0xFFu), sx, ox);
            lya[k] = __builtin_fmaf((float)((qa.z >> (8 * k)) & 0xFFu), sy, oy);
            hya[k] = __builtin_fmaf((float)((qa.w >> (8 * k)) & 0xFFu), sy, oy);
            lza[k] = __builtin_fmaf((float)((qb.x >> (8 * k)) & 0xFFu), sz, oz);
            hza[k] = __builtin_fmaf((float)((qb.y >> (8 * k)) & 0xFFu), sz, oz);
        }
        int r[4] = {rf.x, rf.y, rf.z, rf.w};
        float key[4];
        const float lim = ANY ? 1.0f : best_t;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float tn, tf;
            slab(sr, lxa[k], lya[k], lza[k], hxa[k], hya[k], hza[k], tn, tf);
            const bool h = tn <= tf && tf >= 0.0f && tn <= lim && r[k] != kEmpty4;
            key[k] = h ? tn : INFINITY;
        }
        // nearest first
        cswap(key[0], r[0], key[1], r[1]);
        cswap(key[2], r[2], key[3], r[3]);
        cswap(key[0], r[0], key[2], r[2]);
        cswap(key[1], r[1], key[3], r[3]);
        cswap(key[1], r[1], key[2], r[2]);
        // leaves become pending (tested nearest first)
        pr0 = r[0];
        pr1 = r[1];
        pr2 = r[2];
        pr3 = r[3];
        pk0 = key[0];
        pk1 = key[1];
        pk2 = key[2];
        pk3 = key[3];
#pragma unroll
        for (int k = 0; k < 4; ++k) pend |= (key[k] != INFINITY && r[k] < 0) ? (1u << k) : 0u;
        // internal children: the nearest is visited next, the others pushed
        // farthest first
        node = -1;
#pragma unroll
        for (int k = 3; k >= 0; --k) {
            if (key[k] == INFINITY || r[k] < 0) continue;
            if (node >= 0) stk.push(node);
            node = r[k];
        }
        return exhausted();
    }

    __device__ __forceinline__ bool step2(const Traversal& T, uint32_t& nvis, uint32_t& ntri) {
        if (node < 0) node = stk.pop();
        if (INSTR) ++nvis;
        const BvhNode* N = T.nodes + node;
        float4 a = ldg(&N->a), b = ldg(&N->b), c = ldg(&N->c);
        int4 ch = ldg(&N->d);
        float tn0, tf0, tn1, tf1;
        slab(sr, a.x, a.y, a.z, a.w, b.x, b.y, tn0, tf0);
        slab(sr, b.z, b.w, c.x, c.y, c.z, c.w, tn1, tf1);
        float lim = ANY ? 1.0f : best_t;
        bool h0 = tn0 <= tf0 && tf0 >= 0.0f && tn0 <= lim;
        bool h1 = tn1 <= tf1 && tf1 >= 0.0f && tn1 <= lim;
        if (h0 && ch.x < 0) {
            if (test_leaf<ANY, INSTR>(T, ch.x, o, d, best_t, best_s, best_o, found, ntri)) {
                found = true;
                return true;
            }
            h0 = false;
        }
        if (h1 && ch.y < 0) {
            if (test_leaf<ANY, INSTR>(T, ch.y, o, d, best_t, best_s, best_o, found, ntri)) {
                found = true;
                return true;
            }
            h1 = false;
        }
        node = -1;
        if (h0 && h1) {
            bool first0 = tn0 <= tn1;
            stk.push(first0 ? ch.y : ch.x);
            node = first0 ? ch.x : ch.y;
        } else if (h0) {
            node = ch.x;
        } else if (h1) {
            node = ch.y;
        }
        return node < 0 && stk.sp == 0;
    }
};

// Returns true on a hit; best_* hold the closest hit (ANY = false).
template <int LDS, int BLOCK, int WIDTH, bool ANY, bool INSTR>
__device__ bool traverse(const Traversal& T, V3 o, V3 d, float& best_t, int& best_s, uint32_t& best_o, int* lds,
                         uint32_t gtid, uint32_t& nvis, uint32_t& ntri) {
    Walker<LDS, BLOCK, WIDTH, ANY, INSTR> w(lds, T.gstack, gtid, T.gthreads);
    if (!w.start(T, o, d, best_t)) return false;
    while (!w.step(T, nvis, ntri)) {
    }
    best_t = w.best_t;
    best_s = w.best_s;
    best_o = w.best_o;
    return w.found;
}

// ---------------------------------------------------------------------------
// histogram splat
// ---------------------------------------------------------------------------
__device__ __forceinline__ void acc128(unsigned long long* cell, I128 q) {
    unsigned long long lo = q.lo;
    unsigned long long old = atomicAdd(cell, lo);
    unsigned long long carry = (old + lo) < old ? 1ull : 0ull;
    unsigned long long hi = (unsigned long long)q.hi + carry;
    if (hi) atomicAdd(cell + 1, hi);
}

// Ramp of values (w - i) * A / w at bins s + i, i in [0, w), bins < 0 and
// >= nb dropped, added to the second-difference row h: +w*S at s, -(w+1)*S
// at s+1, +S at s+w+1.  Each point is added as a magnitude to the row's "+"
// or "-" part, so no add is negative and the high word is touched only on a
// carry (one atomic per point instead of two for negative values).  Returns
// the ramp's end (one past its last bin) if bins at or past nb were cut, else
// 0 (a start before bin 0 is the reference's own skip, src/StereoRecorder.cpp:92-95).
// The span of such a ramp: start s clipped to 0 (width ww = w + s then) and
// *cut as splat_ramp returns it; false when no point lands in [0, nb).
__device__ __forceinline__ bool ramp_span(uint32_t nb, int& s, int w, int& ww, uint32_t& cut) {
    cut = 0;
    if (w <= 0) return false;
    ww = w;
    if (s < 0) {
        ww = w + s;
        s = 0;
        if (ww <= 0) return false;
    }
    const uint32_t end = (uint32_t)s + (uint32_t)ww;
    cut = end > nb ? end : 0u;
    return (uint32_t)s < nb;
}

// The three points of a ramp of slope S = A / w (its original width) from
// bin s0 (already clipped, ww bins), as magnitudes: part `hp` of the row gets
// ww * Sq at s0 and Sq at s0 + ww + 1, part `hm` (ww + 1) * Sq at s0 + 1.
struct RampPoints {
    I128 sq, aq, p1;
    bool neg;  // S < 0: the "+" and "-" parts swap
};
__device__ __forceinline__ RampPoints ramp_points(float A, int w, int ww) {
    RampPoints r;
    const double S = (double)A / (double)w;
    r.neg = S < 0.0;
    r.sq = fixed_from_double(r.neg ? -S : S);
    r.aq = i128_mul_small(r.sq, (uint32_t)ww);
    r.p1 = i128_add(r.aq, r.sq);
    return r;
}

// The points of a clipped ramp (s0 < nb) added to row h with global atomics.
__device__ __forceinline__ void splat_span(unsigned long long* h, uint32_t nb, uint32_t s0, int ww, int w, float A) {
    const RampPoints q = ramp_points(A, w, ww);
    unsigned long long* hp = q.neg ? h + 2 * (size_t)nb : h;  // gets +w*S and +S
    unsigned long long* hm = q.neg ? h : h + 2 * (size_t)nb;  // gets (w+1)*S
    acc128(hp + 2 * (size_t)s0, q.aq);
    if (s0 + 1 < nb) acc128(hm + 2 * (size_t)(s0 + 1), q.p1);
    if (s0 + (uint32_t)ww + 1 < nb) acc128(hp + 2 * (size_t)(s0 + ww + 1), q.sq);
}

__device__ uint32_t splat_ramp(unsigned long long* h, uint32_t nb, int s, int w, float A) {
    int ww;
    uint32_t cut;
    if (!ramp_span(nb, s, w, ww, cut)) return cut;
    splat_span(h, nb, (uint32_t)s, ww, w, A);
    return cut;
}

// ---------------------------------------------------------------------------
// time-tiled splat (EAR_SPLAT_TILED)
// ---------------------------------------------------------------------------
// The shadow kernel does not add a visible connection's ramps to the
// histogram itself: it writes one record per ramp, slot-indexed (the shadow
// request's index times the records per request), keyed by (histogram row,
// time tile of kTile bins).  Per bounce iteration a counting sort groups the
// records by key (k_tile_count, k_tile_scan, k_tile_scatter), and
// k_tile_splat cuts the sorted records into chunks of kTileChunk, one
// workgroup each: for every key among its records it adds the three points of
// each ramp into a 128-bit accumulator tile in LDS (64-bit LDS atomics with an
// explicit carry), then adds that tile to the row in HBM, one global 128-bit
// add per touched bin.  The rays of one bounce iteration have travelled
// similar distances, so an iteration's records crowd into a few tiles; fixed
// chunks spread a crowded tile over many workgroups.  The sums are the same
// integers in another order, so the histograms are bitwise those of the
// direct splat.
constexpr uint32_t kTile = 256;             // bins per tile
constexpr uint32_t kTileChunk = 4096;       // sorted records per k_tile_splat workgroup
constexpr uint32_t kHalo = 64;              // bins past its tile that a ramp starting in it may reach
constexpr int kTileMaxW = (int)kHalo - 2;   // wider ramps are added directly
constexpr uint32_t kNoKey = 0xFFFFFFFFu;    // an empty record slot
constexpr uint32_t kLdsKeys = 8192;         // keys counted in LDS (more: global atomics)
constexpr uint32_t kTileDirect = 32;        // keys with at most this many ramps skip the LDS tile
constexpr int kTileBlock = 256;

// where the shadow kernels put their records (key == nullptr: direct splat)
struct TileOut {
    uint32_t* key;    // [request * stride + r]: tile key or kNoKey
    uint2* pay;       // {A bits, offset in tile | ww << 16 | w << 24}
    uint32_t stride;  // records per request: the most channels of one recorder
};

// Row layout of the accumulators.  Without the |c| histogram one signed row
// per channel.  With it, contributions are split by sign: row 0 takes the
// ramps of c >= 0, row 1 the ramps of |c| for c < 0, so each contribution
// updates one row; the finalize step turns them into signed = P - N and
// |.| = P + N (exact: quantisation is odd-symmetric, the sums are integers).
struct Hist {
    unsigned long long* cells;  // [band slot][channel][kind][+/-][bin] x {lo, hi}
    uint32_t n_bins, n_ch, kinds;
    uint32_t key_base, n_tiles;  // the time-tiled splat's keys of this render
    __device__ uint32_t row_index(int band, int ch, int kind) const {
        return ((uint32_t)band * n_ch + (uint32_t)ch) * kinds + (uint32_t)kind;
    }
    // the row's "+" part; its "-" part follows at + 2 * n_bins
    __device__ unsigned long long* row(int band, int ch, int kind) const {
        return cells + 4 * (size_t)n_bins * row_index(band, ch, kind);
    }
};

// Records of one connection for the time-tiled splat: they go to
// out.key / out.pay [base, base + out.stride); `used` so far.
struct TileSink {
    TileOut out;
    uint32_t base, used;
    // the slots this connection left empty
    __device__ void close() {
        for (uint32_t r = used; r < out.stride; ++r) out.key[base + r] = kNoKey;
    }
};

// One ramp into row (band, ch, kind): splat directly, or as a record when a
// sink is given (ramps wider than the tile halo are still splatted).
__device__ __forceinline__ void put_ramp(const Hist& H, int band, int ch, int kind, int s, int w, float A, Drops& dr,
                                         TileSink& sink, bool tiled) {
    if (!tiled) {
        dr.add(splat_ramp(H.row(band, ch, kind), H.n_bins, s, w, A));
        return;
    }
    int ww;
    uint32_t cut;
    const bool any = ramp_span(H.n_bins, s, w, ww, cut);
    dr.add(cut);
    if (!any) return;
    if (w > kTileMaxW) {
        splat_span(H.row(band, ch, kind), H.n_bins, (uint32_t)s, ww, w, A);
        return;
    }
    const uint32_t i = sink.base + sink.used++;
    sink.out.key[i] = H.key_base + H.row_index(band, ch, kind) * H.n_tiles + (uint32_t)s / kTile;
    sink.out.pay[i] = make_uint2(__float_as_uint(A), ((uint32_t)s % kTile) | ((uint32_t)ww << 16) | ((uint32_t)w << 24));
}

// Recorder::Record: MonoRecorder (src/MonoRecorder.cpp:83-97) or
// StereoRecorder (src/StereoRecorder.cpp:97-130).
// MonoRecorder::Record (src/MonoRecorder.cpp:83-97) into channel ch
__device__ __forceinline__ void record_channel(const Hist& H, int band, int ch, float a, float t, float dist,
                                               Drops& dr, TileSink& sink, bool tiled) {
    const bool split = H.kinds > 1;
    const uint32_t kind = split && a < 0.0f ? 1u : 0u;
    int s = (int)((double)t * 44100.0);
    float width = sqrtf(dist);
    float ampl = 2.0f * a / width;
    int w = (int)ceilf(width);
    put_ramp(H, band, ch, (int)kind, s, w, split ? fabsf(ampl) : ampl, dr, sink, tiled);
}
__device__ void record_channel(const Hist& H, int band, int ch, float a, float t, float dist, Drops& dr) {
    TileSink none{};
    record_channel(H, band, ch, a, t, dist, dr, none, false);
}

// Returns the number of ramps (channels) of the connection.
__device__ __forceinline__ uint32_t record(const Hist& H, int band, const RecParams& R, V3 dir, float a, float t,
                                           float dist, Drops& dr, TileSink& sink, bool tiled) {
    const bool split = H.kinds > 1;
    const uint32_t kind = split && a < 0.0f ? 1u : 0u;
    if (R.kind == EAR_REC_MONO) {
        record_channel(H, band, R.channel, a, t, dist, dr, sink, tiled);
        return 1;
    } else {
        float dt = dot(dir, V3{R.rx, R.ry, R.rz});
        float td = R.head_size / 343.0f;
        int s_right = (int)((double)(t - dt * td) * 44100.0);
        int s_left = (int)((double)(t + dt * td) * 44100.0);
        float width = sqrtf(dist);
        float ampl = 2.0f * a / width;
        float al = ampl, ar = ampl;
        float factor = powf_ear_l2(R.head_abs[band], R.l2_head_abs[band], fabsf(dt) * R.head_size);
        if (dt < 0.0f)
            ar *= factor * factor;
        else
            al *= factor * factor;
        int w = (int)ceilf(width);
        put_ramp(H, band, R.channel, (int)kind, s_left, w, split ? fabsf(al) : al, dr, sink, tiled);
        put_ramp(H, band, R.channel + 1, (int)kind, s_right, w, split ? fabsf(ar) : ar, dr, sink, tiled);
        return 2;
    }
}
__device__ uint32_t record(const Hist& H, int band, const RecParams& R, V3 dir, float a, float t, float dist,
                           Drops& dr) {
    TileSink none{};
    return record(H, band, R, dir, a, t, dist, dr, none, false);
}

// Receiver spheres (EAR_EST_SPHERE): the segment o + s u, s in [0, seg),
// deposits into every sphere it crosses I0 air^s0 (s1 - s0) / (4/3 pi r^3)
// at distance L0 + s0, [s0, s1) being the chord clipped to the segment (the
// track-length form of counting crossings).  j = reflections before the
// segment: sign (-1)^j, and dry on a point source's first segment.  Returns
// the number of deposits.
__device__ uint32_t sphere_deposits(const Hist& H, const RenderParams& P, int band, V3 o, V3 u, float seg, float I0,
                                    float L0, uint32_t j, Drops& dr) {
    uint32_t n = 0;
    for (uint32_t r = 0; r < P.n_rec; ++r) {
        const RecParams& R = P.rec[r];
        const int nc = R.kind == EAR_REC_STEREO ? 2 : 1;
        for (int e = 0; e < nc; ++e) {
            const V3 oc = sub(o, V3{R.cen[e][0], R.cen[e][1], R.cen[e][2]});
            const float b = dot(oc, u);
            const float disc = b * b - (dot(oc, oc) - R.r2);
            if (!(disc > 0.0f)) continue;
            const float q = sqrtf(disc);
            const float s0 = fmaxf(-b - q, 0.0f), s1 = fminf(-b + q, seg);
            if (!(s1 > s0)) continue;
            float c = I0;
            c *= powf_ear_l2(P.air[band], P.l2_air[band], s0);
            c *= s1 - s0;
            c *= R.inv_vol;
            if (j == 0 && P.src_kind == EAR_SRC_POINT) c *= P.dry;
            if (!is_normal(c)) continue;
            if (j & 1u) c *= -1.0f;
            const float dist = L0 + s0;
            record_channel(H, band, R.channel + e, c, dist / 343.0f, dist, dr);
            ++n;
        }
    }
    return n;
}

// ---------------------------------------------------------------------------
// per-instance device table
// ---------------------------------------------------------------------------
struct MeshSrc {
    const float* v;     // [n][9]
    const float* area;  // [n]
    const float* nrm;   // [n][3]
};

// direct sound per (band, channel) (src/Scene.cpp:299-311): {visible, s, w, A}
struct DirectRamp {
    int32_t visible, s, w;
    float a;
};

// One render instance: a (scene, render description) pair of a batch.  A
// plain render is a batch of one.
struct InstDev {
    RenderParams P;
    const BvhNode* nodes;
    const Bvh4Node* nodes4;
    const float4* tris;
    const float4* nm;
    const float4* mat_band;
    uint32_t n_tris;
    uint32_t key_base;  // time-tiled splat: first tile key of this render
    uint32_t n_tiles;   // and tiles per histogram row
    uint32_t pad0;
    MeshSrc M;
    unsigned long long* hist;  // [3][n_ch][kinds][bins] x 2
    DirectRamp* direct;        // [3][n_ch]
    double* value;             // [3][n_ch][kinds][bins] or null
    int32_t* log_hits;         // [3][trace_rays][trace_segs] or null
    uint32_t* log_nseg;        // [3][trace_rays]
    double inv_n, gain2;
};

// a run of rays of one (instance, band) context inside a chunk
struct GenSeg {
    uint32_t inst, band, ray0, count;
};

__device__ __forceinline__ Traversal make_trav(const InstDev& I, int* gstack, uint32_t gthreads) {
    return Traversal{I.nodes, I.nodes4, I.tris, I.n_tris, gstack, gthreads};
}
__device__ __forceinline__ Hist make_hist(const InstDev& I) {
    return Hist{I.hist, I.P.n_bins, I.P.n_channels, I.P.hist_kinds, I.key_base, I.n_tiles};
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Emission, k = 0 (AbstractSoundFile::SoundRay, src/SoundFile.cpp:215-228).
// Thread gid of the chunk maps to (instance, band, ray) through the segment
// table (exclusive prefix `seg_off`, binary search).
template <bool BATCH>
__global__ void __launch_bounds__(kBlock) k_generate(const InstDev* __restrict__ inst, const GenSeg* __restrict__ segs,
                                                      const uint32_t* __restrict__ seg_off, uint32_t n_segs,
                                                      uint32_t total, const uint32_t* __restrict__ plan,
                                                      uint32_t max_rec, uint32_t any_mesh, Queue q,
                                                      uint32_t* __restrict__ n_q, ShadowQ sq,
                                                      uint32_t* __restrict__ n_sq, unsigned long long* __restrict__ stats) {
    // rays [first, first + total) of the segment table; a plan written on the
    // device (k_plan) overrides the host's range
    uint32_t first = 0;
    if (plan) {
        first = plan[0];
        total = plan[1];
    }
    __shared__ HemiLds hl[kBlock / 64];
    uint32_t n_done = 0;
    for (uint32_t base = blockIdx.x * kBlock; base < total; base += gridDim.x * kBlock) {
        uint32_t gid = first + base + threadIdx.x;
        const uint32_t end = first + total;
        bool valid = gid < end;
        bool live = false;
        float4 ro{}, rd{};
        uint4 rm{};
        int band = 0;
        uint32_t id = 0;
        Rng rng;
        V3 o{0, 0, 0}, sn{0, 0, 0};
        bool mesh_src = false;
        uint32_t n_rec = 0;
        if (valid) {
            uint32_t lo = 0, hi = n_segs - 1;
            while (lo < hi) {  // last segment with seg_off <= gid
                uint32_t mid = (lo + hi + 1) >> 1;
                if (seg_off[mid] <= gid)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const GenSeg g = segs[lo];
            id = BATCH ? g.inst : 0;
            const InstDev& I = inst[id];
            const RenderParams& P = I.P;
            band = (int)g.band;
            uint32_t ray = g.ray0 + (gid - seg_off[lo]);
            rng.init(P.seed, ray, P.ctx_word[band], P.instance, 0);
            mesh_src = P.src_kind == EAR_SRC_MESH;
            n_rec = P.n_rec;
            if (mesh_src) {
                float x = rng.next() * P.src_total_area;
                uint32_t pick = P.src_n_tris - 1;
                for (uint32_t i = 0; i < P.src_n_tris; ++i) {
                    x -= I.M.area[i];
                    if (x < 0.0f) {
                        pick = i;
                        break;
                    }
                }
                const float* v = I.M.v + 9 * (size_t)pick;
                float r1 = rng.next();
                float r2 = rng.next();
                float sr1 = sqrtf(r1);
                float ca = 1.0f - sr1, cb = sr1 * (1.0f - r2), cc = sr1 * r2;
                o = V3{ca * v[0] + cb * v[3] + cc * v[6], ca * v[1] + cb * v[4] + cc * v[7],
                       ca * v[2] + cb * v[5] + cc * v[8]};
                sn = V3{I.M.nrm[3 * pick], I.M.nrm[3 * pick + 1], I.M.nrm[3 * pick + 2]};
            } else {
                o = V3{P.src_pos[0], P.src_pos[1], P.src_pos[2]};
            }
            live = !((double)1.0f < P.cutoff) && P.max_bounces > 0;
        }
        // a mesh source's hemisphere, a point source's sphere
        const V3 d = sample_hemi_wave(valid, mesh_src, rng, sn, hl[threadIdx.x >> 6]);
        if (valid) {
            ro = make_float4(o.x, o.y, o.z, 1.0f);
            rd = make_float4(d.x, d.y, d.z, 0.0f);
            rm = make_uint4(rng.ray, rng.draw, (uint32_t)band, id);
        }
        // mesh sources connect to every recorder at k = 0 (src/Scene.cpp:185)
        if (any_mesh) {
            for (uint32_t r = 0; r < max_rec; ++r) {
                bool emit = false;
                float c = 0.0f;
                if (valid && mesh_src && r < n_rec && inst[id].P.estimator == EAR_EST_NEE) {
                    const RecParams& R = inst[id].P.rec[r];
                    V3 dv = sub(V3{R.px, R.py, R.pz}, o);
                    float l = length(dv);
                    c = 1.0f;
                    c *= powf_ear_l2(inst[id].P.air[band], inst[id].P.l2_air[band], l);
                    c *= 2.0f / (4.0f * kPi * l * l);
                    emit = is_normal(c);
                }
                uint32_t slot = wave_append(n_sq, emit);
                if (emit) {
                    sq.p[slot] = make_float4(o.x, o.y, o.z, c);
                    sq.m[slot] = make_float4(0.0f, __uint_as_float(r | ((uint32_t)band << 8)), __uint_as_float(id),
                                             0.0f);
                }
            }
        }
        uint32_t slot = wave_append(n_q, live);
        if (live) {
            q.o[slot] = ro;
            q.d[slot] = rd;
            q.m[slot] = rm;
        }
        n_done += valid ? 1u : 0u;
    }
    stat_add(stats, ST_RAYS, n_done);
}

template <int BLOCK, int LDS, int WIDTH, bool INSTR, bool BATCH>
__global__ void __launch_bounds__(BLOCK) k_extend(const InstDev* __restrict__ inst, int* gstack, uint32_t gthreads,
                                                   Queue q, const uint32_t* __restrict__ n_q, int2* __restrict__ hits,
                                                   unsigned long long* __restrict__ stats, Fetch /*fetch*/) {
    __shared__ int lds[LDS * BLOCK];
    uint32_t n = *n_q;
    uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    uint32_t nvis = 0, ntri = 0;
    for (uint32_t base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
        uint32_t i = base + threadIdx.x;
        if (i < n) {
            float4 o = q.o[i], d = q.d[i];
            uint32_t id = BATCH ? q.m[i].w : 0;
            Traversal T = make_trav(inst[id], gstack, gthreads);
            float bt = 1000000.0f;
            int bs = -1;
            uint32_t bo = 0;
            traverse<LDS, BLOCK, WIDTH, false, INSTR>(T, V3{o.x, o.y, o.z}, V3{d.x, d.y, d.z}, bt, bs, bo,
                                                      lds + threadIdx.x, gtid, nvis, ntri);
            hits[i] = make_int2(__float_as_int(bt), bs);
        }
    }
    if (INSTR) {
        stat_add(stats, ST_NV_CL, nvis);
        stat_add(stats, ST_NT_CL, ntri);
    }
}

// Scene::Bounce + Material::Bounce + the energy bookkeeping and NEE weight of
// Scene::Render (src/Scene.cpp:49-82, 151-277, src/Material.cpp:76-83).
// Six waves per SIMD: the register cap that costs no spills (2% faster than
// the five the compiler picks by itself).
// The surface interaction of a segment from ro/rd that hit triangle h.y at
// t = h.x (src/Scene.cpp:143-175): hit point, the normal facing the ray
// (negated for REFRACT), Material::Bounce's uniform (its Philox stream
// starts here), air over the segment and the material's retention.
__device__ __forceinline__ void surface_hit(const InstDev& In, const RenderParams& P, int band, float4 ro, float4 rd,
                                            uint4 rm, int2 h, Rng& rng, V3& p, V3& nrm, V3& prev, V3& d, float& I,
                                            float& L, float& spec, float& seg, int& bt, bool& alive) {
    float t = __int_as_float(h.x);
    V3 o{ro.x, ro.y, ro.z};
    d = V3{rd.x, rd.y, rd.z};
    p = V3{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
    float4 nmv = In.nm[h.y];
    V3 N{nmv.x, nmv.y, nmv.z};
    nrm = dot(N, d) > 0.0f ? V3{N.x * -1.0f, N.y * -1.0f, N.z * -1.0f} : N;
    uint32_t mat = __float_as_uint(nmv.w);
    float4 mb = In.mat_band[3 * mat + band];  // {fl2 | -1, spec, retention, -}
    rng.init(P.seed, rm.x, P.ctx_word[band], P.instance, rm.y);
    bt = kReflect;
    if (mb.x >= 0.0f) bt = (rng.next() <= mb.x) ? kReflect : kRefract;
    spec = mb.y;
    if (bt == kRefract) nrm = neg(nrm);
    float l = length(sub(p, o));
    I = ro.w;
    I *= powf_ear_l2(P.air[band], P.l2_air[band], l);
    L = rd.w + l;
    I *= mb.z;
    alive = is_normal(I);
    prev = normalized(d);
    seg = l;
}

// The next-event weight of the connection from hit point p to recorder R
// (src/Scene.cpp:185-268): false if it faces away or is not FP_NORMAL;
// c carries the sign of the k-th reflection.
__device__ __forceinline__ bool nee_weight(const RenderParams& P, const RecParams& R, int band, V3 p, V3 nrm, V3 prev,
                                           int bt, float spec, float I, uint32_t k, float& c) {
    V3 dv = sub(V3{R.px, R.py, R.pz}, p);
    V3 lsdir = normalized(dv);
    float dt = dot(lsdir, nrm);
    if (!(dt > 0.0f)) return false;
    // factor = spec * 1001 * sf^1000 + (1 - spec) * diff; with spec == 0 the
    // lobe term is exactly +0 (sf^1000 is finite), so its binary64 pow is
    // skipped without changing a bit
    float diff, sf;
    if (bt == kReflect) {
        V3 rv = reflect(prev, nrm);
        diff = -dot(nrm, prev);
        sf = fmaxf(0.0f, dot(rv, lsdir));
    } else {
        diff = dot(nrm, prev);
        sf = fmaxf(0.0f, dot(prev, lsdir));
    }
    const float lobe = spec != 0.0f ? spec * kExpInt * powf_ear(sf, kExp) : 0.0f;
    const float factor = lobe + (1.0f - spec) * diff;
    c = I;
    c *= factor;
    float l = length(dv);
    c *= powf_ear_l2(P.air[band], P.l2_air[band], l);
    c *= 2.0f / (4.0f * kPi * l * l);
    if (!is_normal(c)) return false;
    if (k & 1u) c *= -1.0f;
    return true;
}

template <bool BATCH>
__global__ void __launch_bounds__(kShadeBlock, 6) k_shade(const InstDev* __restrict__ inst, Queue qin,
                                                   const uint32_t* __restrict__ n_in, const int2* __restrict__ hits,
                                                   Queue qout, uint32_t* __restrict__ n_out, ShadowQ sq,
                                                   uint32_t* __restrict__ n_sq, uint32_t max_rec,
                                                   unsigned long long* __restrict__ stats) {
    __shared__ uint32_t sc[4][kShadeBlock / 64 + 1];  // scratch of the block-wide appends, alternated
    __shared__ HemiLds hl[kShadeBlock / 64];
    uint32_t n = *n_in;
    uint32_t n_done = 0, n_sph = 0;
    Drops drops;
    for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
        uint32_t i = base + threadIdx.x;
        bool valid = i < n;
        bool alive = false;  // hit and intensity normal: NEE runs
        bool cont = false;   // continues to the next segment
        uint4 rm{};
        V3 p{0, 0, 0}, nrm{0, 0, 0}, prev{0, 0, 0}, d{0, 0, 0};
        float I = 0.0f, L = 0.0f, spec = 0.0f;
        int bt = kReflect;
        int band = 0;
        uint32_t k = 0, id = 0, n_rec = 0;
        float seg = __int_as_float(0x7f800000);  // segment length; +inf for a miss
        bool nee = true;
        Rng rng;
        if (valid) {
            float4 ro = qin.o[i];
            float4 rd = qin.d[i];
            rm = qin.m[i];
            id = BATCH ? rm.w : 0;
            const InstDev& In = inst[id];
            const RenderParams& P = In.P;
            n_rec = P.n_rec;
            band = (int)(rm.z & 0xFFu);
            k = ((rm.z >> 8) & 0xFFFFu) + 1;  // this segment's bounce number
            int2 h = hits[i];
            if (In.log_hits && rm.x < P.trace_rays) {
                if (k - 1 < P.trace_segs)
                    In.log_hits[((size_t)band * P.trace_rays + rm.x) * P.trace_segs + (k - 1)] =
                        h.y >= 0 ? (int32_t)__float_as_uint(In.tris[3 * h.y + 2].y) : -1;
                In.log_nseg[(size_t)band * P.trace_rays + rm.x] = k;
            }
            if (h.y >= 0) surface_hit(In, P, band, ro, rd, rm, h, rng, p, nrm, prev, d, I, L, spec, seg, bt, alive);
            if (P.estimator == EAR_EST_SPHERE) {
                // hit or miss, the segment crosses the receiver spheres
                n_sph += sphere_deposits(make_hist(In), P, band, V3{ro.x, ro.y, ro.z},
                                         normalized(V3{rd.x, rd.y, rd.z}), seg, ro.w, rd.w, k - 1, drops);
                nee = false;
            }
        }
        // NEE to every recorder (src/Scene.cpp:185-268)
        for (uint32_t r = 0; r < max_rec; ++r) {
            bool emit = false;
            float c = 0.0f;
            if (alive && nee && r < n_rec) {
                const RenderParams& P = inst[id].P;
                emit = nee_weight(P, P.rec[r], band, p, nrm, prev, bt, spec, I, k, c);
            }
            uint32_t slot = block_append<kShadeBlock>(n_sq, emit, sc[r & 1u]);
            if (emit) {
                sq.p[slot] = make_float4(p.x, p.y, p.z, c);
                sq.m[slot] = make_float4(L, __uint_as_float(r | ((uint32_t)band << 8)), __uint_as_float(id), 0.0f);
            }
        }
        if (alive) cont = !((double)I < inst[id].P.cutoff) && k < inst[id].P.max_bounces;
        float4 no{}, nd{};
        uint4 nmeta{};
        const V3 hv = sample_hemi_wave(cont, true, rng, nrm, hl[threadIdx.x >> 6]);
        if (cont) {
            // REFRACT: the negated normal and the incoming direction; REFLECT:
            // the mirror direction about the ray-facing normal (src/Scene.cpp:65-74)
            V3 rdir = bt == kRefract ? d : reflect(d, nrm);
            V3 v = blend_spec(hv, rdir, spec);
            no = make_float4(p.x, p.y, p.z, I);
            nd = make_float4(v.x, v.y, v.z, L);
            nmeta = make_uint4(rm.x, rng.draw, (uint32_t)band | (k << 8), id);
        }
        uint32_t slot = block_append<kShadeBlock>(n_out, cont, sc[2 + ((base / (gridDim.x * kShadeBlock)) & 1u)]);
        if (cont) {
            qout.o[slot] = no;
            qout.d[slot] = nd;
            qout.m[slot] = nmeta;
        }
        n_done += valid ? 1u : 0u;
    }
    stat_add(stats, ST_SEGMENTS, n_done);
    stat_add(stats, ST_CONTRIB, n_sph);
    stat_drops(stats, drops);
}

// The tail of the bounce loop: once every ray has been emitted and few are
// left, each lane carries its rays through all their remaining segments in
// one launch (closest hit as k_extend, the surface interaction and
// next-event weights as k_shade, visibility and splat as k_shadow, the next
// direction as k_shade), instead of one launch of each kernel per segment,
// whose time is then set by the latency of a few walks.  The same device
// functions in the same order: every path, and so every histogram sum, is
// the wavefront loop's.  A path ends by a miss, a non-normal intensity, the
// cut-off or max_bounces, so every lane's loop ends.  Not used with the
// sphere estimator, hit logging or counter instrumentation.
template <int BLOCK, int LDS, int WIDTH, bool BATCH>
__global__ void __launch_bounds__(BLOCK) k_tail(const InstDev* __restrict__ inst, int* gstack, uint32_t gthreads,
                                                Queue q, const uint32_t* __restrict__ n_q,
                                                unsigned long long* __restrict__ stats) {
    __shared__ int lds[LDS * BLOCK];
    const uint32_t n = *n_q;
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    uint32_t n_seg = 0, nsh = 0, ncon = 0, nvis = 0, ntri = 0;
    Drops drops;
    for (uint32_t i = gtid; i < n; i += gridDim.x * BLOCK) {
        float4 ro = q.o[i], rd = q.d[i];
        uint4 rm = q.m[i];
        const uint32_t id = BATCH ? rm.w : 0;
        const InstDev& In = inst[id];
        const RenderParams& P = In.P;
        const Traversal T = make_trav(In, gstack, gthreads);
        const int band = (int)(rm.z & 0xFFu);
        for (;;) {
            const uint32_t k = ((rm.z >> 8) & 0xFFFFu) + 1;  // this segment's bounce number
            float bt_t = 1000000.0f;
            int bs = -1;
            uint32_t bo = 0;
            traverse<LDS, BLOCK, WIDTH, false, false>(T, V3{ro.x, ro.y, ro.z}, V3{rd.x, rd.y, rd.z}, bt_t, bs, bo,
                                                      lds + threadIdx.x, gtid, nvis, ntri);
            ++n_seg;
            if (bs < 0) break;  // no hit ends the path (src/Scene.cpp:166)
            Rng rng;
            V3 p, nrm, prev, d;
            float I = 0.0f, L = 0.0f, spec = 0.0f, seg = 0.0f;
            int bt = kReflect;
            bool alive = false;
            surface_hit(In, P, band, ro, rd, rm, make_int2(__float_as_int(bt_t), bs), rng, p, nrm, prev, d, I, L, spec,
                        seg, bt, alive);
            if (!alive) break;
            for (uint32_t r = 0; r < P.n_rec; ++r) {
                float c = 0.0f;
                const RecParams& R = P.rec[r];
                if (!nee_weight(P, R, band, p, nrm, prev, bt, spec, I, k, c)) continue;
                // Scene::Connect: the segment from the hit point to the recorder
                V3 dv = sub(V3{R.px, R.py, R.pz}, p);
                float st = 0.0f;
                int ss = -1;
                uint32_t so = 0;
                ++nsh;
                const bool occ = traverse<LDS, BLOCK, WIDTH, true, false>(T, p, dv, st, ss, so, lds + threadIdx.x,
                                                                          gtid, nvis, ntri);
                if (!occ) {
                    float l = length(dv);
                    V3 lsdir = normalized(dv);
                    record(make_hist(In), band, R, lsdir, c, (L + l) / 343.0f, L + l, drops);
                    ++ncon;
                }
            }
            if ((double)I < P.cutoff || k >= P.max_bounces) break;
            // REFRACT: the negated normal and the incoming direction; REFLECT:
            // the mirror direction about the ray-facing normal (src/Scene.cpp:65-74)
            V3 rdir = bt == kRefract ? d : reflect(d, nrm);
            V3 v = sample_hemi_spec(rng, nrm, rdir, spec);
            ro = make_float4(p.x, p.y, p.z, I);
            rd = make_float4(v.x, v.y, v.z, L);
            rm = make_uint4(rm.x, rng.draw, (uint32_t)band | (k << 8), id);
        }
    }
    stat_add(stats, ST_SEGMENTS, n_seg);
    stat_add(stats, ST_SHADOW, nsh);
    stat_add(stats, ST_CONTRIB, ncon);
    stat_add(stats, ST_TAIL_SEG, n_seg);
    stat_add(stats, ST_TAIL_SH, nsh);
    stat_drops(stats, drops);
}

// Scene::Connect + Mesh::LineIntersection (src/Scene.cpp:84-96,
// src/Mesh.cpp:58-71), then Recorder::Record for visible connections.
template <int BLOCK, int LDS, int WIDTH, bool INSTR, bool BATCH, bool TILED>
__global__ void __launch_bounds__(BLOCK) k_shadow(const InstDev* __restrict__ inst, int* gstack, uint32_t gthreads,
                                                   ShadowQ sq, const uint32_t* __restrict__ n_sq,
                                                   unsigned long long* __restrict__ stats, Fetch /*fetch*/,
                                                   TileOut to) {
    __shared__ int lds[LDS * BLOCK];
    uint32_t n = *n_sq;
    uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    uint32_t nvis = 0, ntri = 0, ncon = 0, nsh = 0, nramp = 0;
    Drops drops;
    for (uint32_t base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
        uint32_t i = base + threadIdx.x;
        if (i < n) {
            float4 pc = sq.p[i], m = sq.m[i];
            uint32_t packed = __float_as_uint(m.y);
            int r = (int)(packed & 0xFFu), band = (int)((packed >> 8) & 0xFFu);
            uint32_t id = BATCH ? __float_as_uint(m.z) : 0;
            const InstDev& In = inst[id];
            const RecParams& R = In.P.rec[r];
            V3 p{pc.x, pc.y, pc.z};
            V3 dv = sub(V3{R.px, R.py, R.pz}, p);
            float bt = 0.0f;
            int bs = -1;
            uint32_t bo = 0;
            ++nsh;
            Traversal T = make_trav(In, gstack, gthreads);
            bool occ = traverse<LDS, BLOCK, WIDTH, true, INSTR>(T, p, dv, bt, bs, bo, lds + threadIdx.x, gtid, nvis,
                                                                ntri);
            TileSink sink{to, i * to.stride, 0};
            if (!occ) {
                float l = length(dv);
                V3 lsdir = normalized(dv);
                float L = m.x;
                nramp += record(make_hist(In), band, R, lsdir, pc.w, (L + l) / 343.0f, L + l, drops, sink, TILED);
                ++ncon;
            }
            if (TILED) sink.close();
        }
    }
    stat_add(stats, ST_SHADOW, nsh);
    stat_add(stats, ST_CONTRIB, ncon);
    stat_add(stats, ST_SH_CONTRIB, ncon);
    stat_add(stats, ST_SH_RAMPS, nramp);
    stat_drops(stats, drops);
    if (INSTR) {
        stat_add(stats, ST_NV_SH, nvis);
        stat_add(stats, ST_NT_SH, ntri);
    }
}

// Persistent variants of k_extend / k_shadow: a lane whose walk has ended
// takes the next queue entry, so a wave keeps its lanes busy instead of
// waiting for its slowest ray (Aila & Laine, "Understanding the Efficiency of
// Ray Traversal on GPUs", HPG 2009).  A wave hands out entries when at least
// fetch.refill of its lanes are idle, from a run of fetch.chunk entries it
// claims with one atomic.  Each lane keeps its own traversal stack across the
// rays it takes.

template <int BLOCK, int LDS, int WIDTH, bool INSTR, bool BATCH>
__global__ void __launch_bounds__(BLOCK) k_extend_f(const InstDev* __restrict__ inst, int* gstack, uint32_t gthreads,
                                                     Queue q, const uint32_t* __restrict__ n_q,
                                                     int2* __restrict__ hits, unsigned long long* __restrict__ stats,
                                                     Fetch fetch) {
    __shared__ int lds[LDS * BLOCK];
    const uint32_t n = *n_q;
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t lane = lane_id();
    uint32_t nvis = 0, ntri = 0;
    Walker<LDS, BLOCK, WIDTH, false, INSTR> w(lds + threadIdx.x, gstack, gtid, gthreads);
    Traversal T = make_trav(inst[0], gstack, gthreads);
    // a lane is walking (active), holds a finished hit not yet stored
    // (pending) or is free; pending hits are stored together when the wave
    // refills, so the 8-byte records of a claimed run land in its cache lines
    // in a few bursts instead of one lane at a time
    bool active = false, pending = false, exhausted = n == 0;
    uint32_t slot = 0;
    int2 hit = make_int2(0, -1);
    uint32_t c_next = 0, c_end = 0;  // the wave's claimed entries [c_next, c_end)
    uint32_t seg = ((blockIdx.x * BLOCK + threadIdx.x) >> 6) % fetch.n_seg;  // the wave's first segment
    for (;;) {
        uint64_t act = __ballot(active);
        const uint32_t n_idle = 64u - (uint32_t)__popcll(act);
        if (n_idle >= fetch.refill || act == 0) {
            if (pending) {
                hits[slot] = hit;
                pending = false;
            }
            if (!exhausted) {
                const uint32_t mine = take_entries(fetch, n, act, lane, c_next, c_end, exhausted, seg);
                if (mine != 0xFFFFFFFFu) {
                    slot = mine;
                    const float4 o = q.o[mine], d = q.d[mine];
                    if (BATCH) T = make_trav(inst[q.m[mine].w], gstack, gthreads);
                    if (w.start(T, V3{o.x, o.y, o.z}, V3{d.x, d.y, d.z}, 1000000.0f))
                        active = true;
                    else
                        hits[mine] = make_int2(__float_as_int(1000000.0f), -1);
                }
            }
            act = __ballot(active);
        }
        // the loop is left by the whole wave at once, with no hit pending
        // (the refill step above has just stored them)
        if (act == 0) {
            if (exhausted) break;
            continue;
        }
        if (active && w.step(T, nvis, ntri)) {
            hit = make_int2(__float_as_int(w.best_t), w.best_s);
            pending = true;
            active = false;
        }
    }
    if (INSTR) {
        stat_add(stats, ST_NV_CL, nvis);
        stat_add(stats, ST_NT_CL, ntri);
    }
}

template <int BLOCK, int LDS, int WIDTH, bool INSTR, bool BATCH, bool TILED>
__global__ void __launch_bounds__(BLOCK, 6) k_shadow_f(const InstDev* __restrict__ inst, int* gstack, uint32_t gthreads,
                                                     ShadowQ sq, const uint32_t* __restrict__ n_sq,
                                                     unsigned long long* __restrict__ stats, Fetch fetch,
                                                     TileOut to) {
    __shared__ int lds[LDS * BLOCK];
    // per wave: request slots of visible connections awaiting their records
    // (a lane adds at most one per pass of the loop below and each pass
    // starts by recording 64 once 64 wait, so at most 127 ever wait)
    __shared__ uint32_t visq[BLOCK / 64][192];
    const uint32_t n = *n_sq;
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t lane = lane_id();
    uint32_t nvis = 0, ntri = 0, ncon = 0, nsh = 0, nramp = 0;
    Drops drops;
    Walker<LDS, BLOCK, WIDTH, true, INSTR> w(lds + threadIdx.x, gstack, gtid, gthreads);
    Traversal T = make_trav(inst[0], gstack, gthreads);
    // A lane is walking (active) or free.  A walk that ends visible leaves
    // its request slot in the wave's list and the lane takes new work at the
    // next refill; once 64 slots wait, every lane of the wave records one
    // (the splat then runs with the whole wave instead of the few lanes that
    // happened to finish together).  The rest are recorded at the end.
    uint32_t* vq = visq[threadIdx.x >> 6];
    uint32_t nq = 0;  // slots in vq, the same in every lane
    auto push = [&](bool v, uint32_t s) {  // the whole wave
        const uint64_t b = __ballot(v);
        if (v) vq[nq + (uint32_t)__popcll(b & ((1ull << lane) - 1ull))] = s;
        nq += (uint32_t)__popcll(b);
    };
    auto drain = [&](uint32_t take) {  // the whole wave: records the last `take` (<= 64) slots
        wave_sync();
        const bool mine = lane < take;
        const uint32_t rs = mine ? vq[nq - take + lane] : 0u;
        wave_sync();
        nq -= take;
        if (mine) {
            const float4 pc = sq.p[rs], m = sq.m[rs];
            const uint32_t packed = __float_as_uint(m.y);
            const int r = (int)(packed & 0xFFu), band = (int)((packed >> 8) & 0xFFu);
            const uint32_t id = BATCH ? __float_as_uint(m.z) : 0;
            const InstDev& In = inst[id];
            const RecParams& R = In.P.rec[r];
            const V3 dv = sub(V3{R.px, R.py, R.pz}, V3{pc.x, pc.y, pc.z});
            const float l = length(dv);
            TileSink sink{to, rs * to.stride, 0};
            nramp += record(make_hist(In), band, R, normalized(dv), pc.w, (m.x + l) / 343.0f, m.x + l, drops, sink,
                            TILED);
            if (TILED) sink.close();
            ++ncon;
        }
    };
    bool active = false, exhausted = n == 0;
    uint32_t slot = 0;
    uint32_t c_next = 0, c_end = 0;  // the wave's claimed entries [c_next, c_end)
    uint32_t seg = ((blockIdx.x * BLOCK + threadIdx.x) >> 6) % fetch.n_seg;  // the wave's first segment
    for (;;) {
        if (nq >= 64) drain(64);
        uint64_t act = __ballot(active);
        const uint32_t n_free = 64u - (uint32_t)__popcll(act);
        if (n_free >= fetch.refill || act == 0) {
            if (!exhausted) {
                const uint32_t mine = take_entries(fetch, n, act, lane, c_next, c_end, exhausted, seg);
                bool vis = false;
                if (mine != 0xFFFFFFFFu) {
                    slot = mine;
                    ++nsh;
                    const float4 pc = sq.p[mine], m = sq.m[mine];
                    const uint32_t packed = __float_as_uint(m.y);
                    const uint32_t id = BATCH ? __float_as_uint(m.z) : 0;
                    if (BATCH) T = make_trav(inst[id], gstack, gthreads);
                    const RecParams& R = inst[id].P.rec[packed & 0xFFu];
                    const V3 p{pc.x, pc.y, pc.z};
                    if (w.start(T, p, sub(V3{R.px, R.py, R.pz}, p), 0.0f))
                        active = true;
                    else
                        vis = true;  // empty scene: visible
                }
                push(vis, slot);
            }
            act = __ballot(active);
        }
        // the loop is left by the whole wave at once (take_entries is a
        // wave-wide step), after the waiting records
        if (act == 0) {
            if (exhausted) {
                while (nq > 0) drain(nq < 64u ? nq : 64u);
                break;
            }
            continue;
        }
        bool vis = false;
        if (active && w.step(T, nvis, ntri)) {
            active = false;
            vis = !w.found;
            if (TILED && w.found) TileSink{to, slot * to.stride, 0}.close();  // occluded: no records
        }
        push(vis, slot);
    }
    stat_add(stats, ST_SHADOW, nsh);
    stat_add(stats, ST_CONTRIB, ncon);
    stat_add(stats, ST_SH_CONTRIB, ncon);
    stat_add(stats, ST_SH_RAMPS, nramp);
    stat_drops(stats, drops);
    if (INSTR) {
        stat_add(stats, ST_NV_SH, nvis);
        stat_add(stats, ST_NT_SH, ntri);
    }
}

// ---------------------------------------------------------------------------
// time-tiled splat: counting sort of the records by key, then one workgroup
// per key accumulating its tile in LDS
// ---------------------------------------------------------------------------
// Record slots are read kTileBatch per thread at a time (all loads issued
// before the first is used): one wave step then waits for one memory round
// trip instead of kTileBatch.  Slot i of step t of workgroup b is
// (t gridDim + b) kTileBlock kTileBatch + j kTileBlock + threadIdx, j < kTileBatch.
constexpr uint32_t kTileBatch = 16;

// cnt[key] += records of that key among the n_sq * stride slots.  Blocks
// count their share in LDS when the keys fit, then add each non-zero count
// to HBM once.  k_tile_scatter must run with the same grid (it recounts the
// same records per block).
__global__ void __launch_bounds__(kTileBlock) k_tile_count(const uint32_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ n_sq, uint32_t stride,
                                                           uint32_t* __restrict__ cnt, uint32_t K) {
    __shared__ uint32_t h[kLdsKeys];
    const size_t n = (size_t)*n_sq * stride;
    const bool in_lds = K <= kLdsKeys;
    if (in_lds) {
        for (uint32_t j = threadIdx.x; j < K; j += kTileBlock) h[j] = 0;
        __syncthreads();
    }
    constexpr size_t kStep = (size_t)kTileBlock * kTileBatch;
    for (size_t base = (size_t)blockIdx.x * kStep; base < n; base += (size_t)gridDim.x * kStep) {
        uint32_t k[kTileBatch];
#pragma unroll
        for (uint32_t j = 0; j < kTileBatch; ++j) {
            const size_t i = base + j * kTileBlock + threadIdx.x;
            k[j] = i < n ? keys[i] : kNoKey;
        }
#pragma unroll
        for (uint32_t j = 0; j < kTileBatch; ++j) {
            if (k[j] == kNoKey) continue;
            if (in_lds)
                atomicAdd(&h[k[j]], 1u);
            else
                atomicAdd(&cnt[k[j]], 1u);
        }
    }
    if (in_lds) {
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < K; j += kTileBlock)
            if (h[j]) atomicAdd(&cnt[j], h[j]);
    }
}

// Keys with more than kTileDirect records (crowded) get a range of the
// sorted array: off[k] = cur[k] = the sum of the crowded counts before k,
// off[K] = their total; a key with fewer has an empty range and cur[k] =
// kSparse (k_tile_scatter adds its ramps to HBM directly).  cnt is cleared
// for the next iteration; first[c] = the key holding sorted record
// c kTileChunk (for the chunks that start inside the total).  One block of
// 1024 threads.
constexpr uint32_t kSparse = 0xFFFFFFFFu;
__global__ void __launch_bounds__(1024) k_tile_scan(uint32_t* __restrict__ cnt, uint32_t* __restrict__ off,
                                                    uint32_t* __restrict__ cur, uint32_t K,
                                                    uint32_t* __restrict__ first) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < K; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t c = i < K ? cnt[i] : 0u;
        const uint32_t v = c > kTileDirect ? c : 0u;
        if (i < K) cnt[i] = 0;
        // inclusive scan within the wave
        uint32_t x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if ((int)lane >= o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        if (wave == 0) {
            uint32_t t = lane < 16 ? wsum[lane] : 0u;
            for (int o = 1; o < 16; o <<= 1) {
                const uint32_t y = __shfl_up(t, o);
                if ((int)lane >= o) t += y;
            }
            if (lane < 16) wsum[lane] = t;  // inclusive over waves
        }
        __syncthreads();
        const uint32_t excl = carry + (wave ? wsum[wave - 1] : 0u) + x - v;
        if (i < K) {
            off[i] = excl;
            cur[i] = v ? excl : kSparse;
            // the chunks whose first record is one of this key's [excl, excl + v)
            for (uint32_t c = (excl + kTileChunk - 1) / kTileChunk; c * kTileChunk < excl + v; ++c) first[c] = i;
        }
        __syncthreads();
        if (threadIdx.x == 0) carry += wsum[15];
        __syncthreads();
    }
    if (threadIdx.x == 0) off[K] = carry;
}

// The row and first bin of tile key k: its render is the last item with
// kbase <= k (kbase ascending, kbase[n_items] = K).
struct TileRow {
    unsigned long long* row;
    uint32_t nb, t0;
};
__device__ __forceinline__ TileRow tile_row(const InstDev* __restrict__ inst, const uint32_t* __restrict__ kbase,
                                            uint32_t n_items, uint32_t k) {
    uint32_t lo = 0, hi = n_items;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (kbase[mid] <= k)
            lo = mid;
        else
            hi = mid;
    }
    const InstDev& In = inst[lo];
    const uint32_t rel = k - kbase[lo];
    const uint32_t nb = In.P.n_bins;
    return TileRow{In.hist + 4 * (size_t)nb * (rel / In.n_tiles), nb, (rel % In.n_tiles) * kTile};
}

// A record's ramp added to HBM directly.
__device__ __forceinline__ void splat_record(const TileRow& r, uint2 p) {
    splat_span(r.row, r.nb, r.t0 + (p.y & 0xFFFFu), (int)((p.y >> 16) & 0xFFu), (int)(p.y >> 24), __uint_as_float(p.x));
}

// Records of crowded keys to out[] grouped by key: a block recounts its share
// (the records k_tile_count gave it), takes a range per key from cur with one
// atomic each, then places its records there (LDS atomics give the ranks; the
// order within a key does not matter to integer sums).  Records of sparse
// keys (cur[k] = kSparse) are added to HBM here.
__global__ void __launch_bounds__(kTileBlock) k_tile_scatter(const InstDev* __restrict__ inst,
                                                             const uint32_t* __restrict__ kbase, uint32_t n_items,
                                                             const uint32_t* __restrict__ keys,
                                                             const uint2* __restrict__ pay,
                                                             const uint32_t* __restrict__ n_sq, uint32_t stride,
                                                             uint32_t* __restrict__ cur, uint32_t K,
                                                             uint2* __restrict__ out) {
    __shared__ uint32_t h[kLdsKeys];
    const size_t n = (size_t)*n_sq * stride;
    constexpr size_t kStep = (size_t)kTileBlock * kTileBatch;
    const size_t step = (size_t)gridDim.x * kStep;
    const bool in_lds = K <= kLdsKeys;
    if (in_lds) {
        for (uint32_t j = threadIdx.x; j < K; j += kTileBlock) h[j] = 0;
        __syncthreads();
        for (size_t base = (size_t)blockIdx.x * kStep; base < n; base += step) {
            uint32_t k[kTileBatch];
#pragma unroll
            for (uint32_t j = 0; j < kTileBatch; ++j) {
                const size_t i = base + j * kTileBlock + threadIdx.x;
                k[j] = i < n ? keys[i] : kNoKey;
            }
#pragma unroll
            for (uint32_t j = 0; j < kTileBatch; ++j)
                if (k[j] != kNoKey) atomicAdd(&h[k[j]], 1u);
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < K; j += kTileBlock) {
            if (!h[j]) continue;
            const uint32_t c = cur[j];
            h[j] = c == kSparse ? kSparse : atomicAdd(&cur[j], h[j]);
        }
        __syncthreads();
    }
    for (size_t base = (size_t)blockIdx.x * kStep; base < n; base += step) {
        uint32_t k[kTileBatch];
        uint2 v[kTileBatch];
#pragma unroll
        for (uint32_t j = 0; j < kTileBatch; ++j) {
            const size_t i = base + j * kTileBlock + threadIdx.x;
            k[j] = i < n ? keys[i] : kNoKey;
        }
#pragma unroll
        for (uint32_t j = 0; j < kTileBatch; ++j) {
            const size_t i = base + j * kTileBlock + threadIdx.x;
            if (k[j] != kNoKey) v[j] = pay[i];
        }
#pragma unroll
        for (uint32_t j = 0; j < kTileBatch; ++j) {
            if (k[j] == kNoKey) continue;
            const uint32_t c = in_lds ? h[k[j]] : cur[k[j]];
            if (c == kSparse)
                splat_record(tile_row(inst, kbase, n_items, k[j]), v[j]);
            else
                out[in_lds ? atomicAdd(&h[k[j]], 1u) : atomicAdd(&cur[k[j]], 1u)] = v[j];
        }
    }
}

// 128-bit add into the LDS tile (lo at 2 i, hi at 2 i + 1)
__device__ __forceinline__ void lds_acc128(unsigned long long* tile, uint32_t i, I128 q) {
    const unsigned long long lo = q.lo;
    const unsigned long long old = atomicAdd(&tile[2 * i], lo);
    const unsigned long long hi = (unsigned long long)q.hi + ((old + lo) < old ? 1ull : 0ull);
    if (hi) atomicAdd(&tile[2 * i + 1], hi);
}

// The key whose records include sorted position p (off[0] <= p < off[K]):
// the last k with off[k] <= p.
__device__ __forceinline__ uint32_t key_at(const uint32_t* __restrict__ off, uint32_t K, uint32_t p) {
    uint32_t lo = 0, hi = K;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= p)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Workgroup b takes the sorted records [b kTileChunk, (b + 1) kTileChunk) (all
// of crowded keys) and walks the keys among them: the ramps of one (render,
// row, tile) are added into an LDS copy of the tile's bins
// [t0, t0 + kTile + kHalo) of both parts, which is then added to the row in
// HBM (bins < n_bins); a key with at most kTileDirect records in the chunk
// goes straight to HBM.  kbase: the first key of each render of the batch,
// ascending, kbase[n_items] = K.
__global__ void __launch_bounds__(kTileBlock) k_tile_splat(const InstDev* __restrict__ inst,
                                                           const uint32_t* __restrict__ kbase, uint32_t n_items,
                                                           const uint32_t* __restrict__ off, uint32_t K,
                                                           const uint32_t* __restrict__ first,
                                                           const uint2* __restrict__ pay) {
    constexpr uint32_t kSpan = kTile + kHalo;
    constexpr uint32_t kPer = kTileChunk / kTileBlock;  // records per thread and key, at most
    constexpr uint32_t kSplatBatch = 8;
    __shared__ unsigned long long acc[2 * kSpan * 2];  // [part][bin] x {lo, hi}
    const uint32_t total = off[K];
    const uint32_t c0 = blockIdx.x * kTileChunk;
    if (c0 >= total) return;
    const uint32_t c1 = min(c0 + kTileChunk, total);
    for (uint32_t b = c0; b < c1;) {
        const uint32_t k = b == c0 ? first[blockIdx.x] : key_at(off, K, b);
        const uint32_t e = min(off[k + 1], c1);
        const TileRow tr = tile_row(inst, kbase, n_items, k);
        unsigned long long* row = tr.row;
        const uint32_t nb = tr.nb, t0 = tr.t0;
        if (e - b <= kTileDirect) {
            // the chunk's share of the key is small: straight to HBM
            for (uint32_t i = b + threadIdx.x; i < e; i += kTileBlock) splat_record(tr, pay[i]);
        } else {
            for (uint32_t j = threadIdx.x; j < 4 * kSpan; j += kTileBlock) acc[j] = 0;
            __syncthreads();
            // this thread's records of the key, kSplatBatch loads at a time
            for (uint32_t jb = 0; jb < kPer && b + jb * kTileBlock < e; jb += kSplatBatch) {
                uint2 pv[kSplatBatch];
#pragma unroll
                for (uint32_t j = 0; j < kSplatBatch; ++j) {
                    const uint32_t i = b + (jb + j) * kTileBlock + threadIdx.x;
                    if (i < e) pv[j] = pay[i];
                }
#pragma unroll
                for (uint32_t j = 0; j < kSplatBatch; ++j) {
                    if (b + (jb + j) * kTileBlock + threadIdx.x >= e) break;
                    const uint2 p = pv[j];
                    const uint32_t o = p.y & 0xFFFFu;
                    const int ww = (int)((p.y >> 16) & 0xFFu), w = (int)(p.y >> 24);
                    const RampPoints q = ramp_points(__uint_as_float(p.x), w, ww);
                    unsigned long long* tp = acc + (q.neg ? 2 * kSpan : 0u);  // gets +w*S and +S
                    unsigned long long* tm = acc + (q.neg ? 0u : 2 * kSpan);  // gets (w+1)*S
                    lds_acc128(tp, o, q.aq);
                    if (t0 + o + 1 < nb) lds_acc128(tm, o + 1, q.p1);
                    if (t0 + o + (uint32_t)ww + 1 < nb) lds_acc128(tp, o + (uint32_t)ww + 1, q.sq);
                }
            }
            __syncthreads();
            for (uint32_t j = threadIdx.x; j < 2 * kSpan; j += kTileBlock) {
                const uint32_t part = j >= kSpan ? 1u : 0u, loc = j - part * kSpan;
                const unsigned long long vlo = acc[2 * j], vhi = acc[2 * j + 1];
                if ((vlo | vhi) && t0 + loc < nb)
                    acc128(row + 2 * (size_t)nb * part + 2 * (size_t)(t0 + loc), I128{vlo, (int64_t)vhi});
            }
            __syncthreads();  // the tile is cleared again for the next key
        }
        b = e;
    }
}

// p[0, n) = v, grid-stride (clears device memory without a memset node)
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_reset2(uint32_t* a, uint32_t* b) {
    if (threadIdx.x == 0) {
        *a = 0;
        *b = 0;
    }
}

// Ray regeneration: tops the current queue up to `cap` with the next rays of
// the segment table.  plan = {first ray, count} for k_generate; the cursor
// counts the rays taken so far.
__global__ void k_plan(uint32_t* cursor, uint32_t total, uint32_t cap, const uint32_t* n_q, uint32_t* plan) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t taken = *cursor, live = *n_q;
        const uint32_t room = cap > live ? cap - live : 0u;
        const uint32_t take = min(room, total - taken);
        plan[0] = taken;
        plan[1] = take;
        *cursor = taken + take;
    }
}

// End of a bounce-loop iteration: zeroes the counters c[i] for the bits i of
// mask and the n_seg refill counters (a kernel, not memset nodes, so a
// captured loop replays it correctly)
__global__ void k_reset_iter(uint32_t* c, uint32_t mask, uint32_t* seg, uint32_t n_seg) {
    if (threadIdx.x < 32 && ((mask >> threadIdx.x) & 1u)) c[threadIdx.x] = 0;
    for (uint32_t i = threadIdx.x; i < n_seg; i += blockDim.x) seg[i * kSegStride] = 0;
}

// Direct sound (src/Scene.cpp:299-311): one block per instance, one thread
// per (band, recorder).
template <int WIDTH>
__global__ void k_direct(const InstDev* __restrict__ inst, int* gstack, uint32_t gthreads,
                         unsigned long long* __restrict__ stats) {
    __shared__ int lds[kLdsMin * kBlock];
    const InstDev& In = inst[blockIdx.x];
    const RenderParams& P = In.P;
    uint32_t t = threadIdx.x;
    if (t >= 3 * P.n_rec) return;
    int band = (int)(t / P.n_rec);
    if (!((P.band_mask >> band) & 1u)) return;
    const RecParams& R = P.rec[t % P.n_rec];
    DirectRamp* o = In.direct + (size_t)band * P.n_channels + R.channel;
    int nch = R.kind == EAR_REC_STEREO ? 2 : 1;
    for (int c = 0; c < nch; ++c) o[c] = DirectRamp{0, 0, 0, 0.0f};
    if (P.src_kind != EAR_SRC_POINT || P.estimator != EAR_EST_NEE) return;
    V3 rp{R.px, R.py, R.pz}, sp{P.src_pos[0], P.src_pos[1], P.src_pos[2]};
    float bt = 0.0f;
    int bs = -1;
    uint32_t bo = 0, nv = 0, nt = 0;
    Traversal T = make_trav(In, gstack, gthreads);
    if (traverse<kLdsMin, kBlock, WIDTH, true, false>(T, rp, sub(sp, rp), bt, bs, bo, lds + threadIdx.x,
                                                      blockIdx.x * kBlock + t, nv, nt))
        return;
    V3 dist = sub(rp, sp);
    float len = length(dist);
    V3 dir = normalized(dist);
    float a = 1.0f / (4.0f * kPi * len * len) * powf_ear_l2(P.air[band], P.l2_air[band], len) * P.dry;
    float tt = len / 343.0f;
    float width = sqrtf(len);
    float ampl = 2.0f * a / width;
    int w = (int)ceilf(width);
    // a direct ramp cut at the end of the histogram counts as a dropped ramp
    auto cut = [&](int s) {
        const int64_t e = (int64_t)s + (int64_t)w;  // one past the ramp's last bin
        if (e > (int64_t)P.n_bins) {
            atomicAdd(stats + ST_DROPPED, 1ull);
            atomicMax(stats + ST_BINS, (unsigned long long)e);
        }
    };
    if (R.kind == EAR_REC_MONO) {
        o[0] = DirectRamp{1, (int)((double)tt * 44100.0), w, ampl};
        cut(o[0].s);
    } else {
        float dt = dot(dir, V3{R.rx, R.ry, R.rz});
        float td = R.head_size / 343.0f;
        float al = ampl, ar = ampl;
        float factor = powf_ear_l2(R.head_abs[band], R.l2_head_abs[band], fabsf(dt) * R.head_size);
        if (dt < 0.0f)
            ar *= factor * factor;
        else
            al *= factor * factor;
        o[0] = DirectRamp{1, (int)((double)(tt + dt * td) * 44100.0), w, al};
        o[1] = DirectRamp{1, (int)((double)(tt - dt * td) * 44100.0), w, ar};
        cut(o[0].s);
        cut(o[1].s);
    }
}

// Two inclusive prefix sums over a second-difference row (128-bit, exact),
// then value = (raw 2^-80 * inv_n + direct ramp) * gain2.  One block per row;
// row r of the table is {instance, band, channel, kind}.
constexpr int kFinBlock = 1024;
__global__ void __launch_bounds__(kFinBlock) k_finalize(const InstDev* __restrict__ inst, const uint4* __restrict__ rows) {
    __shared__ uint64_t slo[kFinBlock];
    __shared__ int64_t shi[kFinBlock];
    const uint4 rw = rows[blockIdx.x];
    const InstDev& In = inst[rw.x];
    const RenderParams& P = In.P;
    const int band = (int)rw.y;
    const uint32_t ch = rw.z, kind = rw.w, nb = P.n_bins, n_ch = P.n_channels, kinds = P.hist_kinds;
    size_t roff = (((size_t)band * n_ch + ch) * kinds + kind) * 2 * nb;
    unsigned long long* c = In.hist + 2 * roff;      // "+" part, receives the sums
    unsigned long long* cm = c + 2 * (size_t)nb;     // "-" part
    uint32_t per = (nb + kFinBlock - 1) / kFinBlock;
    uint32_t b = threadIdx.x * per, e = min(nb, b + per);
    // pass 0 sums (+) - (-), the second differences; pass 1 their prefix sums
    auto elem = [&](int pass, uint32_t j) {
        I128 v{c[2 * j], (int64_t)c[2 * j + 1]};
        if (pass == 0) v = i128_add(v, i128_neg(I128{cm[2 * j], (int64_t)cm[2 * j + 1]}));
        return v;
    };
    for (int pass = 0; pass < 2; ++pass) {
        I128 sum{0, 0};
        for (uint32_t j = b; j < e; ++j) sum = i128_add(sum, elem(pass, j));
        slo[threadIdx.x] = sum.lo;
        shi[threadIdx.x] = sum.hi;
        __syncthreads();
        for (int off = 1; off < kFinBlock; off <<= 1) {
            I128 v{0, 0};
            if ((int)threadIdx.x >= off) v = I128{slo[threadIdx.x - off], shi[threadIdx.x - off]};
            __syncthreads();
            I128 cur = i128_add(I128{slo[threadIdx.x], shi[threadIdx.x]}, v);
            slo[threadIdx.x] = cur.lo;
            shi[threadIdx.x] = cur.hi;
            __syncthreads();
        }
        I128 run = i128_add(I128{slo[threadIdx.x], shi[threadIdx.x]}, i128_neg(sum));
        for (uint32_t j = b; j < e; ++j) {
            run = i128_add(run, elem(pass, j));
            c[2 * j] = run.lo;
            c[2 * j + 1] = (unsigned long long)run.hi;
        }
        __syncthreads();
    }
}

// Per (instance, band, channel) and bin: the sign-split rows (if any) become
// the signed and |.| sums, then the finished track: 2^-80, 1/N, the direct
// sound (src/Scene.cpp:299-311) and gain^2 (src/Scene.cpp:313-316).
__global__ void __launch_bounds__(256) k_values(const InstDev* __restrict__ inst, const uint4* __restrict__ rows) {
    const uint4 rw = rows[blockIdx.y];
    const InstDev& In = inst[rw.x];
    const RenderParams& P = In.P;
    const int band = (int)rw.y;
    const uint32_t ch = rw.z, nb = P.n_bins, n_ch = P.n_channels, kinds = P.hist_kinds;
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= nb) return;
    const size_t r0 = (((size_t)band * n_ch + ch) * kinds) * nb;  // value rows
    unsigned long long* c0 = In.hist + 4 * r0;                      // kind 0 "+" part
    I128 v[2];
    v[0] = I128{c0[2 * j], (int64_t)c0[2 * j + 1]};
    if (kinds > 1) {
        unsigned long long* c1 = c0 + 4 * (size_t)nb;  // kind 1 "+" part
        const I128 pos = v[0], neg = I128{c1[2 * j], (int64_t)c1[2 * j + 1]};
        v[0] = i128_add(pos, i128_neg(neg));
        v[1] = i128_add(pos, neg);
        c0[2 * j] = v[0].lo;
        c0[2 * j + 1] = (unsigned long long)v[0].hi;
        c1[2 * j] = v[1].lo;
        c1[2 * j + 1] = (unsigned long long)v[1].hi;
    }
    if (!In.value) return;
    const DirectRamp dr = In.direct[(size_t)band * n_ch + ch];
    const double S = dr.visible ? (double)dr.a / (double)dr.w : 0.0;
    for (uint32_t kind = 0; kind < kinds; ++kind) {
        double x = (double)v[kind].hi * 18446744073709551616.0 + (double)v[kind].lo;
        x = x * 0x1p-80;  // 2^-EAR_FRAC_BITS
        x = x * In.inv_n;
        if (dr.visible && (int)j >= dr.s && (int)j < dr.s + dr.w) {
            const double dv = (double)(dr.w - ((int)j - dr.s)) * S;
            x += kind == 1 ? fabs(dv) : dv;
        }
        In.value[r0 + (size_t)kind * nb + j] = x * In.gain2;
    }
}

// batched queries for tests and host tools
template <int WIDTH, bool ANY>
__global__ void __launch_bounds__(kBlock) k_query(Traversal T, uint32_t n, const float* __restrict__ a,
                                                   const float* __restrict__ b, float* __restrict__ tout,
                                                   int32_t* __restrict__ iout, uint8_t* __restrict__ occ) {
    __shared__ int lds[kLdsMin * kBlock];
    uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    uint32_t nv = 0, nt = 0;
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        uint32_t i = base + threadIdx.x;
        if (i >= n) continue;
        V3 o{a[3 * i], a[3 * i + 1], a[3 * i + 2]};
        V3 x{b[3 * i], b[3 * i + 1], b[3 * i + 2]};
        float bt = 1000000.0f;
        int bs = -1;
        uint32_t bo = 0;
        if (ANY) {
            occ[i] = traverse<kLdsMin, kBlock, WIDTH, true, false>(T, o, sub(x, o), bt, bs, bo, lds + threadIdx.x,
                                                                   gtid, nv, nt)
                         ? 1
                         : 0;
        } else {
            bool h = traverse<kLdsMin, kBlock, WIDTH, false, false>(T, o, x, bt, bs, bo, lds + threadIdx.x, gtid, nv,
                                                                    nt);
            tout[i] = h ? bt : 0.0f;
            iout[i] = h ? (int32_t)bo : -1;
        }
    }
}

// Traversal kernels are instantiated for every (block, LDS stack part, tree
// width, instrumentation, batch) combination; the render picks one at run time.
struct TravCfg {
    int block, lds, width;
    bool instr, batch, fetch;
    bool tiled;  // shadow kernels: write the time-tiled splat's records
};

using ExtendFn = void (*)(const InstDev*, int*, uint32_t, Queue, const uint32_t*, int2*, unsigned long long*, Fetch);
using ShadowFn = void (*)(const InstDev*, int*, uint32_t, ShadowQ, const uint32_t*, unsigned long long*, Fetch,
                         TileOut);

template <int B, int L, int W, bool BA>
void pick_t(bool instr, bool fetch, bool tiled, ExtendFn* e, ShadowFn* s) {
    if (fetch) {
        *e = instr ? (ExtendFn)k_extend_f<B, L, W, true, BA> : (ExtendFn)k_extend_f<B, L, W, false, BA>;
        if (tiled)
            *s = instr ? (ShadowFn)k_shadow_f<B, L, W, true, BA, true> : (ShadowFn)k_shadow_f<B, L, W, false, BA, true>;
        else
            *s = instr ? (ShadowFn)k_shadow_f<B, L, W, true, BA, false>
                       : (ShadowFn)k_shadow_f<B, L, W, false, BA, false>;
    } else {
        *e = instr ? (ExtendFn)k_extend<B, L, W, true, BA> : (ExtendFn)k_extend<B, L, W, false, BA>;
        if (tiled)
            *s = instr ? (ShadowFn)k_shadow<B, L, W, true, BA, true> : (ShadowFn)k_shadow<B, L, W, false, BA, true>;
        else
            *s = instr ? (ShadowFn)k_shadow<B, L, W, true, BA, false> : (ShadowFn)k_shadow<B, L, W, false, BA, false>;
    }
}

template <bool BA>
void pick_b(const TravCfg& c, ExtendFn* e, ShadowFn* s) {
    int key = (c.block >= 256 ? 4 : 0) | (c.lds >= 32 ? 2 : 0) | (c.width == 4 ? 1 : 0);
    switch (key) {
        case 0: pick_t<128, 16, 2, BA>(c.instr, c.fetch, c.tiled, e, s); break;
        case 1: pick_t<128, 16, 4, BA>(c.instr, c.fetch, c.tiled, e, s); break;
        case 2: pick_t<128, 32, 2, BA>(c.instr, c.fetch, c.tiled, e, s); break;
        case 3: pick_t<128, 32, 4, BA>(c.instr, c.fetch, c.tiled, e, s); break;
        case 4: pick_t<256, 16, 2, BA>(c.instr, c.fetch, c.tiled, e, s); break;
        case 5: pick_t<256, 16, 4, BA>(c.instr, c.fetch, c.tiled, e, s); break;
        case 6: pick_t<256, 32, 2, BA>(c.instr, c.fetch, c.tiled, e, s); break;
        default: pick_t<256, 32, 4, BA>(c.instr, c.fetch, c.tiled, e, s); break;
    }
}

using TailFn = void (*)(const InstDev*, int*, uint32_t, Queue, const uint32_t*, unsigned long long*);

template <int B, int L>
TailFn tail_for(int width, bool batch) {
    if (width == 4) return batch ? (TailFn)k_tail<B, L, 4, true> : (TailFn)k_tail<B, L, 4, false>;
    return batch ? (TailFn)k_tail<B, L, 2, true> : (TailFn)k_tail<B, L, 2, false>;
}

TailFn pick_tail(const TravCfg& c) {
    if (c.block == 128) return c.lds == 32 ? tail_for<128, 32>(c.width, c.batch) : tail_for<128, 16>(c.width, c.batch);
    return c.lds == 32 ? tail_for<256, 32>(c.width, c.batch) : tail_for<256, 16>(c.width, c.batch);
}

void pick_kernels(const TravCfg& c, ExtendFn* e, ShadowFn* s) {
    if (c.batch)
        pick_b<true>(c, e, s);
    else
        pick_b<false>(c, e, s);
}

// Blocks per CU that can be resident for kernel `fn` (0 on error).
int resident_blocks(const void* fn, int block) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

uint32_t persistent_grid(Ctx* c, uint32_t work, uint32_t waves_per_cu, uint32_t block = kBlock) {
    uint32_t blocks_needed = (work + block - 1) / block;
    uint32_t cap = (uint32_t)c->num_cus * std::max<uint32_t>(1, waves_per_cu) / std::max<uint32_t>(1, block / 64);
    return std::max<uint32_t>(1, std::min(blocks_needed, std::max<uint32_t>(1, cap)));
}

uint32_t grid_for(uint32_t work, uint32_t cap_blocks, uint32_t block) {
    uint32_t blocks_needed = (work + block - 1) / block;
    return std::max<uint32_t>(1, std::min(blocks_needed, std::max<uint32_t>(1, cap_blocks)));
}

// Block cap of a kernel that loops over its queue with the grid's stride (so
// every block gets the same share): three rounds of the blocks one CU holds
// at once.  A grid of whole rounds leaves no CU a last partial round to run
// alone (32 waves per CU of a kernel that fits 24 ran its last quarter of
// blocks at a third of the occupancy).  waves_per_cu, when set, overrides.
uint32_t rounds_cap(Ctx* c, const void* fn, uint32_t waves_per_cu, int block) {
    if (waves_per_cu) return (uint32_t)c->num_cus * waves_per_cu / (uint32_t)std::max(1, block / 64);
    return (uint32_t)c->num_cus * (uint32_t)std::max(1, resident_blocks(fn, block)) * 3u;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
};

float tri_area_host(const float* v) {
    // Triangle::calcArea (src/Triangle.cpp:26-30): |edge(0) x edge(1)| / 2,
    // edge(i) = v[(i+1)%3] - v[i]
    V3 a = sub(V3{v[3], v[4], v[5]}, V3{v[0], v[1], v[2]});
    V3 b = sub(V3{v[6], v[7], v[8]}, V3{v[3], v[4], v[5]});
    return length(cross(a, b)) / 2.0f;
}

// Validates one render description and fills its RenderParams.
ear_status make_params(Ctx* c, const ear_render_desc* D, RenderParams* P) {
    if (!D) return EAR_ERR_INVALID_ARGUMENT;
    if ((D->band_mask & 7u) == 0 || (D->band_mask & ~7u)) {
        c->last_error = "band_mask must select bands among bits 0..2";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if (D->hist_bins == 0 || D->n_recorders == 0 || !D->recorders) {
        c->last_error = "hist_bins and recorders are required";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if (D->n_recorders > (uint32_t)kMaxRecorders) {
        c->last_error = "too many recorders";
        return EAR_ERR_CAPACITY;
    }
    if (D->source.kind != EAR_SRC_POINT && D->source.kind != EAR_SRC_MESH) return EAR_ERR_INVALID_ARGUMENT;
    if (D->source.kind == EAR_SRC_MESH && (!D->source.mesh_v || D->source.mesh_n_tris == 0)) {
        c->last_error = "mesh source without triangles";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if ((uint64_t)D->hist_bins * 3ull * 2ull * kMaxRecorders > (1ull << 36)) {
        c->last_error = "histogram too long";
        return EAR_ERR_CAPACITY;
    }
    for (int b = 0; b < 3; ++b)
        if (!std::isfinite(D->air[b])) return EAR_ERR_INVALID_ARGUMENT;
    if (D->max_bounces > kMaxBounces) {
        // the bounce number travels in 16 bits of a queue entry; past them it
        // would wrap, and a path in a lossless room would never end
        c->last_error = "max_bounces above 65535";
        return EAR_ERR_CAPACITY;
    }
    if (D->estimator > EAR_EST_SPHERE) {
        c->last_error = "unknown estimator";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    std::memset(P, 0, sizeof(*P));
    P->estimator = D->estimator;
    P->seed = D->seed;
    P->instance = D->instance;
    for (int b = 0; b < 3; ++b) {
        P->ctx_word[b] = (uint32_t)b | ((D->sound & 0xFFu) << 8) | ((uint32_t)(D->keyframe + 1) << 16);
        P->air[b] = D->air[b];
        P->l2_air[b] = pow_base_l2(D->air[b]);
    }
    P->dry = D->dry;
    P->cutoff = D->energy_cutoff;
    P->max_bounces = D->max_bounces;
    P->n_bins = D->hist_bins;
    P->n_rec = D->n_recorders;
    P->n_band_slots = 3;
    P->band_mask = D->band_mask;
    P->hist_kinds = (D->flags & EAR_FLAG_ABS_HIST) ? 2 : 1;
    P->src_kind = D->source.kind;
    for (int k = 0; k < 3; ++k) P->src_pos[k] = D->source.pos[k];
    P->flags = D->flags;
    uint32_t nch = 0;
    for (uint32_t r = 0; r < D->n_recorders; ++r) {
        const ear_recorder& R = D->recorders[r];
        if (R.kind != EAR_REC_MONO && R.kind != EAR_REC_STEREO) return EAR_ERR_INVALID_ARGUMENT;
        RecParams& Q = P->rec[r];
        Q.px = R.pos[0];
        Q.py = R.pos[1];
        Q.pz = R.pos[2];
        Q.rx = R.right_ear[0];
        Q.ry = R.right_ear[1];
        Q.rz = R.right_ear[2];
        Q.head_size = R.head_size;
        for (int b = 0; b < 3; ++b) {
            Q.head_abs[b] = R.head_abs[b];
            Q.l2_head_abs[b] = pow_base_l2(R.head_abs[b]);
        }
        Q.kind = R.kind;
        Q.channel = (int32_t)nch;
        nch += R.kind == EAR_REC_STEREO ? 2 : 1;
        if (D->estimator == EAR_EST_SPHERE) {
            // spheres: mono at pos; stereo head_size / 2 either side along
            // the unit right-ear direction (left channel first)
            const float rad = R.radius;
            const float rn = sqrtf(R.right_ear[0] * R.right_ear[0] + R.right_ear[1] * R.right_ear[1] +
                                   R.right_ear[2] * R.right_ear[2]);
            if (!(rad > 0.0f) || !std::isfinite(rad) || (R.kind == EAR_REC_STEREO && !(rn > 0.0f))) {
                c->last_error = "sphere estimator needs radius > 0 (and a right-ear direction for stereo)";
                return EAR_ERR_INVALID_ARGUMENT;
            }
            Q.r2 = rad * rad;
            Q.inv_vol = 3.0f / (4.0f * kPi * rad * rad * rad);
            for (int k = 0; k < 3; ++k) Q.cen[0][k] = Q.cen[1][k] = R.pos[k];
            if (R.kind == EAR_REC_STEREO) {
                const float h = (0.5f * R.head_size) / rn;
                for (int k = 0; k < 3; ++k) {
                    Q.cen[0][k] = R.pos[k] - R.right_ear[k] * h;
                    Q.cen[1][k] = R.pos[k] + R.right_ear[k] * h;
                }
            }
        }
    }
    P->n_channels = nch;
    bool want_log = (D->flags & EAR_FLAG_TRACE_LOG) && D->trace_rays && D->trace_segs;
    P->trace_rays = want_log ? D->trace_rays : 0;
    P->trace_segs = want_log ? D->trace_segs : 0;
    return EAR_OK;
}

}  // namespace

ear_status run_query_closest(Ctx* c, DeviceScene* s, uint32_t n, const float* o, const float* d, float* t,
                             int32_t* tri) {
    if (n == 0) return EAR_OK;
    if (!o || !d || !t || !tri) return EAR_ERR_INVALID_ARGUMENT;
    DevBuf da(c), db(c), dt(c), di(c), gs(c);
    TRY(da.alloc((size_t)n * 12));
    TRY(db.alloc((size_t)n * 12));
    TRY(dt.alloc((size_t)n * 4));
    TRY(di.alloc((size_t)n * 4));
    uint32_t grid = persistent_grid(c, n, 16);
    uint32_t gthreads = grid * kBlock;
    TRY(gs.alloc((size_t)(kStackDepth - kLdsMin) * gthreads * 4));
    hipStream_t st = c->stream;
    EAR_HIP_TRY(c, hipMemcpyAsync(da.p, o, (size_t)n * 12, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(db.p, d, (size_t)n * 12, hipMemcpyHostToDevice, st));
    Traversal T{s->nodes, s->nodes4, s->tris, s->n_tris, gs.as<int>(), gthreads};
    if (s->width == 4)
        k_query<4, false><<<grid, kBlock, 0, st>>>(T, n, da.as<float>(), db.as<float>(), dt.as<float>(),
                                                   di.as<int32_t>(), nullptr);
    else
        k_query<2, false><<<grid, kBlock, 0, st>>>(T, n, da.as<float>(), db.as<float>(), dt.as<float>(),
                                                   di.as<int32_t>(), nullptr);
    EAR_HIP_TRY(c, hipGetLastError());
    EAR_HIP_TRY(c, hipMemcpyAsync(t, dt.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(tri, di.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    EAR_HIP_TRY(c, hipStreamSynchronize(st));
    return EAR_OK;
}

ear_status run_query_occluded(Ctx* c, DeviceScene* s, uint32_t n, const float* p, const float* q, uint8_t* occ) {
    if (n == 0) return EAR_OK;
    if (!p || !q || !occ) return EAR_ERR_INVALID_ARGUMENT;
    DevBuf da(c), db(c), dk(c), gs(c);
    TRY(da.alloc((size_t)n * 12));
    TRY(db.alloc((size_t)n * 12));
    TRY(dk.alloc((size_t)n));
    uint32_t grid = persistent_grid(c, n, 16);
    uint32_t gthreads = grid * kBlock;
    TRY(gs.alloc((size_t)(kStackDepth - kLdsMin) * gthreads * 4));
    hipStream_t st = c->stream;
    EAR_HIP_TRY(c, hipMemcpyAsync(da.p, p, (size_t)n * 12, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(db.p, q, (size_t)n * 12, hipMemcpyHostToDevice, st));
    Traversal T{s->nodes, s->nodes4, s->tris, s->n_tris, gs.as<int>(), gthreads};
    if (s->width == 4)
        k_query<4, true><<<grid, kBlock, 0, st>>>(T, n, da.as<float>(), db.as<float>(), nullptr, nullptr,
                                                  dk.as<uint8_t>());
    else
        k_query<2, true><<<grid, kBlock, 0, st>>>(T, n, da.as<float>(), db.as<float>(), nullptr, nullptr,
                                                  dk.as<uint8_t>());
    EAR_HIP_TRY(c, hipGetLastError());
    EAR_HIP_TRY(c, hipMemcpyAsync(occ, dk.p, n, hipMemcpyDeviceToHost, st));
    EAR_HIP_TRY(c, hipStreamSynchronize(st));
    return EAR_OK;
}

// A batch of renders in one wavefront: every (instance, band) context's rays
// share the queues; per-instance scene data, parameters and histograms come
// from the device instance table.  A plain render is a batch of one.
ear_status run_batch(Ctx* c, uint32_t n_items, DeviceScene* const* scenes, const ear_render_desc* const* descs,
                     ear_hist* const* outs, ear_stats* st_out, Arena* arena) {
    if (n_items == 0) return EAR_ERR_INVALID_ARGUMENT;
    // no-allocation render: every buffer is carved from the arena, device
    // memory is cleared by a kernel, no events, pinned memory or graphs
    const bool noalloc = arena != nullptr;
    const bool batch = n_items > 1;
    const int width = (int)scenes[0]->width;
    std::vector<InstDev> inst(n_items);
    std::vector<std::vector<float>> src_area(n_items), src_nrm(n_items);
    uint64_t total_rays = 0;
    uint32_t max_bounces = 0;
    bool instr = false, any_log = false;
    size_t hist_cells = 0, value_cells = 0, direct_cells = 0, log_hits = 0, log_nseg = 0;
    std::vector<size_t> hist_off(n_items), value_off(n_items), direct_off(n_items), logh_off(n_items),
        logn_off(n_items);
    bool any_value = false;
    size_t max_recs = 0;
    uint32_t any_mesh = 0;
    // time-tiled splat: keys (render, row, tile) of every render, and the
    // most channels of one recorder (records per shadow request)
    uint64_t n_keys = 0;
    uint32_t tile_stride = 1;
    std::vector<uint32_t> kbase(n_items + 1, 0);
    for (uint32_t i = 0; i < n_items; ++i) {
        if (!scenes[i] || !descs[i]) return EAR_ERR_INVALID_ARGUMENT;
        if ((int)scenes[i]->width != width) {
            c->last_error = "all scenes of a batch must have the same BVH width";
            return EAR_ERR_INVALID_ARGUMENT;
        }
        InstDev& I = inst[i];
        std::memset(&I, 0, sizeof(I));
        TRY(make_params(c, descs[i], &I.P));
        const ear_render_desc* D = descs[i];
        const DeviceScene* s = scenes[i];
        I.nodes = s->nodes;
        I.nodes4 = s->nodes4;
        I.tris = s->tris;
        I.nm = s->nm;
        I.mat_band = s->mat_band;
        I.n_tris = s->n_tris;
        I.inv_n = (double)(1.0f / (float)std::max<uint32_t>(1, D->n_rays));  // Recorder::Multiply(1.0f / amount)
        I.gain2 = (double)(D->source.gain * D->source.gain);                  // Recorder::Multiply(gain * gain)
        if (D->source.kind == EAR_SRC_MESH) {
            any_mesh = 1;
            uint32_t m = D->source.mesh_n_tris;
            src_area[i].resize(m);
            src_nrm[i].resize(3 * (size_t)m);
            float total = 0.0f;
            for (uint32_t k = 0; k < m; ++k) {
                const float* v = D->source.mesh_v + 9 * (size_t)k;
                src_area[i][k] = tri_area_host(v);
                total += src_area[i][k];
                V3 nn = normalized(cross(sub(V3{v[3], v[4], v[5]}, V3{v[0], v[1], v[2]}),
                                         sub(V3{v[6], v[7], v[8]}, V3{v[0], v[1], v[2]})));
                src_nrm[i][3 * k] = nn.x;
                src_nrm[i][3 * k + 1] = nn.y;
                src_nrm[i][3 * k + 2] = nn.z;
            }
            I.P.src_n_tris = m;
            I.P.src_total_area = total;
        }
        I.n_tiles = (I.P.n_bins + kTile - 1) / kTile;
        I.key_base = (uint32_t)std::min<uint64_t>(n_keys, 0xFFFFFFFFull);
        kbase[i] = I.key_base;
        n_keys += (uint64_t)3 * I.P.n_channels * I.P.hist_kinds * I.n_tiles;
        for (uint32_t r = 0; r < D->n_recorders; ++r)
            if (D->recorders[r].kind == EAR_REC_STEREO) tile_stride = 2;
        int nb = 0;
        for (int b = 0; b < 3; ++b) nb += (D->band_mask >> b) & 1u;
        total_rays += (uint64_t)D->n_rays * nb;
        max_bounces = std::max(max_bounces, D->max_bounces);
        instr = instr || (D->flags & EAR_FLAG_INSTRUMENT);
        max_recs = std::max<size_t>(max_recs, D->n_recorders);
        size_t cells = (size_t)3 * I.P.n_channels * I.P.hist_kinds * I.P.n_bins;
        hist_off[i] = hist_cells;
        hist_cells += 2 * cells;  // "+" and "-" parts
        value_off[i] = value_cells;
        if (outs[i] && (outs[i]->value || outs[i]->abs_value)) {
            value_cells += cells;
            any_value = true;
        }
        direct_off[i] = direct_cells;
        direct_cells += (size_t)3 * I.P.n_channels;
        logh_off[i] = log_hits;
        logn_off[i] = log_nseg;
        if (I.P.trace_rays) {
            log_hits += (size_t)3 * I.P.trace_rays * I.P.trace_segs;
            log_nseg += (size_t)3 * I.P.trace_rays;
            any_log = true;
        }
    }

    if (total_rays > 0xFFFFFFFFull) {
        c->last_error = "more than 2^32 - 1 rays in one call";
        return EAR_ERR_CAPACITY;
    }
    const ear_tuning& tu = c->tuning;
    kbase[n_items] = (uint32_t)std::min<uint64_t>(n_keys, 0xFFFFFFFFull);
    const uint32_t K = kbase[n_items];
    uint32_t chunk = tu.ray_chunk ? tu.ray_chunk : (1u << 23);
    uint64_t cap64 = std::min<uint64_t>((uint64_t)chunk * 3, std::max<uint64_t>(1, total_rays));
    uint32_t cap_rays = (uint32_t)cap64;
    // the shadow kernels' splat: records binned by time tile and accumulated
    // in LDS, or direct global atomics (always in the no-allocation render:
    // the records would not fit its region).  By default the tiled splat when
    // the call's keys fit the LDS counters and the queues hold at least 2^22
    // rays (on the 200k hall, 3 x 10M rays: 463.9 against 468.5 ms per
    // render); otherwise the atomics: a batch of many renders has too many
    // keys, and smaller renders do not repay the tile kernels' fixed cost per
    // iteration (the 500k hall at 3 x 781k rays: 9% slower tiled).
    const bool tiled = !noalloc && n_keys < 0x7FFFFFFFull &&
                       (tu.splat == EAR_SPLAT_TILED ||
                        (tu.splat == EAR_SPLAT_DEFAULT && n_keys <= kLdsKeys && cap_rays >= (1u << 22)));
    // walks vary in length; refilling idle lanes keeps waves busy (both
    // kernels by default)
    const bool fetch_e = tu.traversal != EAR_TRAVERSAL_STATIC;
    const bool fetch_s = tu.traversal == EAR_TRAVERSAL_FETCH_ALL || tu.traversal == EAR_TRAVERSAL_DEFAULT;
    const TravCfg cfg_e{tu.block_extend == 128 ? 128 : 256, tu.stack_lds >= 32 ? 32 : 16, width, instr, batch, fetch_e,
                        false};
    const TravCfg cfg_s{tu.block_shadow == 128 ? 128 : 256, tu.stack_lds >= 32 ? 32 : 16, width, instr, batch, fetch_s,
                        tiled};
    ExtendFn fe, fe_unused;
    ShadowFn fs, fs_unused;
    pick_kernels(cfg_e, &fe, &fs_unused);
    pick_kernels(cfg_s, &fe_unused, &fs);
    auto cap_for = [&](const void* fn, int block) -> uint32_t {
        uint32_t cap;
        if (tu.waves_per_cu) {
            cap = (uint32_t)c->num_cus * tu.waves_per_cu / std::max(1, block / 64);
        } else {
            int nb = resident_blocks(fn, block);
            cap = (uint32_t)c->num_cus * (uint32_t)std::max(1, nb);
        }
        // the no-allocation render keeps its overflow stacks small: at most
        // kNoallocWaves waves per CU in the traversal grids
        if (noalloc) cap = std::min<uint32_t>(cap, (uint32_t)c->num_cus * kNoallocWaves / std::max(1, block / 64));
        return std::max<uint32_t>(1, cap);
    };
    const uint32_t cap_e = std::max<uint32_t>(1, cap_for((const void*)fe, cfg_e.block));
    const uint32_t cap_s = std::max<uint32_t>(1, cap_for((const void*)fs, cfg_s.block));

    hipStream_t st = c->stream;
    Arena* ar = arena;
    DevBuf qo[2] = {DevBuf(c, ar), DevBuf(c, ar)}, qd[2] = {DevBuf(c, ar), DevBuf(c, ar)},
           qm[2] = {DevBuf(c, ar), DevBuf(c, ar)};
    DevBuf hits(c, ar), so(c, ar), sm(c, ar), counters(c, ar), stats(c, ar), hist(c, ar), gs(c, ar), direct(c, ar),
        value(c, ar), logh(c, ar), logn(c, ar), dinst(c, ar), dsegs(c, ar), dsegoff(c, ar), drows(c, ar),
        dvrows(c, ar), msrc(c, ar), segc(c, ar), tkey(c, ar), tpay(c, ar), tsort(c, ar), tcnt(c, ar), toff(c, ar),
        tcur(c, ar), tkb(c, ar), tfirst(c, ar);
    // device memory that must start at a value: hipMemsetAsync, or in the
    // no-allocation render a fill kernel (no memset nodes if a caller
    // captures the stream)
    auto clear = [&](DevBuf& b, uint8_t byte) -> ear_status {
        if (!noalloc) {
            EAR_HIP_TRY(c, hipMemsetAsync(b.p, byte, b.n, st));
            return EAR_OK;
        }
        const size_t words = b.n / 4;  // every cleared buffer is a whole number of words
        const uint32_t w = 0x01010101u * byte;
        k_fill_u32<<<persistent_grid(c, (uint32_t)std::min<size_t>(words, 0xFFFFFFFFu), 8), kBlock, 0, st>>>(
            b.as<uint32_t>(), words, w);
        EAR_HIP_TRY(c, hipGetLastError());
        return EAR_OK;
    };
    // [0] queue 0, [1] queue 1, [2] shadow queue, [3] generation cursor,
    // [6..7] plan {first, count}
    TRY(counters.alloc(16 * 4));
    // refill counters of the closest-hit and shadow kernels, one per line
    TRY(segc.alloc(2 * kMaxSeg * kSegStride * 4));
    TRY(clear(segc, 0));
    TRY(stats.alloc(ST_N * 8 * kStatCopies));
    TRY(hist.alloc(hist_cells * 16));
    uint32_t gthreads = std::max(cap_e * (uint32_t)cfg_e.block, cap_s * (uint32_t)cfg_s.block);
    gthreads = std::max<uint32_t>(gthreads, n_items * kBlock);
    // overflow stack entries per persistent thread: the full depth, or in the
    // no-allocation render the bound of the scenes' own trees (a 4-wide node
    // visit pushes at most 3 entries, a 2-wide one at most 1, per level)
    size_t gs_entries = kStackDepth - kLdsMin;
    if (noalloc) {
        uint32_t need = 1;
        for (uint32_t i = 0; i < n_items; ++i)
            need = std::max(need, scenes[i]->width == 4 ? 3 * scenes[i]->depth4 + 1 : scenes[i]->max_depth);
        gs_entries = need > (uint32_t)kLdsMin ? need - kLdsMin : 1;
    }
    TRY(gs.alloc(gs_entries * gthreads * 4));
    TRY(direct.alloc(direct_cells * sizeof(DirectRamp)));
    if (any_value) TRY(value.alloc(value_cells * 8));
    if (any_log) {
        TRY(logh.alloc(log_hits * 4));
        TRY(logn.alloc(log_nseg * 4));
        TRY(clear(logh, 0xFE));  // 0xFEFEFEFE read back as -2
        TRY(clear(logn, 0));
    }
    // mesh-source tables, one allocation
    size_t msrc_floats = 0;
    std::vector<size_t> msrc_off(n_items);
    for (uint32_t i = 0; i < n_items; ++i) {
        msrc_off[i] = msrc_floats;
        msrc_floats += 9 * (size_t)inst[i].P.src_n_tris + src_area[i].size() + src_nrm[i].size();
    }
    if (msrc_floats) TRY(msrc.alloc(msrc_floats * 4));
    for (uint32_t i = 0; i < n_items; ++i) {
        InstDev& I = inst[i];
        I.hist = hist.as<unsigned long long>() + 2 * hist_off[i];
        I.direct = direct.as<DirectRamp>() + direct_off[i];
        I.value = (outs[i] && (outs[i]->value || outs[i]->abs_value)) ? value.as<double>() + value_off[i] : nullptr;
        I.log_hits = I.P.trace_rays ? logh.as<int32_t>() + logh_off[i] : nullptr;
        I.log_nseg = I.P.trace_rays ? logn.as<uint32_t>() + logn_off[i] : nullptr;
        if (I.P.src_n_tris) {
            float* base = msrc.as<float>() + msrc_off[i];
            size_t m = I.P.src_n_tris;
            EAR_HIP_TRY(c, hipMemcpyAsync(base, descs[i]->source.mesh_v, m * 36, hipMemcpyHostToDevice, st));
            EAR_HIP_TRY(c, hipMemcpyAsync(base + 9 * m, src_area[i].data(), m * 4, hipMemcpyHostToDevice, st));
            EAR_HIP_TRY(c, hipMemcpyAsync(base + 10 * m, src_nrm[i].data(), m * 12, hipMemcpyHostToDevice, st));
            I.M = MeshSrc{base, base + 9 * m, base + 10 * m};
        }
    }
    TRY(dinst.alloc(sizeof(InstDev) * n_items));
    EAR_HIP_TRY(c, hipMemcpyAsync(dinst.p, inst.data(), sizeof(InstDev) * n_items, hipMemcpyHostToDevice, st));
    TRY(clear(hist, 0));
    TRY(clear(stats, 0));
    TRY(clear(counters, 0));

    // every (instance, band) context as a run of rays, cut into chunks
    std::vector<GenSeg> all;
    for (uint32_t i = 0; i < n_items; ++i)
        for (int b = 0; b < 3; ++b)
            if ((descs[i]->band_mask >> b) & 1u && descs[i]->n_rays) all.push_back(GenSeg{i, (uint32_t)b, 0, descs[i]->n_rays});
    TRY(dsegs.alloc(sizeof(GenSeg) * std::max<size_t>(1, all.size())));
    TRY(dsegoff.alloc(4 * std::max<size_t>(1, all.size())));
    // the histogram rows of the finishing kernels: one per (instance, band,
    // channel, kind) for the prefix sums, one per (instance, band, channel)
    // for the combination step
    std::vector<uint4> rows, vrows;
    uint32_t max_nb = 1;
    for (uint32_t i = 0; i < n_items; ++i)
        for (uint32_t b = 0; b < 3; ++b)
            if ((descs[i]->band_mask >> b) & 1u)
                for (uint32_t ch = 0; ch < inst[i].P.n_channels; ++ch)
                    for (uint32_t kind = 0; kind < inst[i].P.hist_kinds; ++kind) rows.push_back(make_uint4(i, b, ch, kind));
    for (const uint4& r : rows)
        if (r.w == 0) {
            vrows.push_back(r);
            max_nb = std::max(max_nb, inst[r.x].P.n_bins);
        }
    TRY(drows.alloc(sizeof(uint4) * rows.size()));
    TRY(dvrows.alloc(sizeof(uint4) * vrows.size()));

    // the ray queues, their hit records and the shadow queue; a queue's
    // shading emits <= cap * recorders requests, and mesh sources add their
    // k = 0 connections of the rays generated in the same iteration
    const size_t sh_per_ray = std::max<size_t>(1, max_recs) * (any_mesh ? 2 : 1);
    if (noalloc) {
        // as many rays in flight as the rest of the region holds
        const size_t per_ray = 6 * 16 + 8 + sh_per_ray * 32;
        const size_t slack = 9 * kArenaAlign;
        const size_t fit = arena->left() > slack ? (arena->left() - slack) / per_ray : 0;
        if (fit < kNoallocMinRays) {
            c->last_error = "render needs more device memory than the no-allocation scratch region holds (" +
                            std::to_string(arena->size >> 20) + " MiB): histograms, stacks and tables take " +
                            std::to_string(arena->used >> 20) + " MiB";
            return EAR_ERR_CAPACITY;
        }
        cap_rays = (uint32_t)std::min<size_t>(cap_rays, fit / 64 * 64);
    }
    for (int k = 0; k < 2; ++k) {
        TRY(qo[k].alloc((size_t)cap_rays * 16));
        TRY(qd[k].alloc((size_t)cap_rays * 16));
        TRY(qm[k].alloc((size_t)cap_rays * 16));
    }
    TRY(hits.alloc((size_t)cap_rays * 8));
    const size_t cap_sh = (size_t)cap_rays * sh_per_ray;
    TRY(so.alloc(cap_sh * 16));
    TRY(sm.alloc(cap_sh * 16));
    TileOut tile_out{nullptr, nullptr, tile_stride};
    if (tiled) {
        const size_t n_rec = cap_sh * tile_stride;
        TRY(tkey.alloc(n_rec * 4));
        TRY(tpay.alloc(n_rec * 8));
        TRY(tsort.alloc(n_rec * 8));
        TRY(tcnt.alloc((size_t)K * 4));
        TRY(toff.alloc(((size_t)K + 1) * 4));
        TRY(tcur.alloc((size_t)K * 4));
        TRY(tkb.alloc(((size_t)n_items + 1) * 4));
        TRY(tfirst.alloc((n_rec + kTileChunk - 1) / kTileChunk * 4));
        TRY(clear(tcnt, 0));
        EAR_HIP_TRY(c, hipMemcpyAsync(tkb.p, kbase.data(), 4 * kbase.size(), hipMemcpyHostToDevice, st));
        tile_out.key = tkey.as<uint32_t>();
        tile_out.pay = tpay.as<uint2>();
    }

    uint32_t* cnt = counters.as<uint32_t>();  // [0],[1] ray queues, [2] shadow queue
    unsigned long long* stp = stats.as<unsigned long long>();
    const InstDev* di = dinst.as<InstDev>();
    int* gsp = gs.as<int>();

    // the queue counters read by the host: pinned memory, or in the
    // no-allocation render a host array (a synchronous copy)
    uint32_t host_cnt[4] = {0, 0, 0, 0};
    uint32_t* pinned = host_cnt;
    if (!noalloc) EAR_HIP_TRY(c, hipHostMalloc((void**)&pinned, 16, hipHostMallocDefault));
    struct PinGuard {
        uint32_t* p;
        ~PinGuard() {
            if (p) hipHostFree(p);
        }
    } pg{noalloc ? nullptr : pinned};

    std::vector<EventPair> ev_ext, ev_sha, ev_shw, ev_spl;
    auto mk = [&](std::vector<EventPair>& v) -> EventPair& {
        v.emplace_back();
        hipEventCreate(&v.back().a);
        hipEventCreate(&v.back().b);
        return v.back();
    };
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (!noalloc) {
        EAR_HIP_TRY(c, hipEventCreate(&t0));
        EAR_HIP_TRY(c, hipEventCreate(&t1));
        EAR_HIP_TRY(c, hipEventRecord(t0, st));
    }

    // The bounce loop is device driven: every iteration tops the current
    // queue up with new rays (k_plan + k_generate), extends and shades it
    // into the other queue, and connects the shadow requests.  Kernels read
    // their counts from device memory and run on fixed persistent grids, so
    // the host only checks for the end every few iterations, and an
    // iteration block can be captured once as a graph and replayed.
    uint32_t bounces = 0, launches = 0;
    ear_status status = EAR_OK;
    Queue qq[2] = {Queue{qo[0].as<float4>(), qd[0].as<float4>(), qm[0].as<uint4>()},
                   Queue{qo[1].as<float4>(), qd[1].as<float4>(), qm[1].as<uint4>()}};
    ShadowQ SQ{so.as<float4>(), sm.as<float4>()};
    {
        std::vector<uint32_t> off(all.size());
        uint64_t acc = 0;
        for (size_t k = 0; k < all.size(); ++k) {
            off[k] = (uint32_t)acc;
            acc += all[k].count;
        }
        if (!all.empty()) {
            EAR_HIP_TRY(c, hipMemcpyAsync(dsegs.p, all.data(), sizeof(GenSeg) * all.size(), hipMemcpyHostToDevice, st));
            EAR_HIP_TRY(c, hipMemcpyAsync(dsegoff.p, off.data(), 4 * off.size(), hipMemcpyHostToDevice, st));
        }
    }
    const uint32_t n_segs = (uint32_t)std::max<size_t>(1, all.size());
    const uint32_t total32 = (uint32_t)total_rays;
    const uint32_t g_gen = grid_for(
        cap_rays, rounds_cap(c, batch ? (const void*)k_generate<true> : (const void*)k_generate<false>, tu.waves_per_cu, kBlock),
        kBlock);
    const uint32_t g_shade = grid_for(
        cap_rays,
        rounds_cap(c, batch ? (const void*)k_shade<true> : (const void*)k_shade<false>, tu.waves_per_cu, kShadeBlock),
        kShadeBlock);
    const uint32_t g_ext = grid_for(cap_rays, cap_e, cfg_e.block);
    const uint32_t g_shw = grid_for((uint32_t)std::min<size_t>(cap_sh, 0xFFFFFFFFu), cap_s, cfg_s.block);
    // count and scatter: the same grid, two workgroups per CU (at least 16k
    // record slots each).  Fewer, longer-running workgroups give each key
    // longer runs of the sorted array per workgroup, so fewer cache lines are
    // written in part from several XCDs (on the hall: 36.0 -> 30.3 ms of
    // tile kernels per render against four per CU, 31.5 ms at one)
    const uint32_t g_tile = (uint32_t)std::max<size_t>(
        1, std::min<size_t>((size_t)c->num_cus * 2, cap_sh * tile_stride / (kTileBlock * kTileBatch * 4)));
    // splat: one workgroup per kTileChunk records of the most one launch can write
    const uint32_t g_splat = (uint32_t)std::max<size_t>(1, (cap_sh * tile_stride + kTileChunk - 1) / kTileChunk);
    // eager: the bounce loop launched from the host (always in the
    // no-allocation render); timed: with per-kernel events
    const bool eager = tu.graph_iters == 0 || noalloc;
    const bool timed = eager && !noalloc;
    const uint32_t refill = tu.refill ? tu.refill : 24u;
    const uint32_t fchunk = tu.fetch_chunk;  // 0: one entry per idle lane
    const uint32_t nseg = tu.fetch_segments ? std::min<uint32_t>(tu.fetch_segments, kMaxSeg) : 64u;
    const Fetch fq_e{segc.as<uint32_t>(), refill, fchunk, nseg},
        fq_s{segc.as<uint32_t>() + kMaxSeg * kSegStride, refill, fchunk, nseg};
    // one iteration; cur = the queue it consumes
    auto iteration = [&](int cur) {
        uint32_t* nq = cnt + cur;
        k_plan<<<1, 64, 0, st>>>(cnt + 3, total32, cap_rays, nq, cnt + 6);
        if (batch)
            k_generate<true><<<g_gen, kBlock, 0, st>>>(di, dsegs.as<GenSeg>(), dsegoff.as<uint32_t>(), n_segs, 0u,
                                                       cnt + 6, (uint32_t)max_recs, any_mesh, qq[cur], nq, SQ,
                                                       cnt + 2, stp);
        else
            k_generate<false><<<g_gen, kBlock, 0, st>>>(di, dsegs.as<GenSeg>(), dsegoff.as<uint32_t>(), n_segs, 0u,
                                                        cnt + 6, (uint32_t)max_recs, any_mesh, qq[cur], nq, SQ,
                                                        cnt + 2, stp);
        EventPair* e1 = timed ? &mk(ev_ext) : nullptr;
        if (e1) hipEventRecord(e1->a, st);
        hipLaunchKernelGGL(fe, dim3(g_ext), dim3(cfg_e.block), 0, st, di, gsp, gthreads, qq[cur], (const uint32_t*)nq,
                           hits.as<int2>(), stp, fq_e);
        if (e1) hipEventRecord(e1->b, st);
        EventPair* e2 = timed ? &mk(ev_sha) : nullptr;
        if (e2) hipEventRecord(e2->a, st);
        if (batch)
            k_shade<true><<<g_shade, kShadeBlock, 0, st>>>(di, qq[cur], nq, hits.as<int2>(), qq[1 - cur], cnt + (1 - cur),
                                                      SQ, cnt + 2, (uint32_t)max_recs, stp);
        else
            k_shade<false><<<g_shade, kShadeBlock, 0, st>>>(di, qq[cur], nq, hits.as<int2>(), qq[1 - cur],
                                                       cnt + (1 - cur), SQ, cnt + 2, (uint32_t)max_recs, stp);
        if (e2) hipEventRecord(e2->b, st);
        EventPair* e3 = timed ? &mk(ev_shw) : nullptr;
        if (e3) hipEventRecord(e3->a, st);
        hipLaunchKernelGGL(fs, dim3(g_shw), dim3(cfg_s.block), 0, st, di, gsp, gthreads, SQ, (const uint32_t*)(cnt + 2),
                           stp, fq_s, tile_out);
        if (e3) hipEventRecord(e3->b, st);
        if (tiled) {
            // the records of this launch, grouped by key and splatted per tile
            EventPair* e4 = timed ? &mk(ev_spl) : nullptr;
            if (e4) hipEventRecord(e4->a, st);
            k_tile_count<<<g_tile, kTileBlock, 0, st>>>(tkey.as<uint32_t>(), cnt + 2, tile_stride, tcnt.as<uint32_t>(),
                                                        K);
            k_tile_scan<<<1, 1024, 0, st>>>(tcnt.as<uint32_t>(), toff.as<uint32_t>(), tcur.as<uint32_t>(), K,
                                            tfirst.as<uint32_t>());
            k_tile_scatter<<<g_tile, kTileBlock, 0, st>>>(di, tkb.as<uint32_t>(), n_items, tkey.as<uint32_t>(),
                                                          tpay.as<uint2>(), cnt + 2, tile_stride, tcur.as<uint32_t>(),
                                                          K, tsort.as<uint2>());
            k_tile_splat<<<g_splat, kTileBlock, 0, st>>>(di, tkb.as<uint32_t>(), n_items, toff.as<uint32_t>(), K,
                                                         tfirst.as<uint32_t>(), tsort.as<uint2>());
            if (e4) hipEventRecord(e4->b, st);
            launches += 4;
        }
        // the consumed queue, the shadow queue and the fetch counters
        k_reset_iter<<<1, 64, 0, st>>>(cnt, (1u << cur) | 0x4u, segc.as<uint32_t>(), 2 * kMaxSeg);
        launches += 6;
        ++bounces;
    };
    // the render has ended once every ray was taken and no queue holds any
    auto finished = [&](int cur, bool* done) -> ear_status {
        EAR_HIP_TRY(c, hipMemcpyAsync(pinned, cnt, 16, hipMemcpyDeviceToHost, st));
        EAR_HIP_TRY(c, hipStreamSynchronize(st));
        *done = pinned[cur] == 0 && pinned[3] == total32;
        return EAR_OK;
    };
    // the tail: every ray emitted and few live ones left (queue cur, of
    // pinned[cur] rays, read by finished); one launch traces them to their ends
    // (tail_rays 0: 2^20 live rays, measured best on the 200k-triangle hall,
    // or every ray when no scene has more than 64 triangles: their walks are
    // so short that the queues between the wavefront kernels cost more than
    // the phase divergence of one launch (12-triangle box, 3M rays: 3.25e9
    // against 2.11e9 segments/s); 1: no tail launch)
    uint32_t max_tris = 0;
    for (uint32_t i = 0; i < n_items; ++i) max_tris = std::max(max_tris, inst[i].n_tris);
    const uint32_t tail_default = max_tris <= 64 ? 0xFFFFFFFFu : (1u << 20);
    const uint32_t tail_rays = tu.tail_rays == 0 ? tail_default : (tu.tail_rays == 1 ? 0u : tu.tail_rays);
    bool tail_ok = tail_rays > 0 && !instr && !any_log;
    for (uint32_t i = 0; i < n_items; ++i) tail_ok = tail_ok && inst[i].P.estimator == EAR_EST_NEE;
    const TailFn ft = tail_ok ? pick_tail(cfg_e) : nullptr;
    auto tail = [&](int cur, bool* done) -> ear_status {
        if (*done || !ft || pinned[3] != total32 || pinned[cur] > tail_rays) return EAR_OK;
        hipLaunchKernelGGL(ft, dim3(g_ext), dim3(cfg_e.block), 0, st, di, gsp, gthreads, qq[cur],
                           (const uint32_t*)(cnt + cur), stp);
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) return hip_status(c, le, "tail launch");
        launches += 1;
        *done = true;
        return EAR_OK;
    };
    const uint32_t max_iters = max_bounces + 2 + (uint32_t)((total_rays + cap_rays - 1) / cap_rays) * (max_bounces + 1);
    int cur = 0;
    bool done = false;
    if (eager) {
        const uint32_t every = 4;
        while (!done && status == EAR_OK) {
            for (uint32_t k = 0; k < every; ++k) {
                iteration(cur);
                cur = 1 - cur;
            }
            hipError_t le = hipGetLastError();
            if (le != hipSuccess) status = hip_status(c, le, "tracer launch");
            if (status == EAR_OK) status = finished(cur, &done);
            if (status == EAR_OK) status = tail(cur, &done);
            if (!done && bounces > max_iters) {
                c->last_error = "bounce loop did not drain";
                status = EAR_ERR_INTERNAL;
            }
        }
    } else {
        const uint32_t per = std::max<uint32_t>(2, (tu.graph_iters + 1) & ~1u);  // even: replays start on queue 0
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        EAR_HIP_TRY(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (uint32_t k = 0; k < per; ++k) iteration((int)(k & 1u));
        hipError_t ce = hipStreamEndCapture(st, &graph);
        if (ce == hipSuccess) ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ce != hipSuccess) status = hip_status(c, ce, "bounce-loop graph");
        const uint32_t captured = bounces;
        bounces = 0;
        launches = 0;
        while (!done && status == EAR_OK) {
            hipError_t le = hipGraphLaunch(exec, st);
            bounces += captured;
            launches += 1;
            if (le != hipSuccess) status = hip_status(c, le, "bounce-loop graph launch");
            if (status == EAR_OK) status = finished(0, &done);
            if (status == EAR_OK) status = tail(0, &done);  // an even replay ends on queue 0
            if (!done && bounces > max_iters + per) {
                c->last_error = "bounce loop did not drain";
                status = EAR_ERR_INTERNAL;
            }
        }
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
    }
    if (status != EAR_OK) return status;
    if (width == 4)
        k_direct<4><<<n_items, kBlock, 0, st>>>(di, gsp, gthreads, stp);
    else
        k_direct<2><<<n_items, kBlock, 0, st>>>(di, gsp, gthreads, stp);
    EAR_HIP_TRY(c, hipMemcpyAsync(drows.p, rows.data(), sizeof(uint4) * rows.size(), hipMemcpyHostToDevice, st));
    k_finalize<<<(uint32_t)rows.size(), kFinBlock, 0, st>>>(di, drows.as<uint4>());
    EAR_HIP_TRY(c, hipMemcpyAsync(dvrows.p, vrows.data(), sizeof(uint4) * vrows.size(), hipMemcpyHostToDevice, st));
    k_values<<<dim3((max_nb + 255) / 256, (uint32_t)vrows.size()), 256, 0, st>>>(di, dvrows.as<uint4>());
    EAR_HIP_TRY(c, hipGetLastError());
    if (t1) EAR_HIP_TRY(c, hipEventRecord(t1, st));
    EAR_HIP_TRY(c, hipStreamSynchronize(st));

    // copy out [band slot][channel][bin] rows of every rendered band
    for (uint32_t i = 0; i < n_items; ++i) {
        ear_hist* out = outs[i];
        if (!out) continue;
        const RenderParams& P = inst[i].P;
        const size_t nb = P.n_bins;
        const uint32_t nch = P.n_channels;
        for (int b = 0; b < 3; ++b) {
            if (!((descs[i]->band_mask >> b) & 1u)) continue;
            for (uint32_t ch = 0; ch < nch; ++ch) {
                for (uint32_t kind = 0; kind < P.hist_kinds; ++kind) {
                    size_t roff = (((size_t)b * nch + ch) * P.hist_kinds + kind) * nb;
                    size_t ooff = ((size_t)b * nch + ch) * nb;
                    int64_t* raw = kind == 0 ? out->raw : out->abs_raw;
                    double* val = kind == 0 ? out->value : out->abs_value;
                    if (raw)
                        EAR_HIP_TRY(c, hipMemcpyAsync(raw + 2 * ooff,
                                                      hist.as<unsigned long long>() + 2 * (hist_off[i] + 2 * roff),
                                                      nb * 16, hipMemcpyDeviceToHost, st));
                    if (val && inst[i].value)
                        EAR_HIP_TRY(c, hipMemcpyAsync(val + ooff, value.as<double>() + value_off[i] + roff, nb * 8,
                                                      hipMemcpyDeviceToHost, st));
                }
            }
        }
        if (P.trace_rays) {
            std::vector<int32_t> lh((size_t)3 * P.trace_rays * P.trace_segs);
            std::vector<uint32_t> ln((size_t)3 * P.trace_rays);
            EAR_HIP_TRY(c, hipMemcpyAsync(lh.data(), logh.as<int32_t>() + logh_off[i], lh.size() * 4,
                                          hipMemcpyDeviceToHost, st));
            EAR_HIP_TRY(c, hipMemcpyAsync(ln.data(), logn.as<uint32_t>() + logn_off[i], ln.size() * 4,
                                          hipMemcpyDeviceToHost, st));
            EAR_HIP_TRY(c, hipStreamSynchronize(st));
            for (int b = 0; b < 3; ++b) {
                if (!((descs[i]->band_mask >> b) & 1u)) continue;
                size_t per = (size_t)P.trace_rays * P.trace_segs;
                if (out->trace_hits) {
                    for (size_t j = 0; j < per; ++j) {
                        int32_t v = lh[(size_t)b * per + j];
                        out->trace_hits[(size_t)b * per + j] = v == (int32_t)0xFEFEFEFE ? -2 : v;
                    }
                }
                if (out->trace_nseg)
                    std::memcpy(out->trace_nseg + (size_t)b * P.trace_rays, ln.data() + (size_t)b * P.trace_rays,
                                (size_t)P.trace_rays * 4);
            }
        }
    }
    EAR_HIP_TRY(c, hipStreamSynchronize(st));

    if (st_out) {
        std::vector<unsigned long long> hc((size_t)ST_N * kStatCopies);
        EAR_HIP_TRY(c, hipMemcpy(hc.data(), stats.p, hc.size() * 8, hipMemcpyDeviceToHost));
        unsigned long long hs[ST_N] = {};
        for (uint32_t k = 0; k < kStatCopies; ++k)
            for (int j = 0; j < ST_N; ++j)
                hs[j] = j == ST_BINS ? std::max(hs[j], hc[(size_t)k * ST_N + j]) : hs[j] + hc[(size_t)k * ST_N + j];
        std::memset(st_out, 0, sizeof(*st_out));
        st_out->segments = hs[ST_SEGMENTS];
        st_out->shadow_rays = hs[ST_SHADOW];
        st_out->contributions = hs[ST_CONTRIB];
        st_out->rays = hs[ST_RAYS];
        st_out->node_visits_closest = hs[ST_NV_CL];
        st_out->tri_tests_closest = hs[ST_NT_CL];
        st_out->node_visits_shadow = hs[ST_NV_SH];
        st_out->tri_tests_shadow = hs[ST_NT_SH];
        st_out->bounces_launched = bounces;
        st_out->launches = launches;
        float ms = 0;
        if (t0 && t1) hipEventElapsedTime(&ms, t0, t1);
        st_out->trace_ms = ms;
        auto sum = [](std::vector<EventPair>& v) {
            double a = 0;
            for (auto& e : v) {
                float m = 0;
                hipEventElapsedTime(&m, e.a, e.b);
                a += m;
            }
            return a;
        };
        st_out->extend_ms = sum(ev_ext);
        st_out->shade_ms = sum(ev_sha);
        st_out->shadow_ms = sum(ev_shw);
        st_out->splat_ms = sum(ev_spl);
        st_out->shadow_contributions = hs[ST_SH_CONTRIB];
        st_out->shadow_ramps = hs[ST_SH_RAMPS];
        st_out->extend_launches = (uint32_t)ev_ext.size();
        st_out->shadow_launches = (uint32_t)ev_shw.size();
        st_out->extend_queries = hs[ST_SEGMENTS] - hs[ST_TAIL_SEG];  // the timed k_extend launches' share
        st_out->shadow_queries = hs[ST_SHADOW] - hs[ST_TAIL_SH];
        st_out->dropped_ramps = hs[ST_DROPPED];
        st_out->bins_needed = (uint32_t)std::min<unsigned long long>(hs[ST_BINS], 0xFFFFFFFFull);
    }
    for (auto* v : {&ev_ext, &ev_sha, &ev_shw, &ev_spl})
        for (auto& e : *v) {
            hipEventDestroy(e.a);
            hipEventDestroy(e.b);
        }
    if (t0) hipEventDestroy(t0);
    if (t1) hipEventDestroy(t1);
    return EAR_OK;
}

ear_status run_render(Ctx* c, DeviceScene* s, const ear_render_desc* d, ear_hist* out, ear_stats* st) {
    DeviceScene* sc[1] = {s};
    const ear_render_desc* dd[1] = {d};
    ear_hist* oo[1] = {out};
    return run_batch(c, 1, sc, dd, oo, st);
}

}  // namespace ear

// ===========================================================================
// Adjoint: gradient of a loss on the histograms with respect to the rendered
// bands' per-material reflection coefficient and specularity.
// ===========================================================================
namespace ear {
namespace {

constexpr int kAdMaxMat = 8;
// The gradient is accumulated into kGradCopies copies, a block adding into
// copy blockIdx % kGradCopies, and the copies are summed at the end: the
// per-wave partials of every launch would otherwise all land on the same
// 2 n_mat addresses, and same-address atomics serialise in the L2.
constexpr uint32_t kGradCopies = 128;

// Tangent state of the rays of a queue, SoA [q][field][slot] with fields
// {o.xyz, d.xyz, dL, dlogI} per specularity parameter q, then
// dlogI_refl[q][slot].
struct Tangents {
    float* t;  // [n_mat][8][cap]
    float* r;  // [n_mat][cap]
    uint32_t cap;
    __device__ float& f(uint32_t q, int field, uint32_t slot) const {
        return t[((size_t)q * 8 + field) * cap + slot];
    }
    __device__ float& refl(uint32_t q, uint32_t slot) const { return r[(size_t)q * cap + slot]; }
    __device__ float& row(uint32_t j, uint32_t slot) const { return t[(size_t)j * cap + slot]; }  // score form
};

// The adjoint histograms as prefix sums: row (band, channel) holds bins + 1
// pairs {sum_{b<j} a[b], sum_{b<j} b a[b]}, so that a ramp's weighted sum
// sum_b (w + s - b) a[b] over its bins is four loads at any width.
struct AdjArgs {
    const double2* adj_s;  // [3][n_ch][bins + 1] prefix sums of dL / d raw (signed), or null
    const double2* adj_a;  // [3][n_ch][bins + 1] prefix sums of dL / d raw (|.|), or null
    double* grad;         // [3][2][n_mat]: band, {refl, spec}, material (copy 0)
    uint32_t n_mat;
    uint32_t copy_stride;  // doubles from one gradient copy to the next
    __device__ double* grad_copy() const { return grad + (size_t)(blockIdx.x % kGradCopies) * copy_stride; }
};

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    return v;
}

// The end-of-launch gradient partials of a block (single instance, one
// band): per-wave sums, then summed over the block's waves in LDS, one
// atomic per value into the block's gradient copy.
__device__ void block_grad_flush(const AdjArgs& A, int band_acc, const double* g_refl, const double* g_spec,
                                 double* red /* LDS [kBlock / 64][2 * kAdMaxMat] */) {
    int band = __shfl(band_acc, 0);
    for (int l = 1; l < 64 && band < 0; ++l) band = __shfl(band_acc, l);
    const uint32_t nm = A.n_mat;
    const uint32_t wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < kAdMaxMat; ++q) {
        const double sr = wave_sum_d((uint32_t)q < nm ? g_refl[q] : 0.0);
        const double ss = wave_sum_d((uint32_t)q < nm ? g_spec[q] : 0.0);
        if (lane_id() == 0) {
            red[wave * 2 * kAdMaxMat + q] = sr;
            red[wave * 2 * kAdMaxMat + kAdMaxMat + q] = ss;
        }
    }
    // the band of the block: any wave's (one band per launch)
    __shared__ int bband;
    if (threadIdx.x == 0) bband = -1;
    __syncthreads();
    if (lane_id() == 0 && band >= 0) bband = band;
    __syncthreads();
    const int b = bband;
    if (threadIdx.x < 2 * nm && b >= 0) {
        const uint32_t q = threadIdx.x % nm, which = threadIdx.x / nm;  // 0 refl, 1 spec
        double v = 0.0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) v += red[w * 2 * kAdMaxMat + which * kAdMaxMat + q];
        if (v != 0.0) atomicAdd(A.grad_copy() + ((size_t)b * 2 + which) * nm + q, v);
    }
}

__global__ void k_sum_grad_copies(double* g, uint32_t n, uint32_t copy_stride) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double v = 0.0;
        for (uint32_t c = 0; c < kGradCopies; ++c) v += g[(size_t)c * copy_stride + i];
        g[i] = v;
    }
}

// sum over the ramp's bins b = s .. s + w - 1 inside [0, nb) of
// a[b] (w - (b - s)) / w * per_c, from the row's prefix sums
__device__ double ramp_adjoint(const double2* pre, uint32_t nb, int s, int w, double per_c) {
    const int lo = s < 0 ? 0 : s;
    const int hi = s + w > (int)nb ? (int)nb : s + w;
    if (lo >= hi) return 0.0;
    const double2 a = pre[lo], b = pre[hi];
    return ((double)(w + s) * (b.x - a.x) - (b.y - a.y)) / (double)w * per_c;
}

// dL/dc of a contribution c at time t, distance dist, direction dir.
__device__ double contribution_adjoint(const RenderParams& P, const AdjArgs& A, int band, const RecParams& R,
                                       V3 dir, float c, float t, float dist) {
    const uint32_t nb = P.n_bins, nch = P.n_channels;
    float width = sqrtf(dist);
    int w = (int)ceilf(width);
    double per_c = 2.0 / (double)width;
    double sgn = c < 0.0f ? -1.0 : 1.0;
    double a = 0.0;
    auto row = [&](const double2* base, int ch) { return base + ((size_t)band * nch + ch) * (nb + 1); };
    if (R.kind == EAR_REC_MONO) {
        int s = (int)((double)t * 44100.0);
        if (A.adj_s) a += ramp_adjoint(row(A.adj_s, R.channel), nb, s, w, per_c);
        if (A.adj_a) a += sgn * ramp_adjoint(row(A.adj_a, R.channel), nb, s, w, per_c);
    } else {
        float dt = dot(dir, V3{R.rx, R.ry, R.rz});
        float td = R.head_size / 343.0f;
        int s_right = (int)((double)(t - dt * td) * 44100.0);
        int s_left = (int)((double)(t + dt * td) * 44100.0);
        float factor = powf_ear_l2(R.head_abs[band], R.l2_head_abs[band], fabsf(dt) * R.head_size);
        double gl = 1.0, gr = 1.0;
        if (dt < 0.0f)
            gr = (double)(factor * factor);
        else
            gl = (double)(factor * factor);
        if (A.adj_s) {
            a += gl * ramp_adjoint(row(A.adj_s, R.channel), nb, s_left, w, per_c);
            a += gr * ramp_adjoint(row(A.adj_s, R.channel + 1), nb, s_right, w, per_c);
        }
        if (A.adj_a) {
            a += sgn * gl * ramp_adjoint(row(A.adj_a, R.channel), nb, s_left, w, per_c);
            a += sgn * gr * ramp_adjoint(row(A.adj_a, R.channel + 1), nb, s_right, w, per_c);
        }
    }
    return a;
}

// The primal of k_shade (same draws, same arithmetic) plus forward-mode
// tangents of every specularity parameter through the reparameterised
// direction v = normalize(h (1 - s) + r s) with h and the hit triangles held
// fixed, and of log I with respect to every reflection coefficient.  Each
// visible connection adds (dL/dc)(dc/dtheta) to the gradient.  Arrival-time
// shifts and visibility changes are not differentiated.
// BATCH: the queue holds the rays of several instances (rm.w), each with
// its own adjoint and gradient (adj[id]); the per-lane gradient partials are
// then reduced per instance at the end of every grid-stride step.
template <int WIDTH, bool BATCH>
__global__ void __launch_bounds__(kBlock) k_shade_ad(const InstDev* __restrict__ inst, int* gstack, uint32_t gthreads,
                                                      Queue qin, const uint32_t* __restrict__ n_in,
                                                      const int2* __restrict__ hits, Tangents tin, Queue qout,
                                                      uint32_t* __restrict__ n_out, Tangents tout, AdjArgs A0,
                                                      const AdjArgs* __restrict__ adj,
                                                      unsigned long long* __restrict__ stats) {
    __shared__ int lds[kLdsMin * kBlock];
    uint32_t n = *n_in;
    uint32_t n_done = 0;
    uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    double g_refl[kAdMaxMat], g_spec[kAdMaxMat];
    for (int q = 0; q < kAdMaxMat; ++q) g_refl[q] = g_spec[q] = 0.0;
    int band_acc = -1;
    uint32_t nvis = 0, ntri = 0;
    uint32_t my_id = 0;
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        uint32_t i = base + threadIdx.x;
        bool cont = false;
        float4 no{}, nd{};
        uint4 nmeta{};
        float dtan[kAdMaxMat][8];
        float drefl[kAdMaxMat];
        uint32_t nm = A0.n_mat;
        if (i < n) {
            float4 ro = qin.o[i];
            float4 rd = qin.d[i];
            uint4 rm = qin.m[i];
            const uint32_t id = BATCH ? rm.w : 0u;
            my_id = id;
            const InstDev& In = inst[id];
            const RenderParams& P = In.P;
            const AdjArgs& A = BATCH ? adj[id] : A0;
            nm = A.n_mat;
            int band = (int)(rm.z & 0xFFu);
            band_acc = band;
            uint32_t k = ((rm.z >> 8) & 0xFFFFu) + 1;
            int2 h = hits[i];
            if (h.y >= 0) {
                float t = __int_as_float(h.x);
                V3 o{ro.x, ro.y, ro.z};
                V3 d{rd.x, rd.y, rd.z};
                V3 p{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
                float4 nmv = In.nm[h.y];
                V3 N{nmv.x, nmv.y, nmv.z};
                V3 nrm = dot(N, d) > 0.0f ? V3{N.x * -1.0f, N.y * -1.0f, N.z * -1.0f} : N;
                uint32_t mat = __float_as_uint(nmv.w);
                float4 mb = In.mat_band[3 * mat + band];
                Rng rng;
                rng.init(P.seed, rm.x, P.ctx_word[band], P.instance, rm.y);
                int bt = kReflect;
                if (mb.x >= 0.0f) bt = (rng.next() <= mb.x) ? kReflect : kRefract;
                float spec = mb.y;
                if (bt == kRefract) nrm = neg(nrm);
                V3 dist3 = sub(p, o);
                float l = length(dist3);
                float I = ro.w;
                I *= powf_ear_l2(P.air[band], P.l2_air[band], l);
                float L = rd.w + l;
                I *= mb.z;
                V3 prev = normalized(d);
                const double ln_air = log((double)P.air[band]);
                // tangents at the hit point (plane of the hit triangle)
                const double Nd = (double)N.x * d.x + (double)N.y * d.y + (double)N.z * d.z;
                const double dn = (double)length(d);
                double dprev[kAdMaxMat][3];
                for (uint32_t q = 0; q < nm; ++q) {
                    double to[3] = {tin.f(q, 0, i), tin.f(q, 1, i), tin.f(q, 2, i)};
                    double td[3] = {tin.f(q, 3, i), tin.f(q, 4, i), tin.f(q, 5, i)};
                    double dL = tin.f(q, 6, i), dlogI = tin.f(q, 7, i);
                    double Ndo = N.x * to[0] + N.y * to[1] + N.z * to[2];
                    double Ndd = N.x * td[0] + N.y * td[1] + N.z * td[2];
                    double dtt = -(Ndo + (double)t * Ndd) / Nd;
                    double dp[3] = {to[0] + td[0] * t + (double)d.x * dtt, to[1] + td[1] * t + (double)d.y * dtt,
                                    to[2] + td[2] * t + (double)d.z * dtt};
                    double dl = ((double)dist3.x * (dp[0] - to[0]) + (double)dist3.y * (dp[1] - to[1]) +
                                 (double)dist3.z * (dp[2] - to[2])) /
                                (double)l;
                    dL += dl;
                    dlogI += ln_air * dl;
                    double pd = ((double)d.x * td[0] + (double)d.y * td[1] + (double)d.z * td[2]) / dn;
                    dprev[q][0] = (td[0] - (double)d.x / dn * pd) / dn;
                    dprev[q][1] = (td[1] - (double)d.y / dn * pd) / dn;
                    dprev[q][2] = (td[2] - (double)d.z / dn * pd) / dn;
                    dtan[q][0] = (float)dp[0];
                    dtan[q][1] = (float)dp[1];
                    dtan[q][2] = (float)dp[2];
                    dtan[q][3] = (float)td[0];  // incoming direction tangent, replaced below
                    dtan[q][4] = (float)td[1];
                    dtan[q][5] = (float)td[2];
                    dtan[q][6] = (float)dL;
                    dtan[q][7] = (float)dlogI;
                    drefl[q] = tin.refl(q, i) + ((uint32_t)q == mat ? 1.0f / mb.z : 0.0f);
                }
                if (is_normal(I)) {
                    for (uint32_t r = 0; r < P.n_rec; ++r) {
                        const RecParams& R = P.rec[r];
                        V3 dv = sub(V3{R.px, R.py, R.pz}, p);
                        V3 lsdir = normalized(dv);
                        float dtd = dot(lsdir, nrm);
                        if (!(dtd > 0.0f)) continue;
                        V3 rv{0, 0, 0};
                        float diff, sf;
                        if (bt == kReflect) {
                            rv = reflect(prev, nrm);
                            diff = -dot(nrm, prev);
                            sf = fmaxf(0.0f, dot(rv, lsdir));
                        } else {
                            diff = dot(nrm, prev);
                            sf = fmaxf(0.0f, dot(prev, lsdir));
                        }
                        float pw = powf_ear(sf, kExp);
                        float factor = spec * kExpInt * pw + (1.0f - spec) * diff;
                        float c = I;
                        c *= factor;
                        float ls = length(dv);
                        float geo_air = powf_ear_l2(P.air[band], P.l2_air[band], ls);
                        float geo_inv = 2.0f / (4.0f * kPi * ls * ls);
                        c *= geo_air;
                        c *= geo_inv;
                        if (!is_normal(c)) continue;
                        if (k & 1u) c *= -1.0f;
                        float bt_ = 0.0f;
                        int bs_ = -1;
                        uint32_t bo_ = 0;
                        Traversal T = make_trav(In, gstack, gthreads);
                        if (traverse<kLdsMin, kBlock, WIDTH, true, false>(T, p, dv, bt_, bs_, bo_, lds + threadIdx.x,
                                                                          gtid, nvis, ntri))
                            continue;
                        double a_c = contribution_adjoint(P, A, band, R, lsdir, c, (L + ls) / 343.0f, L + ls);
                        if (a_c == 0.0) continue;
                        const double cnof = ((k & 1u) ? -1.0 : 1.0) * (double)I * (double)geo_air * (double)geo_inv;
                        const double lobe = (double)kExpInt * (double)pw;
                        for (uint32_t q = 0; q < nm; ++q) {
                            double dvv[3] = {-(double)dtan[q][0], -(double)dtan[q][1], -(double)dtan[q][2]};
                            double dls = (dvv[0] * dv.x + dvv[1] * dv.y + dvv[2] * dv.z) / (double)ls;
                            double ld = lsdir.x * dvv[0] + lsdir.y * dvv[1] + lsdir.z * dvv[2];
                            double dld[3] = {(dvv[0] - lsdir.x * ld) / ls, (dvv[1] - lsdir.y * ld) / ls,
                                             (dvv[2] - lsdir.z * ld) / ls};
                            double pn = dprev[q][0] * nrm.x + dprev[q][1] * nrm.y + dprev[q][2] * nrm.z;
                            double ddiff, dsf = 0.0;
                            if (bt == kReflect) {
                                ddiff = -pn;
                                if (sf > 0.0f) {
                                    double drv[3] = {dprev[q][0] - 2.0 * pn * nrm.x, dprev[q][1] - 2.0 * pn * nrm.y,
                                                     dprev[q][2] - 2.0 * pn * nrm.z};
                                    dsf = drv[0] * lsdir.x + drv[1] * lsdir.y + drv[2] * lsdir.z + dld[0] * rv.x +
                                          dld[1] * rv.y + dld[2] * rv.z;
                                }
                            } else {
                                ddiff = pn;
                                if (sf > 0.0f)
                                    dsf = dprev[q][0] * lsdir.x + dprev[q][1] * lsdir.y + dprev[q][2] * lsdir.z +
                                          dld[0] * prev.x + dld[1] * prev.y + dld[2] * prev.z;
                            }
                            double dlobe = sf > 0.0f ? (double)kExpInt * 1000.0 * pow((double)sf, 999.0) * dsf : 0.0;
                            double dfactor = (double)spec * dlobe + (1.0 - (double)spec) * ddiff;
                            if (q == mat) dfactor += lobe - (double)diff;
                            double dgeo = ln_air * dls - 2.0 * dls / (double)ls;
                            double dc = cnof * dfactor + (double)c * ((double)dtan[q][7] + dgeo);
                            g_spec[q] += a_c * dc;
                            g_refl[q] += a_c * (double)c * (double)drefl[q];
                        }
                    }
                    cont = !((double)I < P.cutoff) && k < P.max_bounces;
                }
                if (cont) {
                    // hemisphere sample and blend exactly as sample_hemi_spec
                    V3 hs = sample_hemi(rng, nrm);
                    V3 rdir = bt == kRefract ? d : reflect(d, nrm);
                    float gg = 1.0f - spec;
                    V3 w{hs.x * gg + rdir.x * spec, hs.y * gg + rdir.y * spec, hs.z * gg + rdir.z * spec};
                    float wl = length(w);
                    V3 v = normalized(w);
                    const bool local = (P.flags & EAR_FLAG_ADJ_LOCAL) != 0;
                    for (uint32_t q = 0; q < nm; ++q) {
                        if (local) {
                            for (int f = 0; f < 8; ++f) dtan[q][f] = 0.0f;
                            continue;
                        }
                        double td[3] = {dtan[q][3], dtan[q][4], dtan[q][5]};
                        double dr[3];
                        if (bt == kRefract) {
                            dr[0] = td[0];
                            dr[1] = td[1];
                            dr[2] = td[2];
                        } else {
                            double tn = td[0] * nrm.x + td[1] * nrm.y + td[2] * nrm.z;
                            dr[0] = td[0] - 2.0 * tn * nrm.x;
                            dr[1] = td[1] - 2.0 * tn * nrm.y;
                            dr[2] = td[2] - 2.0 * tn * nrm.z;
                        }
                        double dw[3] = {dr[0] * spec, dr[1] * spec, dr[2] * spec};
                        if (q == mat) {
                            dw[0] += (double)rdir.x - (double)hs.x;
                            dw[1] += (double)rdir.y - (double)hs.y;
                            dw[2] += (double)rdir.z - (double)hs.z;
                        }
                        double vdw = v.x * dw[0] + v.y * dw[1] + v.z * dw[2];
                        dtan[q][3] = (float)((dw[0] - v.x * vdw) / wl);
                        dtan[q][4] = (float)((dw[1] - v.y * vdw) / wl);
                        dtan[q][5] = (float)((dw[2] - v.z * vdw) / wl);
                    }
                    no = make_float4(p.x, p.y, p.z, I);
                    nd = make_float4(v.x, v.y, v.z, L);
                    nmeta = make_uint4(rm.x, rng.draw, (uint32_t)band | (k << 8), id);
                }
            }
        }
        if (BATCH) {
            // reduce the step's partials per instance present in the wave
            const uint32_t nmx = adj[my_id].n_mat;
            uint64_t todo = __ballot(i < n);
            while (todo) {
                const int leader = __ffsll((unsigned long long)todo) - 1;
                const uint32_t key = __shfl(my_id, leader);
                const int kband = __shfl(band_acc, leader);
                const bool mine = i < n && my_id == key;
                const uint32_t nmk = __shfl(nmx, leader);
                for (uint32_t q = 0; q < nmk; ++q) {
                    const double sr = wave_sum_d(mine ? g_refl[q] : 0.0);
                    const double ss = wave_sum_d(mine ? g_spec[q] : 0.0);
                    if (lane_id() == 0) {
                        double* gk = adj[key].grad_copy();
                        if (sr != 0.0) atomicAdd(gk + ((size_t)kband * 2 + 0) * nmk + q, sr);
                        if (ss != 0.0) atomicAdd(gk + ((size_t)kband * 2 + 1) * nmk + q, ss);
                    }
                }
                todo &= ~__ballot(mine);
            }
            for (int q = 0; q < kAdMaxMat; ++q) g_refl[q] = g_spec[q] = 0.0;
        }
        uint32_t slot = wave_append(n_out, cont);
        if (cont) {
            qout.o[slot] = no;
            qout.d[slot] = nd;
            qout.m[slot] = nmeta;
            for (uint32_t q = 0; q < nm; ++q) {
                for (int f = 0; f < 8; ++f) tout.f(q, f, slot) = dtan[q][f];
                tout.refl(q, slot) = drefl[q];
            }
        }
        n_done += i < n ? 1u : 0u;
    }
    stat_add(stats, ST_SEGMENTS, n_done);
    if (BATCH) return;
    __shared__ double red[kBlock / 64][2 * kAdMaxMat];
    block_grad_flush(A0, band_acc, g_refl, g_spec, &red[0][0]);
}

// Connection requests of the local adjoint: the hit point, the gradient
// weights the connection adds when it is visible, and which recorder,
// band, material and instance it belongs to.
struct AdjShadowQ {
    float4* p;  // {px, py, pz, -}
    uint4* m;   // {rec | band << 8 | mat << 16, instance, -, -}
    float4* w;  // [4 per request]: d/drefl_q (q = 0..7), then d/dspec_q (q = 0..7), each times dL/dc
};

// The score form (EAR_FLAG_ADJ_SCORE): a queue entry whose metadata carries
// kContFlag is a continuation path.  Its tangent rows hold S (the weight of
// its spawning material, zero elsewhere) and, in row n_mat, the vertex that
// spawned it (its stream's key).
constexpr uint32_t kContFlag = 1u << 31;
// a vertex spawns a continuation with probability min(1, kScoreRate F), F
// its rim flux (the rim term's weight: the least variance for a given
// number of continuations)
constexpr float kScoreRate = 0.05f;

struct ContOut {  // where a score-form vertex appends its continuation
    Queue q;
    uint32_t* n;
    Tangents t;
};

// The adjoint with EAR_FLAG_ADJ_LOCAL (or EAR_FLAG_ADJ_SCORE) for every ray:
// no direction or position tangents travel between vertices, so a ray
// carries only dlogI/drefl (n_mat floats) and, in the score form, S (n_mat
// floats).  At each connection of a hit on material m,
// d c / d spec_m = sign I air^l 2 / (4 pi l^2) (lobe - diffuse) and
// d c / d refl_q = c dlogI/drefl_q; the score form adds c S_q to d c / dspec_q.
// Times dL/dc they are queued as a request whose visibility k_adj_vis
// decides, as k_shade queues the forward's.  Same primal and draws as
// k_shade; arrays are indexed at compile time only.
//
// SCORE: after its connections, a vertex whose path goes on adds to S_m
// (m its material) d log p(v)/ds - F for the direction v it drew, and with
// probability q = min(1, kScoreRate F) appends a continuation leaving the
// vertex along a rim direction (score_rim_dir) with weight -F / q on spec_m,
// adding F / q to S_m.  A continuation replays transport only:
// no refl tangents, no explicit lobe term, no score, no spawns.
// CONT: every entry of the queue is a continuation (the continuations' own
// queues), so the main paths' steps compile away.
template <bool BATCH, bool SCORE, bool CONT = false>
__global__ void __launch_bounds__(kBlock) k_shade_adl(const InstDev* __restrict__ inst, Queue qin,
                                                       const uint32_t* __restrict__ n_in,
                                                       const int2* __restrict__ hits, Tangents tin, Queue qout,
                                                       uint32_t* __restrict__ n_out, Tangents tout, AdjShadowQ aq,
                                                       uint32_t* __restrict__ n_aq, uint32_t max_rec, AdjArgs A0,
                                                       const AdjArgs* __restrict__ adj, ContOut co,
                                                       unsigned long long* __restrict__ stats) {
    __shared__ uint32_t sc[6][kBlock / 64 + 1];  // scratch of the block-wide appends, alternated
    __shared__ HemiLds hl[kBlock / 64];
    const uint32_t n = *n_in;
    uint32_t n_done = 0;
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t parity = (base / (gridDim.x * kBlock)) & 1u;
        bool cont = false, alive = false, contn = false, spawn = false;
        Rng rng;
        float4 no{}, nd{};
        uint4 nmeta{};
        float drefl[kAdMaxMat], sacc[kAdMaxMat];
#pragma unroll
        for (int q = 0; q < kAdMaxMat; ++q) drefl[q] = sacc[q] = 0.0f;
        uint32_t nm = A0.n_mat, id = 0, k = 0, mat = 0, n_rec = 0, kspawn = 0, sdraw = 0, ray = 0;
        int band = 0, bt = kReflect;
        float I = 0.0f, L = 0.0f, spec = 0.0f, wcont = 0.0f;
        V3 p{0, 0, 0}, nrm{0, 0, 0}, prev{0, 0, 0}, vdir{0, 0, 0}, rdir{0, 0, 0}, vb{0, 0, 0};
        if (i < n) {
            const float4 ro = qin.o[i];
            const float4 rd = qin.d[i];
            const uint4 rm = qin.m[i];
            id = BATCH ? rm.w : 0u;
            ray = rm.x;
            const InstDev& In = inst[id];
            const RenderParams& P = In.P;
            nm = (BATCH ? adj[id] : A0).n_mat;
            n_rec = P.n_rec;
            band = (int)(rm.z & 0xFFu);
            k = ((rm.z >> 8) & 0xFFFFu) + 1;
            contn = CONT || (SCORE && (rm.z & kContFlag) != 0u);
            const int2 h = hits[i];
            if (h.y >= 0) {
                const float t = __int_as_float(h.x);
                const V3 o{ro.x, ro.y, ro.z};
                const V3 d{rd.x, rd.y, rd.z};
                p = V3{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
                const float4 nmv = In.nm[h.y];
                const V3 N{nmv.x, nmv.y, nmv.z};
                nrm = dot(N, d) > 0.0f ? V3{N.x * -1.0f, N.y * -1.0f, N.z * -1.0f} : N;
                mat = __float_as_uint(nmv.w);
                const float4 mb = In.mat_band[3 * mat + band];
                if (contn) {
                    kspawn = __float_as_uint(tin.row(nm, i));
                    score_rng_init(rng, P.seed, rm.x, P.ctx_word[band], P.instance, kspawn, rm.y);
                } else {
                    rng.init(P.seed, rm.x, P.ctx_word[band], P.instance, rm.y);
                }
                if (mb.x >= 0.0f) bt = (rng.next() <= mb.x) ? kReflect : kRefract;
                spec = mb.y;
                if (bt == kRefract) nrm = neg(nrm);
                const float l = length(sub(p, o));
                I = ro.w;
                I *= powf_ear_l2(P.air[band], P.l2_air[band], l);
                L = rd.w + l;
                I *= mb.z;
                prev = normalized(d);
#pragma unroll
                for (int q = 0; q < kAdMaxMat; ++q) {
                    if ((uint32_t)q < nm) {
                        drefl[q] = contn ? 0.0f : tin.refl(q, i) + ((uint32_t)q == mat ? 1.0f / mb.z : 0.0f);
                        if (SCORE) sacc[q] = tin.row(q, i);
                    }
                }
                alive = is_normal(I);
                if (alive) {
                    cont = !((double)I < P.cutoff) && k < P.max_bounces;
                    if (cont) rdir = bt == kRefract ? d : reflect(d, nrm);
                }
            }
        }
        {
            const V3 hv = sample_hemi_wave(cont, true, rng, nrm, hl[threadIdx.x >> 6]);
            if (cont) {
                vdir = blend_spec(hv, rdir, spec);
                no = make_float4(p.x, p.y, p.z, I);
                nd = make_float4(vdir.x, vdir.y, vdir.z, L);
                nmeta = make_uint4(ray, rng.draw, (uint32_t)band | (k << 8) | (contn ? kContFlag : 0u), id);
            }
        }
        // connections to every recorder (src/Scene.cpp:185-268), queued with
        // their gradient weights
        for (uint32_t r = 0; r < max_rec; ++r) {
            bool emit = false;
            float ds = 0.0f, ac = 0.0f;
            if (alive && r < n_rec) {
                const RenderParams& P = inst[id].P;
                const AdjArgs& A = BATCH ? adj[id] : A0;
                const RecParams& R = P.rec[r];
                const V3 dv = sub(V3{R.px, R.py, R.pz}, p);
                const V3 lsdir = normalized(dv);
                if (dot(lsdir, nrm) > 0.0f) {
                    float diff, sf;
                    if (bt == kReflect) {
                        const V3 rv = reflect(prev, nrm);
                        diff = -dot(nrm, prev);
                        sf = fmaxf(0.0f, dot(rv, lsdir));
                    } else {
                        diff = dot(nrm, prev);
                        sf = fmaxf(0.0f, dot(prev, lsdir));
                    }
                    const float pw = powf_ear(sf, kExp);
                    const float factor = spec * kExpInt * pw + (1.0f - spec) * diff;
                    float c = I;
                    c *= factor;
                    const float ls = length(dv);
                    const float geo_air = powf_ear_l2(P.air[band], P.l2_air[band], ls);
                    const float geo_inv = 2.0f / (4.0f * kPi * ls * ls);
                    c *= geo_air;
                    c *= geo_inv;
                    if (is_normal(c)) {
                        if (k & 1u) c *= -1.0f;
                        const double a_c =
                            contribution_adjoint(P, A, band, R, lsdir, c, (L + ls) / 343.0f, L + ls);
                        if (a_c != 0.0) {
                            const double cnof =
                                ((k & 1u) ? -1.0 : 1.0) * (double)I * (double)geo_air * (double)geo_inv;
                            ds = contn ? 0.0f : (float)(a_c * cnof * ((double)kExpInt * (double)pw - (double)diff));
                            ac = (float)(a_c * (double)c);
                            emit = true;
                        }
                    }
                }
            }
            const uint32_t slot = block_append<kBlock>(n_aq, emit, sc[r & 1u]);
            if (emit) {
                float ws[kAdMaxMat];
#pragma unroll
                for (int q = 0; q < kAdMaxMat; ++q) ws[q] = ac * sacc[q] + ((uint32_t)q == mat ? ds : 0.0f);
                aq.p[slot] = make_float4(p.x, p.y, p.z, 0.0f);
                aq.m[slot] = make_uint4(r | ((uint32_t)band << 8) | (mat << 16), id, 0u, 0u);
                aq.w[4 * (size_t)slot] = make_float4(ac * drefl[0], ac * drefl[1], ac * drefl[2], ac * drefl[3]);
                aq.w[4 * (size_t)slot + 1] =
                    make_float4(ac * drefl[4], ac * drefl[5], ac * drefl[6], ac * drefl[7]);
                aq.w[4 * (size_t)slot + 2] = make_float4(ws[0], ws[1], ws[2], ws[3]);
                aq.w[4 * (size_t)slot + 3] = make_float4(ws[4], ws[5], ws[6], ws[7]);
            }
        }
        if (SCORE && !CONT && cont && !contn) {
            // the direction just drawn leaves this vertex: its score, the rim
            // flux and perhaps a continuation (after this vertex's connections,
            // which that direction does not reach)
            const RenderParams& P = inst[id].P;
            const double cdir = (double)vdir.x * (double)rdir.x + (double)vdir.y * (double)rdir.y +
                                (double)vdir.z * (double)rdir.z;
            const double cr = (double)rdir.x * (double)nrm.x + (double)rdir.y * (double)nrm.y +
                              (double)rdir.z * (double)nrm.z;
            const double F = score_flux(cr, (double)spec);
            double coef = score_dlogp(cdir, (double)spec) - F;
            Rng rs;
            score_rng_init(rs, P.seed, ray, P.ctx_word[band], P.instance, k, 0u);
            const float qk = fminf(1.0f, (float)(F * (double)kScoreRate));
            if (rs.next() < qk) {
                vb = score_rim_dir(rs, nrm, rdir, spec);
                const double wgt = F / (double)qk;
                coef += wgt;
                wcont = (float)(-wgt);
                sdraw = rs.draw;
                spawn = true;
            }
#pragma unroll
            for (int q = 0; q < kAdMaxMat; ++q)
                if ((uint32_t)q == mat) sacc[q] += (float)coef;
        }
        const uint32_t slot = block_append<kBlock>(n_out, cont, sc[2 + parity]);
        if (cont) {
            qout.o[slot] = no;
            qout.d[slot] = nd;
            qout.m[slot] = nmeta;
#pragma unroll
            for (int q = 0; q < kAdMaxMat; ++q) {
                if ((uint32_t)q < nm) {
                    tout.refl(q, slot) = drefl[q];
                    if (SCORE) tout.row(q, slot) = sacc[q];
                }
            }
            if (SCORE && contn) tout.row(nm, slot) = __uint_as_float(kspawn);
        }
        if (SCORE && !CONT) {
            const uint32_t cs = block_append<kBlock>(co.n, spawn, sc[4 + parity]);
            if (spawn) {
                co.q.o[cs] = no;
                co.q.d[cs] = make_float4(vb.x, vb.y, vb.z, L);
                co.q.m[cs] = make_uint4(ray, sdraw, (uint32_t)band | (k << 8) | kContFlag, id);
#pragma unroll
                for (int q = 0; q < kAdMaxMat; ++q) {
                    if ((uint32_t)q < nm) {
                        co.t.refl(q, cs) = 0.0f;
                        co.t.row(q, cs) = (uint32_t)q == mat ? wcont : 0.0f;
                    }
                }
                co.t.row(nm, cs) = __uint_as_float(k);
            }
        }
        n_done += i < n ? 1u : 0u;
    }
    stat_add(stats, ST_SEGMENTS, n_done);
}

// Visibility of the local adjoint's connection requests (Scene::Connect,
// src/Scene.cpp:84-96): the lane-refilling any-hit walk of k_shadow_f; a
// visible request adds its weights to the gradient of its instance and band.
template <int WIDTH, bool BATCH>
__global__ void __launch_bounds__(kBlock) k_adj_vis(const InstDev* __restrict__ inst, int* gstack, uint32_t gthreads,
                                                     AdjShadowQ aq, const uint32_t* __restrict__ n_aq, AdjArgs A0,
                                                     const AdjArgs* __restrict__ adj,
                                                     unsigned long long* __restrict__ stats, Fetch fetch) {
    constexpr int LDS = 16;
    __shared__ int lds[LDS * kBlock];
    const uint32_t n = *n_aq;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t lane = lane_id();
    uint32_t nvis = 0, ntri = 0, nsh = 0;
    Walker<LDS, kBlock, WIDTH, true, false> w(lds + threadIdx.x, gstack, gtid, gthreads);
    Traversal T = make_trav(inst[0], gstack, gthreads);
    double g_refl[kAdMaxMat], g_spec[kAdMaxMat];
#pragma unroll
    for (int q = 0; q < kAdMaxMat; ++q) g_refl[q] = g_spec[q] = 0.0;
    uint32_t key = 0xFFFFFFFFu;  // instance of the accumulated partials
    int kband = -1;
    bool active = false, pending = false, exhausted = n == 0;
    uint32_t slot = 0;
    uint32_t c_next = 0, c_end = 0;
    uint32_t seg = (gtid >> 6) % fetch.n_seg;
    for (;;) {
        uint64_t act = __ballot(active);
        const uint32_t n_free = 64u - (uint32_t)__popcll(act);
        if (n_free >= fetch.refill || act == 0) {
            if (pending) {
                const uint4 m = aq.m[slot];
                const uint32_t id = BATCH ? m.y : 0u;
                const int band = (int)((m.x >> 8) & 0xFFu);
                if (BATCH && key != 0xFFFFFFFFu && (id != key || band != kband)) {
                    // a batch lane moving to another instance or band hands its partials on
                    const AdjArgs& A = adj[key];
                    double* gk = A.grad_copy();
#pragma unroll
                    for (int q = 0; q < kAdMaxMat; ++q) {
                        if ((uint32_t)q < A.n_mat) {
                            if (g_refl[q] != 0.0) atomicAdd(gk + ((size_t)kband * 2 + 0) * A.n_mat + q, g_refl[q]);
                            if (g_spec[q] != 0.0) atomicAdd(gk + ((size_t)kband * 2 + 1) * A.n_mat + q, g_spec[q]);
                        }
                        g_refl[q] = g_spec[q] = 0.0;
                    }
                }
                key = id;
                kband = band;
                const float4 w0 = aq.w[4 * (size_t)slot], w1 = aq.w[4 * (size_t)slot + 1];
                const float4 w2 = aq.w[4 * (size_t)slot + 2], w3 = aq.w[4 * (size_t)slot + 3];
                const float wq[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
                const float wsq[8] = {w2.x, w2.y, w2.z, w2.w, w3.x, w3.y, w3.z, w3.w};
#pragma unroll
                for (int q = 0; q < kAdMaxMat; ++q) {
                    g_refl[q] += (double)wq[q];
                    g_spec[q] += (double)wsq[q];
                }
                pending = false;
            }
            if (!exhausted) {
                const uint32_t mine = take_entries(fetch, n, act, lane, c_next, c_end, exhausted, seg);
                if (mine != 0xFFFFFFFFu) {
                    slot = mine;
                    ++nsh;
                    const float4 pw = aq.p[mine];
                    const uint4 m = aq.m[mine];
                    const uint32_t id = BATCH ? m.y : 0u;
                    if (BATCH) T = make_trav(inst[id], gstack, gthreads);
                    const RecParams& R = inst[id].P.rec[m.x & 0xFFu];
                    const V3 p{pw.x, pw.y, pw.z};
                    if (w.start(T, p, sub(V3{R.px, R.py, R.pz}, p), 0.0f))
                        active = true;
                    else
                        pending = true;  // empty scene: visible
                }
            }
            act = __ballot(active);
        }
        // the loop is left by the whole wave at once (take_entries is a
        // wave-wide step)
        if (act == 0) {
            if (exhausted && __ballot(pending) == 0) break;
            continue;
        }
        if (active && w.step(T, nvis, ntri)) {
            active = false;
            pending = !w.found;
        }
    }
    stat_add(stats, ST_SHADOW, nsh);
    if (!BATCH) {
        __shared__ double red[kBlock / 64][2 * kAdMaxMat];
        block_grad_flush(A0, kband, g_refl, g_spec, &red[0][0]);
        return;
    }
    // what is left, reduced per instance and band present in the wave
    uint64_t todo = __ballot(key != 0xFFFFFFFFu);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t lk = __shfl(key, leader);
        const int lb = __shfl(kband, leader);
        const bool mine = key == lk && kband == lb;
        const uint32_t nmk = adj[lk].n_mat;
        double* gk = adj[lk].grad_copy();
#pragma unroll
        for (int q = 0; q < kAdMaxMat; ++q) {
            if ((uint32_t)q >= nmk) break;
            const double sr = wave_sum_d(mine ? g_refl[q] : 0.0);
            const double ss = wave_sum_d(mine ? g_spec[q] : 0.0);
            if (lane == 0) {
                if (sr != 0.0) atomicAdd(gk + ((size_t)lb * 2 + 0) * nmk + q, sr);
                if (ss != 0.0) atomicAdd(gk + ((size_t)lb * 2 + 1) * nmk + q, ss);
            }
        }
        todo &= ~__ballot(mine);
    }
}

__global__ void k_zero_f32(float* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

}  // namespace

// Gradient of sum(adj_value * value) + sum(adj_abs * abs_value) with respect
// to refl and spec of every material, for each band of the mask, for every
// item of a batch.  The bands are replayed one after another (the tangent
// arrays and the per-wave partials belong to one band at a time); within a
// band the rays of all items share the queues, in chunks of the ray chunk.
ear_status run_backward_batch(Ctx* c, uint32_t n_items, DeviceScene* const* scenes,
                              const ear_render_desc* const* descs, const double* const* adj_value,
                              const double* const* adj_abs, double* const* grad_refl, double* const* grad_spec) {
    if (n_items == 0) return EAR_ERR_INVALID_ARGUMENT;
    const bool batch = n_items > 1;
    const int width = (int)scenes[0]->width;
    std::vector<InstDev> inst(n_items);
    uint32_t nm_max = 0, max_rec = 0, max_bounces = 0;
    size_t adj_cells = 0;
    std::vector<size_t> adj_off(n_items);
    for (uint32_t i = 0; i < n_items; ++i) {
        const DeviceScene* s = scenes[i];
        const ear_render_desc* D = descs[i];
        if (!s || !D) return EAR_ERR_INVALID_ARGUMENT;
        if ((int)s->width != width) {
            c->last_error = "all scenes of a batch must have the same BVH width";
            return EAR_ERR_INVALID_ARGUMENT;
        }
        InstDev& I = inst[i];
        std::memset(&I, 0, sizeof(I));
        TRY(make_params(c, D, &I.P));
        if (D->source.kind != EAR_SRC_POINT) {
            c->last_error = "the adjoint supports point sources";
            return EAR_ERR_INVALID_ARGUMENT;
        }
        if (s->n_mat > (uint32_t)kAdMaxMat || s->n_mat == 0) {
            c->last_error = "the adjoint supports 1..8 materials";
            return EAR_ERR_CAPACITY;
        }
        if (!adj_value[i] && !adj_abs[i]) {
            c->last_error = "no adjoint given";
            return EAR_ERR_INVALID_ARGUMENT;
        }
        if ((D->flags & EAR_FLAG_ADJ_SCORE) && (D->flags & EAR_FLAG_ADJ_LOCAL)) {
            c->last_error = "EAR_FLAG_ADJ_SCORE and EAR_FLAG_ADJ_LOCAL exclude each other";
            return EAR_ERR_INVALID_ARGUMENT;
        }
        if (D->flags & EAR_FLAG_ADJ_SCORE) {
            // the blend's direction map is one-to-one only below 1/2
            for (uint32_t m = 0; m < s->n_mat; ++m)
                for (uint32_t b = 0; b < 3; ++b)
                    if (((D->band_mask >> b) & 1u) && !(s->mat_band_host[3 * m + b].y < 0.5f)) {
                        c->last_error = "EAR_FLAG_ADJ_SCORE needs every specularity below 0.5";
                        return EAR_ERR_INVALID_ARGUMENT;
                    }
        }
        I.nodes = s->nodes;
        I.nodes4 = s->nodes4;
        I.tris = s->tris;
        I.nm = s->nm;
        I.mat_band = s->mat_band;
        I.n_tris = s->n_tris;
        nm_max = std::max(nm_max, s->n_mat);
        max_rec = std::max(max_rec, I.P.n_rec);
        max_bounces = std::max(max_bounces, D->max_bounces);
        adj_off[i] = adj_cells;
        adj_cells += (size_t)3 * I.P.n_channels * I.P.n_bins;
    }
    // value = raw * (1/N) * gain^2 (k_finalize), so dL/draw = dL/dvalue * (1/N) * gain^2;
    // each row is then turned into prefix sums (ramp_adjoint)
    std::vector<double> as(adj_cells, 0.0), aa(adj_cells, 0.0);
    bool any_s = false, any_a = false;
    for (uint32_t i = 0; i < n_items; ++i) {
        const ear_render_desc* D = descs[i];
        const double scale = (double)(1.0f / (float)std::max<uint32_t>(1, D->n_rays)) *
                             (double)(D->source.gain * D->source.gain);
        const size_t hn = (size_t)3 * inst[i].P.n_channels * inst[i].P.n_bins;
        if (adj_value[i]) {
            any_s = true;
            for (size_t k = 0; k < hn; ++k) as[adj_off[i] + k] = adj_value[i][k] * scale;
        }
        if (adj_abs[i]) {
            any_a = true;
            for (size_t k = 0; k < hn; ++k) aa[adj_off[i] + k] = adj_abs[i][k] * scale;
        }
    }
    const ear_tuning& tu = c->tuning;
    const uint32_t chunk = tu.ray_chunk ? tu.ray_chunk : (1u << 23);
    uint64_t rays_any = 0;
    for (uint32_t i = 0; i < n_items; ++i) rays_any = std::max<uint64_t>(rays_any, descs[i]->n_rays);
    uint64_t rays_band_max = 0;
    for (int b = 0; b < 3; ++b) {
        uint64_t t = 0;
        for (uint32_t i = 0; i < n_items; ++i)
            if ((inst[i].P.band_mask >> b) & 1u) t += descs[i]->n_rays;
        rays_band_max = std::max(rays_band_max, t);
    }
    const uint32_t cap = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(chunk, rays_band_max));
    hipStream_t st = c->stream;
    DevBuf qo[2] = {DevBuf(c), DevBuf(c)}, qd[2] = {DevBuf(c), DevBuf(c)}, qm[2] = {DevBuf(c), DevBuf(c)};
    DevBuf tt[2] = {DevBuf(c), DevBuf(c)}, tr[2] = {DevBuf(c), DevBuf(c)};
    DevBuf hits(c), counters(c), stats(c), gs(c), dinst(c), dsegs(c), dsegoff(c), dplan(c), dadj_s(c), dadj_a(c),
        dgrad(c), dargs(c), sqp(c), sqm(c);
    for (int k = 0; k < 2; ++k) {
        TRY(qo[k].alloc((size_t)cap * 16));
        TRY(qd[k].alloc((size_t)cap * 16));
        TRY(qm[k].alloc((size_t)cap * 16));
        TRY(tt[k].alloc((size_t)cap * nm_max * 8 * 4));
        TRY(tr[k].alloc((size_t)cap * nm_max * 4));
    }
    TRY(counters.alloc(32));
    TRY(stats.alloc(ST_N * 8 * kStatCopies));
    TRY(sqp.alloc(16));
    TRY(sqm.alloc(16));
    // connection requests of the local adjoint (one per ray and recorder)
    uint32_t n_local = 0, n_score = 0;
    for (uint32_t i = 0; i < n_items; ++i) {
        n_local += (inst[i].P.flags & EAR_FLAG_ADJ_LOCAL) ? 1u : 0u;
        n_score += (inst[i].P.flags & EAR_FLAG_ADJ_SCORE) ? 1u : 0u;
    }
    if (n_score != 0 && n_score != n_items) {
        c->last_error = "a batch mixes EAR_FLAG_ADJ_SCORE items with others";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    const bool all_score = n_score == n_items;
    const bool all_local = all_score || n_local == n_items;  // the per-vertex kernels
    DevBuf aqp(c), aqm(c), aqw(c), segv(c);
    // the score form's continuations: traced one bounce per main bounce in
    // their own queues (qc), then to their ends after the chunk's main rays;
    // room for four chunks (about the most that are alive at once: spawns
    // build up as ~0.05 F per live ray and bounce), a drain when full
    const uint32_t cap_c = all_score ? (uint32_t)std::min<uint64_t>(4ull * cap, 1ull << 25) : 1u;
    const uint32_t cap_q = std::max(cap, cap_c);  // rays of one bounce: the chunk or a drain
    TRY(hits.alloc((size_t)cap_q * 8));
    const size_t cap_aq = all_local ? (size_t)cap_q * std::max<uint32_t>(1, max_rec) : 1;
    TRY(aqp.alloc(cap_aq * 16));
    TRY(aqm.alloc(cap_aq * 16));
    TRY(aqw.alloc(cap_aq * 64));
    DevBuf qco[2] = {DevBuf(c), DevBuf(c)}, qcd[2] = {DevBuf(c), DevBuf(c)}, qcm[2] = {DevBuf(c), DevBuf(c)};
    DevBuf tct[2] = {DevBuf(c), DevBuf(c)}, tcr[2] = {DevBuf(c), DevBuf(c)}, cntc(c);
    for (int k = 0; k < 2; ++k) {
        TRY(qco[k].alloc((size_t)cap_c * 16));
        TRY(qcd[k].alloc((size_t)cap_c * 16));
        TRY(qcm[k].alloc((size_t)cap_c * 16));
        TRY(tct[k].alloc((size_t)cap_c * (nm_max + 1) * 4));
        TRY(tcr[k].alloc((size_t)cap_c * nm_max * 4));
    }
    TRY(cntc.alloc(8));
    EAR_HIP_TRY(c, hipMemsetAsync(cntc.p, 0, cntc.n, st));
    TRY(segv.alloc(kMaxSeg * kSegStride * 4));
    EAR_HIP_TRY(c, hipMemsetAsync(segv.p, 0, segv.n, st));
    const AdjShadowQ AQ{aqp.as<float4>(), aqm.as<uint4>(), aqw.as<float4>()};
    // static closest-hit traversal by default (the replayed bands run one
    // at a time, with fewer rays in flight than the forward's, where the
    // refill bookkeeping does not pay); the refilling one on request
    const bool fetch_e = tu.traversal == EAR_TRAVERSAL_FETCH || tu.traversal == EAR_TRAVERSAL_FETCH_ALL;
    ExtendFn fe;
    ShadowFn fs_unused;
    const TravCfg cfg{256, 16, width, false, batch, fetch_e, false};
    pick_kernels(cfg, &fe, &fs_unused);
    DevBuf segc(c);
    TRY(segc.alloc(kMaxSeg * kSegStride * 4));
    EAR_HIP_TRY(c, hipMemsetAsync(segc.p, 0, segc.n, st));
    const uint32_t nseg = tu.fetch_segments ? std::min<uint32_t>(tu.fetch_segments, kMaxSeg) : 64u;
    const Fetch fq_e{segc.as<uint32_t>(), tu.refill ? tu.refill : 24u, tu.fetch_chunk, nseg};
    const uint32_t waves = tu.waves_per_cu ? tu.waves_per_cu : 32;
    const uint32_t cap_e = (uint32_t)c->num_cus * (uint32_t)std::max(1, resident_blocks((const void*)fe, 256));
    const uint32_t gmax = persistent_grid(c, 0xFFFFFFFFu, waves);
    // every thread of either traversal launch owns one overflow stack
    const uint32_t gthreads = std::max(cap_e * 256u, gmax * (uint32_t)kBlock);
    TRY(gs.alloc((size_t)(kStackDepth - kLdsMin) * gthreads * 4));
    TRY(dinst.alloc(sizeof(InstDev) * n_items));
    TRY(dsegs.alloc(sizeof(GenSeg) * n_items));
    TRY(dsegoff.alloc(4 * (size_t)n_items));
    TRY(dplan.alloc(8));
    const uint32_t grad_n = 3 * 2 * nm_max * n_items;  // one copy
    TRY(dgrad.alloc(sizeof(double) * (size_t)grad_n * kGradCopies));
    TRY(dargs.alloc(sizeof(AdjArgs) * n_items));
    // prefix sums of every row (bins + 1 pairs), accumulated in long double
    std::vector<size_t> pre_off(n_items);
    size_t pre_pairs = 0;
    for (uint32_t i = 0; i < n_items; ++i) {
        pre_off[i] = pre_pairs;
        pre_pairs += (size_t)3 * inst[i].P.n_channels * (inst[i].P.n_bins + 1);
    }
    auto prefix = [&](const std::vector<double>& a) {
        std::vector<double> pre(2 * pre_pairs, 0.0);
        for (uint32_t i = 0; i < n_items; ++i) {
            const uint32_t nb = inst[i].P.n_bins, rows = 3 * inst[i].P.n_channels;
            for (uint32_t r = 0; r < rows; ++r) {
                const double* src = a.data() + adj_off[i] + (size_t)r * nb;
                double* dst = pre.data() + 2 * (pre_off[i] + (size_t)r * (nb + 1));
                long double s1 = 0.0L, s2 = 0.0L;
                for (uint32_t b = 0; b <= nb; ++b) {
                    dst[2 * b] = (double)s1;
                    dst[2 * b + 1] = (double)s2;
                    if (b < nb) {
                        s1 += (long double)src[b];
                        s2 += (long double)b * (long double)src[b];
                    }
                }
            }
        }
        return pre;
    };
    if (any_s) {
        const std::vector<double> pre = prefix(as);
        TRY(dadj_s.alloc(pre.size() * 8));
        EAR_HIP_TRY(c, hipMemcpyAsync(dadj_s.p, pre.data(), pre.size() * 8, hipMemcpyHostToDevice, st));
        EAR_HIP_TRY(c, hipStreamSynchronize(st));  // pre leaves scope
    }
    if (any_a) {
        const std::vector<double> pre = prefix(aa);
        TRY(dadj_a.alloc(pre.size() * 8));
        EAR_HIP_TRY(c, hipMemcpyAsync(dadj_a.p, pre.data(), pre.size() * 8, hipMemcpyHostToDevice, st));
        EAR_HIP_TRY(c, hipStreamSynchronize(st));
    }
    std::vector<AdjArgs> args(n_items);
    for (uint32_t i = 0; i < n_items; ++i)
        args[i] = AdjArgs{adj_value[i] ? dadj_s.as<double2>() + pre_off[i] : nullptr,
                          adj_abs[i] ? dadj_a.as<double2>() + pre_off[i] : nullptr,
                          dgrad.as<double>() + (size_t)i * 3 * 2 * nm_max, scenes[i]->n_mat, grad_n};
    EAR_HIP_TRY(c, hipMemcpyAsync(dargs.p, args.data(), sizeof(AdjArgs) * n_items, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(dinst.p, inst.data(), sizeof(InstDev) * n_items, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemsetAsync(dgrad.p, 0, dgrad.n, st));
    EAR_HIP_TRY(c, hipMemsetAsync(counters.p, 0, counters.n, st));
    EAR_HIP_TRY(c, hipMemsetAsync(stats.p, 0, stats.n, st));
    const InstDev* di = dinst.as<InstDev>();
    uint32_t* cnt = counters.as<uint32_t>();
    unsigned long long* stp = stats.as<unsigned long long>();
    uint32_t pinned[4];
    Queue q[2] = {Queue{qo[0].as<float4>(), qd[0].as<float4>(), qm[0].as<uint4>()},
                  Queue{qo[1].as<float4>(), qd[1].as<float4>(), qm[1].as<uint4>()}};
    Tangents T[2] = {Tangents{tt[0].as<float>(), tr[0].as<float>(), cap},
                     Tangents{tt[1].as<float>(), tr[1].as<float>(), cap}};
    Queue qc[2] = {Queue{qco[0].as<float4>(), qcd[0].as<float4>(), qcm[0].as<uint4>()},
                   Queue{qco[1].as<float4>(), qcd[1].as<float4>(), qcm[1].as<uint4>()}};
    Tangents Tc[2] = {Tangents{tct[0].as<float>(), tcr[0].as<float>(), cap_c},
                      Tangents{tct[1].as<float>(), tcr[1].as<float>(), cap_c}};
    uint32_t* ncc = cntc.as<uint32_t>();
    int ci = 0;              // the continuations' input buffer
    uint32_t n_cont = 0;     // continuations waiting there
    bool cont_live = false;  // any may be waiting (a main bounce has run since the last drain)
    uint32_t pinned_c[2];
    const ShadowQ SQ{sqp.as<float4>(), sqm.as<float4>()};
    const AdjArgs A0 = args[0];
    const AdjArgs* dA = dargs.as<AdjArgs>();
    // the visibility kernel: lane refilling over the request queue
    using VisFn = void (*)(const InstDev*, int*, uint32_t, AdjShadowQ, const uint32_t*, AdjArgs, const AdjArgs*,
                           unsigned long long*, Fetch);
    const VisFn fv = width == 4 ? (batch ? k_adj_vis<4, true> : k_adj_vis<4, false>)
                                : (batch ? k_adj_vis<2, true> : k_adj_vis<2, false>);
    const uint32_t cap_v = (uint32_t)c->num_cus * (uint32_t)std::max(1, resident_blocks((const void*)fv, kBlock));
    // grids of the per-vertex shading kernels (no walk inside them, so no
    // overflow stacks to cover) and of the continuations' shading
    const uint32_t cap_sl =
        rounds_cap(c,
                   batch ? (all_score ? (const void*)k_shade_adl<true, true> : (const void*)k_shade_adl<true, false>)
                         : (all_score ? (const void*)k_shade_adl<false, true> : (const void*)k_shade_adl<false, false>),
                   tu.waves_per_cu, kBlock);
    const uint32_t cap_sc = rounds_cap(
        c, batch ? (const void*)k_shade_adl<true, true, true> : (const void*)k_shade_adl<false, true, true>,
        tu.waves_per_cu, kBlock);
    const Fetch fq_v{segv.as<uint32_t>(), tu.refill ? tu.refill : 24u, tu.fetch_chunk, nseg};
    // one bounce of the per-vertex adjoint (local or score form): closest
    // hits of queue qi, the shade (continuing rays into qo, continuations
    // into co), the visibility of the connection requests
    auto adl_bounce = [&](const Queue& qi, uint32_t* ni, const Tangents& ti, const Queue& qo, uint32_t* no,
                          const Tangents& to, const ContOut& co, uint32_t n_now, uint32_t gsd, uint32_t ge,
                          bool conts) {
        hipLaunchKernelGGL(fe, dim3(ge), dim3(256), 0, st, di, gs.as<int>(), gthreads, qi, (const uint32_t*)ni,
                           hits.as<int2>(), stp, fq_e);
        if (fetch_e) k_reset_iter<<<1, 64, 0, st>>>(cnt, 0u, segc.as<uint32_t>(), nseg);
        if (conts) {
            if (batch)
                k_shade_adl<true, true, true><<<gsd, kBlock, 0, st>>>(di, qi, ni, hits.as<int2>(), ti, qo, no, to,
                                                                      AQ, cnt + 2, max_rec, A0, dA, co, stp);
            else
                k_shade_adl<false, true, true><<<gsd, kBlock, 0, st>>>(di, qi, ni, hits.as<int2>(), ti, qo, no, to,
                                                                       AQ, cnt + 2, max_rec, A0, dA, co, stp);
        } else if (batch) {
            if (all_score)
                k_shade_adl<true, true><<<gsd, kBlock, 0, st>>>(di, qi, ni, hits.as<int2>(), ti, qo, no, to, AQ,
                                                                cnt + 2, max_rec, A0, dA, co, stp);
            else
                k_shade_adl<true, false><<<gsd, kBlock, 0, st>>>(di, qi, ni, hits.as<int2>(), ti, qo, no, to, AQ,
                                                                 cnt + 2, max_rec, A0, dA, co, stp);
        } else {
            if (all_score)
                k_shade_adl<false, true><<<gsd, kBlock, 0, st>>>(di, qi, ni, hits.as<int2>(), ti, qo, no, to, AQ,
                                                                 cnt + 2, max_rec, A0, dA, co, stp);
            else
                k_shade_adl<false, false><<<gsd, kBlock, 0, st>>>(di, qi, ni, hits.as<int2>(), ti, qo, no, to, AQ,
                                                                  cnt + 2, max_rec, A0, dA, co, stp);
        }
        const uint32_t n_req = (uint32_t)std::min<size_t>(cap_aq, (size_t)n_now * max_rec);
        hipLaunchKernelGGL(fv, dim3(grid_for(n_req, cap_v, kBlock)), dim3(kBlock), 0, st, di, gs.as<int>(), gthreads,
                           AQ, (const uint32_t*)(cnt + 2), A0, dA, stp, fq_v);
        k_reset_iter<<<1, 64, 0, st>>>(cnt, 0u, segv.as<uint32_t>(), nseg);
        k_reset2<<<1, 64, 0, st>>>(ni, cnt + 2);
    };
    // the waiting continuations traced to their ends (they spawn nothing,
    // so a queue never grows within a drain)
    auto drain = [&]() -> ear_status {
        cont_live = false;
        while (n_cont > 0) {
            const uint32_t gsd = grid_for(n_cont, cap_sc, kBlock);
            const uint32_t ge = grid_for(n_cont, cap_e, 256);
            for (int gi = 0; gi < 4; ++gi) {
                adl_bounce(qc[ci], ncc + ci, Tc[ci], qc[1 - ci], ncc + (1 - ci), Tc[1 - ci],
                           ContOut{qc[ci], ncc + ci, Tc[ci]}, n_cont, gsd, ge, true);
                ci = 1 - ci;
            }
            EAR_HIP_TRY(c, hipMemcpyAsync(pinned_c, ncc, 8, hipMemcpyDeviceToHost, st));
            EAR_HIP_TRY(c, hipStreamSynchronize(st));
            n_cont = pinned_c[ci];
        }
        return EAR_OK;
    };
    for (uint32_t band = 0; band < 3; ++band) {
        // this band's contexts as runs of rays
        std::vector<GenSeg> segs;
        std::vector<uint32_t> off;
        uint32_t total = 0;
        for (uint32_t i = 0; i < n_items; ++i) {
            if (!((inst[i].P.band_mask >> band) & 1u) || descs[i]->n_rays == 0) continue;
            segs.push_back(GenSeg{i, band, 0, descs[i]->n_rays});
            off.push_back(total);
            total += descs[i]->n_rays;
        }
        if (segs.empty()) continue;
        EAR_HIP_TRY(c, hipMemcpyAsync(dsegs.p, segs.data(), sizeof(GenSeg) * segs.size(), hipMemcpyHostToDevice, st));
        EAR_HIP_TRY(c, hipMemcpyAsync(dsegoff.p, off.data(), 4 * off.size(), hipMemcpyHostToDevice, st));
        for (uint32_t r0 = 0; r0 < total; r0 += cap) {
            const uint32_t n_chunk = std::min(cap, total - r0);
            const uint32_t plan[2] = {r0, n_chunk};
            EAR_HIP_TRY(c, hipMemcpyAsync(dplan.p, plan, 8, hipMemcpyHostToDevice, st));
            if (batch)
                k_generate<true><<<persistent_grid(c, n_chunk, waves), kBlock, 0, st>>>(
                    di, dsegs.as<GenSeg>(), dsegoff.as<uint32_t>(), (uint32_t)segs.size(), n_chunk,
                    dplan.as<uint32_t>(), max_rec, 0u, q[0], cnt + 0, SQ, cnt + 2, stp);
            else
                k_generate<false><<<persistent_grid(c, n_chunk, waves), kBlock, 0, st>>>(
                    di, dsegs.as<GenSeg>(), dsegoff.as<uint32_t>(), (uint32_t)segs.size(), n_chunk,
                    dplan.as<uint32_t>(), max_rec, 0u, q[0], cnt + 0, SQ, cnt + 2, stp);
            if (!all_local || all_score) k_zero_f32<<<1024, 256, 0, st>>>(T[0].t, (size_t)cap * nm_max * 8);
            k_zero_f32<<<1024, 256, 0, st>>>(T[0].r, (size_t)cap * nm_max);
            EAR_HIP_TRY(c, hipMemcpyAsync(pinned, cnt, 12, hipMemcpyDeviceToHost, st));
            EAR_HIP_TRY(c, hipStreamSynchronize(st));
            uint32_t n_live = pinned[0];
            int cur = 0;
            // bounces in groups of 4 between looks at the queue count: the
            // kernels read their counts on the device, and a queue never grows
            // within a chunk, so grids sized at a group's start cover it
            uint32_t k = 1;
            while (k <= max_bounces && n_live > 0) {
              uint32_t group = std::min<uint32_t>(4u, max_bounces - k + 1);
              if (all_score) {
                  // every bounce may append one continuation per live ray, and
                  // the continuations' own bounce never adds any
                  if (cap_c - n_cont < n_live) TRY(drain());
                  group = std::min<uint32_t>(group, (cap_c - n_cont) / n_live);
              }
              // k_shade_ad walks inside: its grid stays within the overflow stacks' gmax
              const uint32_t gsd = all_local ? grid_for(n_live, cap_sl, kBlock) : persistent_grid(c, n_live, waves);
              const uint32_t ge = grid_for(n_live, cap_e, 256);
              const uint32_t n_cmax = all_score ? n_cont + group * n_live : 0u;
              const uint32_t gsd_c = grid_for(std::max(1u, n_cmax), cap_sc, kBlock);
              const uint32_t ge_c = grid_for(std::max(1u, n_cmax), cap_e, 256);
              for (uint32_t gi = 0; gi < group; ++gi, ++k) {
                if (all_local) {
                    // the main bounce: its spawns join the continuations' next input
                    adl_bounce(q[cur], cnt + cur, T[cur], q[1 - cur], cnt + (1 - cur), T[1 - cur],
                               ContOut{qc[1 - ci], ncc + (1 - ci), Tc[1 - ci]}, n_live, gsd, ge, false);
                    if (all_score) {
                        // one bounce of the waiting continuations, beside the main rays
                        if (cont_live)
                            adl_bounce(qc[ci], ncc + ci, Tc[ci], qc[1 - ci], ncc + (1 - ci), Tc[1 - ci],
                                       ContOut{qc[ci], ncc + ci, Tc[ci]}, n_cmax, gsd_c, ge_c, true);
                        cont_live = true;
                        ci = 1 - ci;
                    }
                } else {
                    hipLaunchKernelGGL(fe, dim3(ge), dim3(256), 0, st, di, gs.as<int>(), gthreads, q[cur],
                                       (const uint32_t*)(cnt + cur), hits.as<int2>(), stp, fq_e);
                    if (fetch_e) k_reset_iter<<<1, 64, 0, st>>>(cnt, 0u, segc.as<uint32_t>(), nseg);
#define EAR_AD_LAUNCH(W, B)                                                                                      \
    k_shade_ad<W, B><<<gsd, kBlock, 0, st>>>(di, gs.as<int>(), gthreads, q[cur], cnt + cur, hits.as<int2>(), \
                                             T[cur], q[1 - cur], cnt + (1 - cur), T[1 - cur], A0, dA, stp)
                    if (width == 4) {
                        if (batch)
                            EAR_AD_LAUNCH(4, true);
                        else
                            EAR_AD_LAUNCH(4, false);
                    } else {
                        if (batch)
                            EAR_AD_LAUNCH(2, true);
                        else
                            EAR_AD_LAUNCH(2, false);
                    }
#undef EAR_AD_LAUNCH
                    k_reset2<<<1, 64, 0, st>>>(cnt + cur, cnt + 2);
                }
                cur = 1 - cur;
              }
              EAR_HIP_TRY(c, hipMemcpyAsync(pinned, cnt, 12, hipMemcpyDeviceToHost, st));
              if (all_score) EAR_HIP_TRY(c, hipMemcpyAsync(pinned_c, ncc, 8, hipMemcpyDeviceToHost, st));
              EAR_HIP_TRY(c, hipStreamSynchronize(st));
              n_live = pinned[cur];
              if (all_score) n_cont = pinned_c[ci];
            }
            if (all_score) TRY(drain());
            k_reset2<<<1, 64, 0, st>>>(cnt + 0, cnt + 1);
            EAR_HIP_TRY(c, hipGetLastError());
        }
    }
    k_sum_grad_copies<<<std::min<uint32_t>((grad_n + 255) / 256, 1024), 256, 0, st>>>(dgrad.as<double>(), grad_n,
                                                                                      grad_n);
    EAR_HIP_TRY(c, hipGetLastError());
    std::vector<double> g(grad_n);
    EAR_HIP_TRY(c, hipMemcpyAsync(g.data(), dgrad.p, g.size() * 8, hipMemcpyDeviceToHost, st));
    EAR_HIP_TRY(c, hipStreamSynchronize(st));
    for (uint32_t i = 0; i < n_items; ++i) {
        const uint32_t nm = scenes[i]->n_mat;
        const double* gi = g.data() + (size_t)i * 3 * 2 * nm_max;
        for (uint32_t b = 0; b < 3; ++b) {
            if (!((inst[i].P.band_mask >> b) & 1u)) continue;
            for (uint32_t m = 0; m < nm; ++m) {
                if (grad_refl[i]) grad_refl[i][(size_t)m * 3 + b] = gi[((size_t)b * 2 + 0) * nm + m];
                if (grad_spec[i]) grad_spec[i][(size_t)m * 3 + b] = gi[((size_t)b * 2 + 1) * nm + m];
            }
        }
    }
    return EAR_OK;
}

ear_status run_backward(Ctx* c, DeviceScene* s, const ear_render_desc* D, const double* adj_value,
                        const double* adj_abs, double* grad_refl, double* grad_spec) {
    return run_backward_batch(c, 1, &s, &D, &adj_value, &adj_abs, &grad_refl, &grad_spec);
}

}  // namespace ear
