// ear_internal.hpp -- structures shared by the translation units of libear_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "ear.h"
#include "ear_math.hpp"

namespace ear {

// BVH2 node, 64 B: both children's padded AABBs and child references.
//   a = {lo0.x, lo0.y, lo0.z, hi0.x}  b = {hi0.y, hi0.z, lo1.x, lo1.y}
//   c = {lo1.z, hi1.x, hi1.y, hi1.z}  d = {child0, child1, -, -}
// child >= 0: internal node index; child < 0: leaf
//   0x80000000 | (count - 1) << 24 | first   (first < 2^24, count <= 128)
// kEmptyChild: a child slot whose box is empty (single-triangle scenes).
struct alignas(64) BvhNode {
    float4 a, b, c;
    int4 d;
};
constexpr int32_t kEmptyChild = (int32_t)0x80000000u;
constexpr uint32_t kMaxTris = 1u << 24;
// Traversal stack entries per thread (LDS part + global overflow part).  A
// 2-wide tree pushes at most one entry per level (depth <= 30 Morton bits +
// 24 index bits); a 4-wide tree at most three per level; the build checks
// the bound for the tree it produced.
constexpr int kStackDepth = 96;

// BVH4 node, 64 B: the four child boxes quantised to 8 bits per plane
// against the node's own box, with a power-of-two scale per axis (Ylitie,
// Karras, Laine, "Efficient incoherent ray traversal on GPUs through
// compressed wide BVHs", HPG 2017).  A plane decodes as
// org[axis] + q * 2^(e[axis] - 127): the product is exact, so the value
// rounds once and host and device get the same float.  The build rounds every
// lo plane down and every hi plane up, so a decoded box contains the exact
// one: traversal may visit more nodes, never fewer, and hits are unchanged.
// Byte c of each q word belongs to child c; ref as in BvhNode; kEmpty4 marks
// an unused slot.
struct alignas(64) Bvh4Node {
    uint4 org;  // x, y, z: origin (float bits); w: biased exponents ex | ey << 8 | ez << 16
    uint4 qa;   // qlo.x, qhi.x, qlo.y, qhi.y
    uint4 qb;   // qlo.z, qhi.z, 0, 0
    int4 ref;
};
constexpr int32_t kEmpty4 = 0x7FFFFFFF;
constexpr int kMaxRecorders = 16;
constexpr int kMaxChannels = 32;

// Triangles in BVH (Morton) order, 48 B each:
//   t0 = {v0.x, v0.y, v0.z, e1.x}  t1 = {e1.y, e1.z, e2.x, e2.y}  t2 = {e2.z, orig, -, -}
// plus nm = {n.x, n.y, n.z, material} with n = normalize(cross(e1, e2)).

struct DeviceScene {
    uint32_t n_tris = 0;
    uint32_t n_nodes = 0;
    uint32_t width = 2;
    uint32_t n_nodes4 = 0;
    uint32_t n_nodes4_alloc = 0;  // nodes4 entries allocated (>= n_nodes4)
    uint32_t depth4 = 0;
    double collapse_ms = 0;
    Bvh4Node* nodes4 = nullptr;
    uint32_t n_mat = 0;
    uint32_t builder = EAR_BVH_LBVH;
    uint32_t leaf_size = 4;
    uint32_t max_depth = 0;
    float bounds[6] = {0, 0, 0, 0, 0, 0};
    double build_ms = 0;
    BvhNode* nodes = nullptr;
    float4* tris = nullptr;  // 3 * n_tris
    float4* nm = nullptr;    // n_tris
    // per (material, band): {fl2 (or -1: no draw), spec, retention, -}
    float4* mat_band = nullptr;
    std::vector<float4> mat_band_host;
};

struct Ctx;
struct SceneHandle {
    Ctx* ctx = nullptr;
    DeviceScene ds;
};

// Launch parameters of one render (passed by value to kernels).
struct RecParams {
    float px, py, pz;
    float rx, ry, rz;  // right ear
    float head_size;
    float head_abs[3];
    int32_t kind;
    int32_t channel;  // first channel
    // receiver spheres (EAR_EST_SPHERE): radius^2, 3 / (4 pi r^3), centres
    // (mono: one; stereo: left and right channel)
    float r2, inv_vol;
    float cen[2][3];
    double l2_head_abs[3];  // pow_base_l2(head_abs[b])
};

struct RenderParams {
    uint64_t seed;
    uint32_t instance;
    uint32_t ctx_word[3];  // Philox counter word 2 per band
    float air[3];
    double l2_air[3];  // pow_base_l2(air[b])
    float dry;
    double cutoff;
    uint32_t max_bounces;
    uint32_t n_bins;
    uint32_t n_channels;
    uint32_t n_rec;
    uint32_t n_band_slots;  // always 3 (histograms laid out per band slot)
    uint32_t band_mask;
    uint32_t hist_kinds;    // 1 signed, 2 signed + abs
    int32_t src_kind;
    float src_pos[3];
    uint32_t src_n_tris;
    float src_total_area;
    uint32_t trace_rays, trace_segs;
    uint32_t flags;
    uint32_t estimator;  // EAR_EST_*
    RecParams rec[kMaxRecorders];
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t mem_cap = 0;
    uint64_t mem_used = 0;
    int num_cus = 256;
    std::string last_error;
    ear_tuning tuning{};
};

// A fixed device region that a render's buffers are carved from instead of
// allocated (ear_render_noalloc): bump allocation in 256-byte steps, nothing
// is freed.  `size` bytes at `base`, `used` of them taken so far.
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    size_t used = 0;
    size_t left() const { return size - used; }
};
constexpr size_t kArenaAlign = 256;

// memory helpers honouring the context's cap
ear_status dev_alloc(Ctx* c, void** p, size_t bytes);
void dev_free(Ctx* c, void* p, size_t bytes);
ear_status hip_status(Ctx* c, hipError_t e, const char* what);

// lbvh.hip
ear_status build_scene(Ctx* c, const float* v0, const float* v1, const float* v2, const uint16_t* mat_id,
                       uint32_t n_tris, const ear_material* mats, uint32_t n_mat, uint32_t leaf_size,
                       uint32_t width, uint32_t builder, DeviceScene* out);
void free_scene(Ctx* c, DeviceScene* s);
ear_status set_scene_materials(Ctx* c, DeviceScene* s, const ear_material* mats, uint32_t n_mat);

// tracer.hip
ear_status run_render(Ctx* c, DeviceScene* s, const ear_render_desc* d, ear_hist* out, ear_stats* st);
ear_status run_convolve(Ctx* c, const float* const* tracks, const uint32_t* n_track, const float* const* signals,
                        uint32_t terms, uint32_t n_sig, float* out, uint32_t n_out);
ear_status run_backward(Ctx* c, DeviceScene* s, const ear_render_desc* d, const double* adj_value,
                        const double* adj_abs, double* grad_refl, double* grad_spec);
ear_status run_backward_batch(Ctx* c, uint32_t n_items, DeviceScene* const* scenes,
                              const ear_render_desc* const* descs, const double* const* adj_value,
                              const double* const* adj_abs, double* const* grad_refl, double* const* grad_spec);
// arena != nullptr: every device buffer comes from the arena (no allocation);
// the ray queues shrink to what it holds, EAR_ERR_CAPACITY if too little.
ear_status run_batch(Ctx* c, uint32_t n_items, DeviceScene* const* scenes, const ear_render_desc* const* descs,
                     ear_hist* const* outs, ear_stats* st, Arena* arena = nullptr);
ear_status run_query_closest(Ctx* c, DeviceScene* s, uint32_t n, const float* o, const float* d, float* t,
                             int32_t* tri);
ear_status run_query_occluded(Ctx* c, DeviceScene* s, uint32_t n, const float* p, const float* q,
                              uint8_t* occ);

#define EAR_HIP_TRY(ctx, expr)                                      \
    do {                                                            \
        hipError_t _e = (expr);                                     \
        if (_e != hipSuccess) return ::ear::hip_status(ctx, _e, #expr); \
    } while (0)

// A device allocation owned by a scope (through dev_alloc / dev_free), or,
// with an arena, a piece of the arena (never freed: the arena's owner reuses
// the whole region on its next call).
struct DevBuf {
    Ctx* c;
    Arena* arena = nullptr;
    void* p = nullptr;
    size_t n = 0;
    explicit DevBuf(Ctx* cc, Arena* a = nullptr) : c(cc), arena(a) {}
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && !arena) dev_free(c, p, n);
        p = nullptr;
        n = 0;
    }
    ear_status alloc(size_t bytes) {
        release();
        n = bytes ? bytes : 16;
        if (arena) {
            const size_t take = (n + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
            if (take > arena->left()) {
                c->last_error = "render needs more device memory than the no-allocation scratch region holds (" +
                                std::to_string(arena->size >> 20) + " MiB)";
                n = 0;
                return EAR_ERR_CAPACITY;
            }
            p = arena->base + arena->used;
            arena->used += take;
            return EAR_OK;
        }
        return dev_alloc(c, &p, n);
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
};

#ifndef TRY
#define TRY(x)                       \
    do {                             \
        ear_status _s = (x);         \
        if (_s != EAR_OK) return _s; \
    } while (0)
#endif

}  // namespace ear
