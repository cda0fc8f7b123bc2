// lbvh.hip -- the BVH builds over the scene's triangle soup: the device
// binned-SAH build (the default for 4-wide trees), the device PLOC build, the
// host binned-SAH build and the device linear BVH.
//
// Replaces the reference's absence of an acceleration structure: the
// reference scans every triangle per query (src/Mesh.cpp:39-45, 61-69; the
// missing grid is noted at src/Scene.h:37-41).  Build:
//   1. per-triangle prep: e1, e2, unit normal, AABB, centroid
//   2. centroid bounds (two-level reduction)
//   3. 30-bit Morton codes of centroids
//   4. stable LSD radix sort, 4-bit digits, 8 passes (own implementation)
//   5. Karras (HPG 2012) binary radix tree, duplicate keys split by index
//   6. bottom-up AABB refit with per-node arrival counters
//   7. emit 64-byte BVH2 nodes (both child boxes in the parent, padded so the
//      box test never rejects a hit the exact triangle test accepts); any
//      subtree of at most leaf_size triangles becomes one leaf
//   8. gather triangles into Morton order as (v0, e1, e2, original index)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "ear_internal.hpp"

namespace ear {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr uint32_t kLeafFlag = 0x80000000u;

__global__ void k_prep(uint32_t n, const float* __restrict__ v0, const float* __restrict__ v1,
                       const float* __restrict__ v2, float* __restrict__ box, float* __restrict__ cen) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a[3] = {v0[3 * i], v0[3 * i + 1], v0[3 * i + 2]};
    float b[3] = {v1[3 * i], v1[3 * i + 1], v1[3 * i + 2]};
    float c[3] = {v2[3 * i], v2[3 * i + 1], v2[3 * i + 2]};
    for (int k = 0; k < 3; ++k) {
        float lo = fminf(a[k], fminf(b[k], c[k]));
        float hi = fmaxf(a[k], fmaxf(b[k], c[k]));
        box[6 * i + k] = lo;
        box[6 * i + 3 + k] = hi;
        cen[3 * i + k] = 0.5f * lo + 0.5f * hi;
    }
}

// two-level min/max reduction of centroid (6 floats) and vertex box (6 floats)
__global__ void k_bounds_partial(uint32_t n, const float* __restrict__ box, const float* __restrict__ cen,
                                 float* __restrict__ part) {
    __shared__ float s[12][256];
    float r[12];
    for (int k = 0; k < 3; ++k) {
        r[k] = INFINITY;
        r[3 + k] = -INFINITY;
        r[6 + k] = INFINITY;
        r[9 + k] = -INFINITY;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int k = 0; k < 3; ++k) {
            r[k] = fminf(r[k], cen[3 * i + k]);
            r[3 + k] = fmaxf(r[3 + k], cen[3 * i + k]);
            r[6 + k] = fminf(r[6 + k], box[6 * i + k]);
            r[9 + k] = fmaxf(r[9 + k], box[6 * i + 3 + k]);
        }
    }
    for (int k = 0; k < 12; ++k) s[k][threadIdx.x] = r[k];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            for (int k = 0; k < 12; ++k) {
                float o = s[k][threadIdx.x + st];
                bool mx = (k >= 3 && k < 6) || k >= 9;
                s[k][threadIdx.x] = mx ? fmaxf(s[k][threadIdx.x], o) : fminf(s[k][threadIdx.x], o);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 12) part[blockIdx.x * 12 + threadIdx.x] = s[threadIdx.x][0];
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void k_morton(uint32_t n, const float* __restrict__ cen, const float* __restrict__ bnd,
                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float lo = bnd[k], hi = bnd[3 + k];
        float ext = hi - lo;
        float f = ext > 0.0f ? (cen[3 * i + k] - lo) / ext : 0.0f;
        f = fminf(fmaxf(f * 1024.0f, 0.0f), 1023.0f);
        q[k] = (uint32_t)f;
    }
    keys[i] = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
    vals[i] = i;
}

// ---- radix sort ---------------------------------------------------------
__global__ void __launch_bounds__(kSortThreads) k_radix_count(uint32_t n, const uint32_t* __restrict__ keys,
                                                             uint32_t shift, uint32_t* __restrict__ counts,
                                                             uint32_t nblocks) {
    __shared__ uint32_t h[16];
    if (threadIdx.x < 16) h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile + threadIdx.x * kSortItems;
    for (int j = 0; j < kSortItems; ++j) {
        uint32_t i = base + j;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 15u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 16) counts[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// single-block exclusive scan of m uint32 values (in place)
__global__ void __launch_bounds__(1024) k_scan_u32(uint32_t m, uint32_t* __restrict__ a) {
    __shared__ uint32_t s[1024];
    uint32_t per = (m + 1023) / 1024;
    uint32_t b = threadIdx.x * per, e = min(m, b + per);
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += a[i];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - sum;
    for (uint32_t i = b; i < e; ++i) {
        uint32_t x = a[i];
        a[i] = run;
        run += x;
    }
}

__global__ void __launch_bounds__(kSortThreads) k_radix_scatter(uint32_t n, const uint32_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin,
                                                               uint32_t* __restrict__ kout,
                                                               uint32_t* __restrict__ vout, uint32_t shift,
                                                               const uint32_t* __restrict__ offs,
                                                               uint32_t nblocks) {
    __shared__ uint32_t pre[16][kSortThreads];
    uint32_t base = blockIdx.x * kSortTile + threadIdx.x * kSortItems;
    uint32_t k[kSortItems], v[kSortItems];
    uint32_t cnt[16];
    for (int d = 0; d < 16; ++d) cnt[d] = 0;
    for (int j = 0; j < kSortItems; ++j) {
        uint32_t i = base + j;
        if (i < n) {
            k[j] = kin[i];
            v[j] = vin[i];
            cnt[(k[j] >> shift) & 15u]++;
        }
    }
    for (int d = 0; d < 16; ++d) pre[d][threadIdx.x] = cnt[d];
    __syncthreads();
    // inclusive Hillis-Steele scan over threads for each digit
    for (int off = 1; off < kSortThreads; off <<= 1) {
        uint32_t t[16];
        for (int d = 0; d < 16; ++d) t[d] = threadIdx.x >= (uint32_t)off ? pre[d][threadIdx.x - off] : 0;
        __syncthreads();
        for (int d = 0; d < 16; ++d) pre[d][threadIdx.x] += t[d];
        __syncthreads();
    }
    uint32_t run[16];
    for (int d = 0; d < 16; ++d) run[d] = offs[d * nblocks + blockIdx.x] + pre[d][threadIdx.x] - cnt[d];
    for (int j = 0; j < kSortItems; ++j) {
        uint32_t i = base + j;
        if (i < n) {
            uint32_t d = (k[j] >> shift) & 15u;
            uint32_t pos = run[d]++;
            kout[pos] = k[j];
            vout[pos] = v[j];
        }
    }
}

// ---- Karras 2012 hierarchy ---------------------------------------------
__device__ __forceinline__ int delta(const uint32_t* __restrict__ keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

__global__ void k_karras(int n, const uint32_t* __restrict__ keys, uint32_t* __restrict__ left,
                         uint32_t* __restrict__ right, int* __restrict__ first, int* __restrict__ last,
                         int* __restrict__ parent_int, int* __restrict__ parent_leaf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0;
    int div = 2;
    int t = (l + div - 1) / div;
    while (true) {
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
        div <<= 1;
        t = (l + div - 1) / div;
    }
    int g = i + s * d + (d < 0 ? -1 : 0);
    int lo = min(i, j), hi = max(i, j);
    uint32_t lc = (lo == g) ? (kLeafFlag | (uint32_t)g) : (uint32_t)g;
    uint32_t rc = (hi == g + 1) ? (kLeafFlag | (uint32_t)(g + 1)) : (uint32_t)(g + 1);
    left[i] = lc;
    right[i] = rc;
    first[i] = lo;
    last[i] = hi;
    if (lc & kLeafFlag)
        parent_leaf[g] = i;
    else
        parent_int[g] = i;
    if (rc & kLeafFlag)
        parent_leaf[g + 1] = i;
    else
        parent_int[g + 1] = i;
    if (i == 0) parent_int[0] = -1;
}

__device__ __forceinline__ float coherent_load(const float* p) {
    uint32_t u = __hip_atomic_load((const uint32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __uint_as_float(u);
}

// bottom-up refit: box of internal node = union of children
__global__ void k_refit(int n, const uint32_t* __restrict__ vals, const float* __restrict__ tri_box,
                        const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                        const int* __restrict__ parent_int, const int* __restrict__ parent_leaf,
                        float* node_box, uint32_t* __restrict__ arrive) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int node = parent_leaf[s];
    while (node >= 0) {
        __threadfence();
        if (atomicAdd(&arrive[node], 1u) == 0) return;
        __threadfence();
        float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        uint32_t ch[2] = {left[node], right[node]};
        for (int c = 0; c < 2; ++c) {
            const float* src;
            bool coherent;
            if (ch[c] & kLeafFlag) {
                src = tri_box + 6 * (size_t)vals[ch[c] & ~kLeafFlag];
                coherent = false;
            } else {
                src = node_box + 6 * (size_t)ch[c];
                coherent = true;
            }
            for (int k = 0; k < 3; ++k) {
                float lo = coherent ? coherent_load(src + k) : src[k];
                float hi = coherent ? coherent_load(src + 3 + k) : src[3 + k];
                b[k] = fminf(b[k], lo);
                b[3 + k] = fmaxf(b[3 + k], hi);
            }
        }
        for (int k = 0; k < 6; ++k)
            __hip_atomic_store((uint32_t*)(node_box + 6 * (size_t)node + k), __float_as_uint(b[k]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        node = parent_int[node];
    }
}

__global__ void k_depth(int n, const int* __restrict__ parent_int, const int* __restrict__ parent_leaf,
                        uint32_t* __restrict__ max_depth) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t dpt = 1;
    int node = parent_leaf[s];
    while (node >= 0) {
        ++dpt;
        node = parent_int[node];
    }
    atomicMax(max_depth, dpt);
}

__device__ __forceinline__ void padded(const float* b, float pad, float out[6]) {
    for (int k = 0; k < 3; ++k) {
        out[k] = b[k] - (pad + fabsf(b[k]) * 1e-6f);
        out[3 + k] = b[3 + k] + (pad + fabsf(b[3 + k]) * 1e-6f);
    }
}

__global__ void k_emit(int n, uint32_t leaf_size, float pad, const uint32_t* __restrict__ vals,
                       const float* __restrict__ tri_box, const uint32_t* __restrict__ left,
                       const uint32_t* __restrict__ right, const int* __restrict__ first,
                       const int* __restrict__ last, const float* __restrict__ node_box,
                       BvhNode* __restrict__ nodes) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    uint32_t ch[2] = {left[i], right[i]};
    float bx[2][6];
    int32_t ref[2];
    for (int c = 0; c < 2; ++c) {
        if (ch[c] & kLeafFlag) {
            uint32_t s = ch[c] & ~kLeafFlag;
            padded(tri_box + 6 * (size_t)vals[s], pad, bx[c]);
            ref[c] = (int32_t)(kLeafFlag | s);
        } else {
            uint32_t m = ch[c];
            padded(node_box + 6 * (size_t)m, pad, bx[c]);
            int cnt = last[m] - first[m] + 1;
            if ((uint32_t)cnt <= leaf_size)
                ref[c] = (int32_t)(kLeafFlag | ((uint32_t)(cnt - 1) << 24) | (uint32_t)first[m]);
            else
                ref[c] = (int32_t)m;
        }
    }
    BvhNode nd;
    nd.a = make_float4(bx[0][0], bx[0][1], bx[0][2], bx[0][3]);
    nd.b = make_float4(bx[0][4], bx[0][5], bx[1][0], bx[1][1]);
    nd.c = make_float4(bx[1][2], bx[1][3], bx[1][4], bx[1][5]);
    nd.d = make_int4(ref[0], ref[1], 0, 0);
    nodes[i] = nd;
}

__global__ void k_gather(uint32_t n, const uint32_t* __restrict__ vals, const float* __restrict__ v0,
                         const float* __restrict__ v1, const float* __restrict__ v2,
                         const uint16_t* __restrict__ mat, float4* __restrict__ tris, float4* __restrict__ nm) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t o = vals[s];
    V3 a{v0[3 * o], v0[3 * o + 1], v0[3 * o + 2]};
    V3 b{v1[3 * o], v1[3 * o + 1], v1[3 * o + 2]};
    V3 c{v2[3 * o], v2[3 * o + 1], v2[3 * o + 2]};
    V3 e1 = sub(b, a), e2 = sub(c, a);
    V3 nrm = normalized(cross(e1, e2));
    tris[3 * s + 0] = make_float4(a.x, a.y, a.z, e1.x);
    tris[3 * s + 1] = make_float4(e1.y, e1.z, e2.x, e2.y);
    tris[3 * s + 2] = make_float4(e2.z, __uint_as_float(o), 0.0f, 0.0f);
    nm[s] = make_float4(nrm.x, nrm.y, nrm.z, __uint_as_float((uint32_t)mat[o]));
}

struct Buf {
    Ctx* c;
    void* p = nullptr;
    size_t n = 0;
    explicit Buf(Ctx* cc) : c(cc) {}
    ~Buf() {
        if (p) dev_free(c, p, n);
    }
    ear_status alloc(size_t bytes) {
        n = bytes ? bytes : 16;
        return dev_alloc(c, &p, n);
    }
    template <class T>
    T* as() {
        return (T*)p;
    }
};

// Binned surface-area-heuristic build on the host (Wald, "On fast
// Construction of SAH-based Bounding Volume Hierarchies", RT 2007): each node
// bins its triangles' centroids into kSahBins slabs along each axis and takes
// the split plane with the least SAH cost sum(area x count); ranges of at most
// leaf_size triangles become leaves.  Output is the same 2-wide node array the
// device build emits (root at 0, padded child boxes, leaf references), plus
// the triangle order of the leaves.
constexpr int kSahBins = 32;

struct HostBox {
    float lo[3], hi[3];
    void reset() {
        for (int k = 0; k < 3; ++k) {
            lo[k] = INFINITY;
            hi[k] = -INFINITY;
        }
    }
    void grow(const HostBox& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    double area() const {
        double dx = std::max(0.0, (double)hi[0] - lo[0]), dy = std::max(0.0, (double)hi[1] - lo[1]),
               dz = std::max(0.0, (double)hi[2] - lo[2]);
        return dx * dy + dy * dz + dz * dx;
    }
};

struct SahBuild {
    const std::vector<HostBox>& tb;
    const std::vector<float>& cen;  // [n][3]
    std::vector<uint32_t>& order;
    std::vector<BvhNode>& nodes;
    uint32_t leaf_size;
    float pad;
    uint32_t max_depth = 0;

    void padded(const HostBox& b, float out[6]) const {
        for (int k = 0; k < 3; ++k) {
            out[k] = b.lo[k] - (pad + std::fabs(b.lo[k]) * 1e-6f);
            out[3 + k] = b.hi[k] + (pad + std::fabs(b.hi[k]) * 1e-6f);
        }
    }
    HostBox bounds(uint32_t b, uint32_t e) const {
        HostBox r;
        r.reset();
        for (uint32_t i = b; i < e; ++i) r.grow(tb[order[i]]);
        return r;
    }
    // splits [b, e) in place; returns the first index of the right part
    uint32_t split(uint32_t b, uint32_t e) {
        const uint32_t n = e - b;
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = b; i < e; ++i)
            for (int k = 0; k < 3; ++k) {
                clo[k] = std::min(clo[k], cen[3 * order[i] + k]);
                chi[k] = std::max(chi[k], cen[3 * order[i] + k]);
            }
        double best = INFINITY;
        int best_axis = -1, best_bin = 0;
        for (int k = 0; k < 3; ++k) {
            const float ext = chi[k] - clo[k];
            if (!(ext > 0.0f)) continue;
            HostBox bb[kSahBins];
            uint32_t bc[kSahBins] = {0};
            for (int j = 0; j < kSahBins; ++j) bb[j].reset();
            const float scale = (float)kSahBins / ext;
            for (uint32_t i = b; i < e; ++i) {
                int j = (int)((cen[3 * order[i] + k] - clo[k]) * scale);
                j = std::min(std::max(j, 0), kSahBins - 1);
                bb[j].grow(tb[order[i]]);
                ++bc[j];
            }
            // sweep: right-to-left areas, then left-to-right costs
            double ra[kSahBins];
            uint32_t rc[kSahBins];
            HostBox acc;
            acc.reset();
            uint32_t cnt = 0;
            for (int j = kSahBins - 1; j > 0; --j) {
                acc.grow(bb[j]);
                cnt += bc[j];
                ra[j] = acc.area();
                rc[j] = cnt;
            }
            acc.reset();
            cnt = 0;
            for (int j = 0; j < kSahBins - 1; ++j) {
                acc.grow(bb[j]);
                cnt += bc[j];
                if (cnt == 0 || rc[j + 1] == 0) continue;
                const double cost = acc.area() * cnt + ra[j + 1] * rc[j + 1];
                if (cost < best) {
                    best = cost;
                    best_axis = k;
                    best_bin = j;
                }
            }
        }
        if (best_axis < 0) return b + n / 2;  // all centroids coincide: split by count
        const int k = best_axis;
        const float scale = (float)kSahBins / (chi[k] - clo[k]);
        uint32_t* mid = std::partition(order.data() + b, order.data() + e, [&](uint32_t t) {
            int j = (int)((cen[3 * t + k] - clo[k]) * scale);
            j = std::min(std::max(j, 0), kSahBins - 1);
            return j <= best_bin;
        });
        uint32_t m = (uint32_t)(mid - order.data());
        if (m == b || m == e) m = b + n / 2;
        return m;
    }
    // reference of the subtree over [b, e) (a leaf or a new internal node)
    int32_t build(uint32_t b, uint32_t e, uint32_t depth, HostBox* box_out) {
        HostBox box = bounds(b, e);
        *box_out = box;
        max_depth = std::max(max_depth, depth);
        if (e - b <= leaf_size) return (int32_t)(kLeafFlag | ((e - b - 1) << 24) | b);
        // below depth 32 split by count, so the tree stays within the
        // traversal stack whatever the geometry
        const uint32_t m = depth < 32 ? split(b, e) : b + (e - b) / 2;
        const int32_t idx = (int32_t)nodes.size();
        nodes.emplace_back();
        HostBox lb, rb;
        const int32_t l = build(b, m, depth + 1, &lb);
        const int32_t r = build(m, e, depth + 1, &rb);
        float pl[6], pr[6];
        padded(lb, pl);
        padded(rb, pr);
        BvhNode& nd = nodes[idx];
        nd.a = make_float4(pl[0], pl[1], pl[2], pl[3]);
        nd.b = make_float4(pl[4], pl[5], pr[0], pr[1]);
        nd.c = make_float4(pr[2], pr[3], pr[4], pr[5]);
        nd.d = make_int4(l, r, 0, 0);
        return idx;
    }
};

}  // namespace

void free_scene(Ctx* c, DeviceScene* s) {
    if (s->nodes) dev_free(c, s->nodes, sizeof(BvhNode) * std::max<uint32_t>(1, s->n_nodes));
    if (s->nodes4) dev_free(c, s->nodes4, sizeof(Bvh4Node) * std::max<uint32_t>(1, s->n_nodes4_alloc));
    s->nodes4 = nullptr;
    if (s->tris) dev_free(c, s->tris, sizeof(float4) * 3 * std::max<uint32_t>(1, s->n_tris));
    if (s->nm) dev_free(c, s->nm, sizeof(float4) * std::max<uint32_t>(1, s->n_tris));
    if (s->mat_band) dev_free(c, s->mat_band, sizeof(float4) * 3 * std::max<uint32_t>(1, s->n_mat));
    s->nodes = nullptr;
    s->tris = s->nm = s->mat_band = nullptr;
}

// TRY: ear_internal.hpp

// A quantised plane as the traversal decodes it: org + q * 2^(eb - 127).
// The product is exact, so this rounds once, as the traversal's fused
// multiply-add does.
__host__ __device__ static inline float qplane(float org, uint32_t eb, uint32_t q) {
    const uint32_t bits = eb << 23;
    float scale;
    std::memcpy(&scale, &bits, 4);
    return org + (float)q * scale;
}

// Quantise the used children's boxes of one 4-wide node (Bvh4Node): per
// axis the origin is the lowest lo plane and the scale the smallest power of
// two for which every lo plane has a code at or below it and every hi plane a
// code at or above it within 0..255.  Returns false if a scale beyond 2^127
// would be needed (coordinates near the float range).  The same function runs
// in the host collapse and in the device build; whatever the first guess of
// the exponent, the codes are accepted only once decoding (qplane) contains
// every plane.
__host__ __device__ static inline bool quantize_node(const float lo[3][4], const float hi[3][4], const int32_t ref[4],
                                                     Bvh4Node& N) {
    uint32_t q[6][4] = {};  // qlo.x, qhi.x, qlo.y, qhi.y, qlo.z, qhi.z
    uint32_t org_bits[3], eb_all = 0;
    for (int k = 0; k < 3; ++k) {
        float mn = INFINITY, mx = -INFINITY;
        for (int i = 0; i < 4; ++i)
            if (ref[i] != kEmpty4) {
                mn = fminf(mn, lo[k][i]);
                mx = fmaxf(mx, hi[k][i]);
            }
        if (!(mn <= mx)) mn = mx = 0.0f;  // no used child (not produced by the collapse)
        const double ext = (double)mx - (double)mn;
        int e2 = ext > 0.0 ? (int)ceil(log2(ext / 255.0)) - 1 : -126;
        e2 = e2 < -126 ? -126 : e2;
        for (;; ++e2) {
            if (e2 > 127) return false;
            const uint32_t eb = (uint32_t)(e2 + 127);
            const double sc = ldexp(1.0, e2);
            bool ok = true;
            for (int i = 0; i < 4 && ok; ++i) {
                if (ref[i] == kEmpty4) continue;
                double fl = floor(((double)lo[k][i] - (double)mn) / sc);
                uint32_t ql = (uint32_t)fmin(255.0, fmax(0.0, fl));
                while (ql > 0 && qplane(mn, eb, ql) > lo[k][i]) --ql;
                double ce = ceil(((double)hi[k][i] - (double)mn) / sc);
                uint32_t qh = (uint32_t)fmin(255.0, fmax(0.0, ce));
                while (qh < 255 && qplane(mn, eb, qh) < hi[k][i]) ++qh;
                if (qplane(mn, eb, ql) > lo[k][i] || qplane(mn, eb, qh) < hi[k][i]) {
                    ok = false;
                    break;
                }
                q[2 * k][i] = ql;
                q[2 * k + 1][i] = qh;
            }
            if (ok) {
                eb_all |= eb << (8 * k);
                break;
            }
        }
        std::memcpy(&org_bits[k], &mn, 4);
    }
    uint32_t w[6];
    for (int r = 0; r < 6; ++r) w[r] = q[r][0] | q[r][1] << 8 | q[r][2] << 16 | q[r][3] << 24;
    N.org = make_uint4(org_bits[0], org_bits[1], org_bits[2], eb_all);
    N.qa = make_uint4(w[0], w[1], w[2], w[3]);
    N.qb = make_uint4(w[4], w[5], 0u, 0u);
    N.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
    return true;
}

// ---------------------------------------------------------------------------
// PLOC: parallel locally-ordered clustering (Meister and Bittner, "Parallel
// Locally-Ordered Clustering for Bounding Volume Hierarchy Construction",
// IEEE TVCG 2018), then a device top-down pass that collapses the binary tree
// into the traversal's 4-wide quantised nodes.
//   1. the triangles in Morton order are the first clusters (the LBVH steps
//      1-4: boxes, centroid bounds, 30-bit codes, radix sort)
//   2. per iteration every cluster finds, among the kPlocRadius clusters on
//      either side of it, the one whose merged box has the smallest surface
//      area (ties by the pair's indices: a strict order, so nearest-neighbour
//      chains end in mutual pairs); mutual pairs merge into a new node, which
//      takes the lower position; a prefix sum compacts the clusters
//   3. when one cluster is left it is the root; each node knows its triangle
//      count
//   4. level by level from the root: a 4-wide node takes a binary node's two
//      children and, while fewer than four, opens the child with the largest
//      surface area that holds more than leaf_size triangles; a child of at
//      most leaf_size triangles becomes a leaf.  Children's triangle ranges
//      partition their parent's, so every leaf is a contiguous run of the
//      final triangle order.  Nodes are numbered breadth first (a prefix sum
//      per level), their child boxes padded and quantised outward.
// ---------------------------------------------------------------------------
namespace {

constexpr int kPlocRadius = 16;
constexpr int kPlocBlock = 256;

__device__ __forceinline__ float box_area(const float* b) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dy * dz + dz * dx;
}
// in double: finite for any finite coordinates, so every pair is comparable
__device__ __forceinline__ double merged_area(const float* a, const float* b) {
    const double dx = (double)fmaxf(a[3], b[3]) - (double)fminf(a[0], b[0]);
    const double dy = (double)fmaxf(a[4], b[4]) - (double)fminf(a[1], b[1]);
    const double dz = (double)fmaxf(a[5], b[5]) - (double)fminf(a[2], b[2]);
    return dx * dy + dy * dz + dz * dx;
}

// cluster i = leaf i of the Morton order (ref kLeafFlag | i), its triangle's box
__global__ void k_ploc_init(uint32_t n, const uint32_t* __restrict__ vals, const float* __restrict__ tri_box,
                            uint32_t* __restrict__ cref, float* __restrict__ cbox, uint32_t* __restrict__ ccnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cref[i] = kLeafFlag | i;
    ccnt[i] = 1;
    const float* b = tri_box + 6 * (size_t)vals[i];
    for (int k = 0; k < 6; ++k) cbox[6 * (size_t)i + k] = b[k];
}

// nn[i] = the nearest cluster within kPlocRadius positions (smallest merged
// surface area; ties to the smaller pair (min(i, j), max(i, j)))
__global__ void __launch_bounds__(kPlocBlock) k_ploc_nn(uint32_t C, const float* __restrict__ cbox,
                                                        uint32_t* __restrict__ nn) {
    constexpr int kSpan = kPlocBlock + 2 * kPlocRadius;
    __shared__ float sb[6][kSpan];
    const int i0 = (int)(blockIdx.x * kPlocBlock) - kPlocRadius;
    for (int t = threadIdx.x; t < kSpan; t += kPlocBlock) {
        const int j = i0 + t;
        for (int k = 0; k < 6; ++k) sb[k][t] = (j >= 0 && j < (int)C) ? cbox[6 * (size_t)j + k] : 0.0f;
    }
    __syncthreads();
    const int i = (int)(blockIdx.x * kPlocBlock + threadIdx.x);
    if (i >= (int)C) return;
    const int li = threadIdx.x + kPlocRadius;
    float bi[6];
    for (int k = 0; k < 6; ++k) bi[k] = sb[k][li];
    double best = INFINITY;
    int bj = -1;
    for (int o = -kPlocRadius; o <= kPlocRadius; ++o) {
        const int j = i + o;
        if (o == 0 || j < 0 || j >= (int)C) continue;
        float bj6[6];
        for (int k = 0; k < 6; ++k) bj6[k] = sb[k][li + o];
        const double d = merged_area(bi, bj6);
        // pairs are ordered by (d, min(i, j), max(i, j)); for a fixed i and
        // equal d the smaller pair is the one with the smaller j, which is
        // the first one met as j runs upwards
        if (d < best) {
            best = d;
            bj = j;
        }
    }
    nn[i] = (uint32_t)(bj < 0 ? i : bj);
}

// Mutual nearest neighbours merge: the lower position takes the new node
// (left = the lower cluster), the higher one leaves.  Writes the surviving
// clusters to t* at their old positions, keep[i], its exclusive prefix within
// the block in pre[i], and the block's survivors in bsum[block].
__global__ void __launch_bounds__(kPlocBlock) k_ploc_merge(uint32_t C, const uint32_t* __restrict__ nn,
                                                           const uint32_t* __restrict__ cref,
                                                           const float* __restrict__ cbox,
                                                           const uint32_t* __restrict__ ccnt,
                                                           uint32_t* __restrict__ node_ctr, int2* __restrict__ nd_child,
                                                           float* __restrict__ nd_box, uint32_t* __restrict__ nd_cnt,
                                                           uint32_t* __restrict__ tref, float* __restrict__ tbox,
                                                           uint32_t* __restrict__ tcnt, uint32_t* __restrict__ keep,
                                                           uint32_t* __restrict__ pre, uint32_t* __restrict__ bsum) {
    __shared__ uint32_t wsum[kPlocBlock / 64];
    const uint32_t i = blockIdx.x * kPlocBlock + threadIdx.x;
    uint32_t kp = 0;
    if (i < C) {
        const uint32_t j = nn[i];
        const bool mutual = j != i && nn[j] == i;
        if (mutual && i > j) {
            kp = 0;
        } else if (mutual) {
            const uint32_t id = atomicAdd(node_ctr, 1u);
            const float* a = cbox + 6 * (size_t)i;
            const float* b = cbox + 6 * (size_t)j;
            float u[6];
            for (int k = 0; k < 3; ++k) {
                u[k] = fminf(a[k], b[k]);
                u[3 + k] = fmaxf(a[3 + k], b[3 + k]);
            }
            nd_child[id] = make_int2((int32_t)cref[i], (int32_t)cref[j]);
            const uint32_t cn = ccnt[i] + ccnt[j];
            nd_cnt[id] = cn;
            for (int k = 0; k < 6; ++k) {
                nd_box[6 * (size_t)id + k] = u[k];
                tbox[6 * (size_t)i + k] = u[k];
            }
            tref[i] = id;
            tcnt[i] = cn;
            kp = 1;
        } else {
            tref[i] = cref[i];
            tcnt[i] = ccnt[i];
            for (int k = 0; k < 6; ++k) tbox[6 * (size_t)i + k] = cbox[6 * (size_t)i + k];
            kp = 1;
        }
        keep[i] = kp;
    }
    // exclusive prefix of kp within the block
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(kp != 0);
    const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += wsum[w];
    if (i < C) pre[i] = base + below;
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kPlocBlock / 64; ++w) t += wsum[w];
        bsum[blockIdx.x] = t;
    }
}

// survivors to their compacted positions (bsum now exclusive over blocks);
// *count = the number of clusters left
__global__ void __launch_bounds__(kPlocBlock) k_ploc_compact(uint32_t C, const uint32_t* __restrict__ keep,
                                                             const uint32_t* __restrict__ pre,
                                                             const uint32_t* __restrict__ boff,
                                                             const uint32_t* __restrict__ tref,
                                                             const float* __restrict__ tbox,
                                                             const uint32_t* __restrict__ tcnt,
                                                             uint32_t* __restrict__ cref, float* __restrict__ cbox,
                                                             uint32_t* __restrict__ ccnt, uint32_t* __restrict__ count) {
    const uint32_t i = blockIdx.x * kPlocBlock + threadIdx.x;
    if (i >= C) return;
    const uint32_t d = boff[blockIdx.x] + pre[i];
    if (keep[i]) {
        cref[d] = tref[i];
        ccnt[d] = tcnt[i];
        for (int k = 0; k < 6; ++k) cbox[6 * (size_t)d + k] = tbox[6 * (size_t)i + k];
    }
    if (i == C - 1) *count = d + keep[i];
}

// The up to four children of binary node x as the collapse takes them (see
// above): refs (binary node id, or kLeafFlag | Morton position).
__device__ int gather4(uint32_t x, const int2* __restrict__ nd_child, const float* __restrict__ nd_box,
                       const uint32_t* __restrict__ nd_cnt, uint32_t leaf_size, uint32_t kid[4]) {
    const int2 c = nd_child[x];
    kid[0] = (uint32_t)c.x;
    kid[1] = (uint32_t)c.y;
    int nk = 2;
    while (nk < 4) {
        int best = -1;
        float ba = -1.0f;
        for (int i = 0; i < nk; ++i) {
            if ((kid[i] & kLeafFlag) || nd_cnt[kid[i]] <= leaf_size) continue;
            const float a = box_area(nd_box + 6 * (size_t)kid[i]);
            if (a > ba) {
                ba = a;
                best = i;
            }
        }
        if (best < 0) break;
        const int2 cc = nd_child[kid[best]];
        kid[best] = (uint32_t)cc.x;
        kid[nk++] = (uint32_t)cc.y;
    }
    return nk;
}

__device__ __forceinline__ bool is_inner4(uint32_t ref, const uint32_t* __restrict__ nd_cnt, uint32_t leaf_size) {
    return !(ref & kLeafFlag) && nd_cnt[ref] > leaf_size;
}

// kcount[f] = the 4-wide children of frontier node f that are inner nodes
__global__ void k_c4_count(uint32_t F, const uint32_t* __restrict__ fnode, const int2* __restrict__ nd_child,
                           const float* __restrict__ nd_box, const uint32_t* __restrict__ nd_cnt, uint32_t leaf_size,
                           uint32_t* __restrict__ kcount) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    uint32_t kid[4];
    const int nk = gather4(fnode[f], nd_child, nd_box, nd_cnt, leaf_size, kid);
    uint32_t m = 0;
    for (int i = 0; i < nk; ++i) m += is_inner4(kid[i], nd_cnt, leaf_size) ? 1u : 0u;
    kcount[f] = m;
}

// Frontier node f (binary node fnode[f], 4-wide slot fslot[f], triangle
// range from foff[f]) becomes its 4-wide node: children in gather4 order take
// consecutive triangle ranges; leaves list their triangles (original
// indices, depth first) in `order`; inner children go to the next frontier at
// koff[f] + rank (koff: the exclusive prefix of k_c4_count's counts), slot
// next_base + koff[f] + rank.  *next = the next frontier's size; *err set if
// a node cannot be quantised.
__global__ void k_c4_emit(uint32_t F, const uint32_t* __restrict__ fnode, const uint32_t* __restrict__ fslot,
                          const uint32_t* __restrict__ foff, const uint32_t* __restrict__ koff, uint32_t next_base,
                          const int2* __restrict__ nd_child,
                          const float* __restrict__ nd_box, const uint32_t* __restrict__ nd_cnt,
                          const float* __restrict__ tri_box, const uint32_t* __restrict__ vals, uint32_t leaf_size,
                          float pad, Bvh4Node* __restrict__ nodes4, uint32_t* __restrict__ order,
                          uint32_t* __restrict__ nnode, uint32_t* __restrict__ nslot, uint32_t* __restrict__ noff,
                          uint32_t* __restrict__ next, uint32_t* __restrict__ err) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    uint32_t kid[4];
    const int nk = gather4(fnode[f], nd_child, nd_box, nd_cnt, leaf_size, kid);
    float lo[3][4], hi[3][4];
    int32_t ref[4];
    uint32_t off = foff[f], rank = 0;
    for (int i = 0; i < 4; ++i) {
        if (i >= nk) {
            for (int k = 0; k < 3; ++k) lo[k][i] = hi[k][i] = 1e30f;
            ref[i] = kEmpty4;
            continue;
        }
        const uint32_t r = kid[i];
        const bool leaf = (r & kLeafFlag) != 0;
        const float* b = leaf ? tri_box + 6 * (size_t)vals[r & ~kLeafFlag] : nd_box + 6 * (size_t)r;
        float pb[6];
        padded(b, pad, pb);
        for (int k = 0; k < 3; ++k) {
            lo[k][i] = pb[k];
            hi[k][i] = pb[3 + k];
        }
        const uint32_t cnt = leaf ? 1u : nd_cnt[r];
        if (!leaf && cnt > leaf_size) {
            const uint32_t q = koff[f] + rank++;
            nnode[q] = r;
            nslot[q] = next_base + q;
            noff[q] = off;
            ref[i] = (int32_t)(next_base + q);
        } else {
            // a leaf: its triangles in depth-first order (at most leaf_size <= 16)
            uint32_t stk[16];
            int sp = 0;
            stk[sp++] = r;
            uint32_t t = off;
            while (sp > 0) {
                const uint32_t u = stk[--sp];
                if (u & kLeafFlag) {
                    order[t++] = vals[u & ~kLeafFlag];
                } else {
                    const int2 cc = nd_child[u];
                    stk[sp++] = (uint32_t)cc.y;
                    stk[sp++] = (uint32_t)cc.x;
                }
            }
            ref[i] = (int32_t)(kLeafFlag | ((cnt - 1) << 24) | off);
        }
        off += cnt;
    }
    if (!quantize_node(lo, hi, ref, nodes4[fslot[f]])) atomicOr(err, 1u);
    if (f == F - 1) *next = koff[f] + rank;
}

}  // namespace

// Collapse the 2-wide tree into a 4-wide one: each 4-wide node takes a 2-wide
// node's two children and, while fewer than four, replaces the internal child
// with the largest box surface area by its two children.  The (already
// padded) boxes are quantised outward (quantize_node), leaf references carry
// over unchanged; nodes are laid out in breadth-first order so siblings sit
// next to each other.
static ear_status collapse_to_bvh4(Ctx* c, DeviceScene* s) {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<BvhNode> n2(s->n_nodes);
    EAR_HIP_TRY(c, hipMemcpy(n2.data(), s->nodes, sizeof(BvhNode) * s->n_nodes, hipMemcpyDeviceToHost));
    struct Child {
        float b[6];
        int32_t ref;
    };
    auto kids = [&](int32_t node, Child out[2]) {
        const BvhNode& N = n2[node];
        float a[12] = {N.a.x, N.a.y, N.a.z, N.a.w, N.b.x, N.b.y, N.b.z, N.b.w, N.c.x, N.c.y, N.c.z, N.c.w};
        for (int k = 0; k < 6; ++k) {
            out[0].b[k] = a[k];
            out[1].b[k] = a[6 + k];
        }
        out[0].ref = N.d.x;
        out[1].ref = N.d.y;
    };
    auto area = [](const Child& ch) {
        float dx = ch.b[3] - ch.b[0], dy = ch.b[4] - ch.b[1], dz = ch.b[5] - ch.b[2];
        return dx * dy + dy * dz + dz * dx;
    };
    std::vector<Bvh4Node> out;
    out.reserve(s->n_nodes / 2 + 2);
    struct Work {
        int32_t n2, slot, depth;
    };
    std::vector<Work> queue;
    queue.push_back({0, 0, 1});
    out.emplace_back();
    uint32_t depth = 1;
    for (size_t qi = 0; qi < queue.size(); ++qi) {
        Work w = queue[qi];
        depth = std::max<uint32_t>(depth, (uint32_t)w.depth);
        Child ch[4];
        int nch = 2;
        kids(w.n2, ch);
        while (nch < 4) {
            int best = -1;
            float ba = -1.0f;
            for (int i = 0; i < nch; ++i)
                if (ch[i].ref >= 0 && area(ch[i]) > ba) {
                    ba = area(ch[i]);
                    best = i;
                }
            if (best < 0) break;
            Child sub[2];
            kids(ch[best].ref, sub);
            ch[best] = sub[0];
            ch[nch++] = sub[1];
        }
        float lo[3][4], hi[3][4];
        int32_t ref[4];
        for (int i = 0; i < 4; ++i) {
            // the 2-wide tree's placeholder child (a box at 1e30) is no child
            if (i < nch && ch[i].b[0] < 1e29f) {
                for (int k = 0; k < 3; ++k) {
                    lo[k][i] = ch[i].b[k];
                    hi[k][i] = ch[i].b[3 + k];
                }
                if (ch[i].ref >= 0) {
                    int32_t slot = (int32_t)out.size();
                    out.emplace_back();
                    queue.push_back({ch[i].ref, slot, w.depth + 1});
                    ref[i] = slot;
                } else {
                    ref[i] = ch[i].ref;
                }
            } else {
                for (int k = 0; k < 3; ++k) lo[k][i] = hi[k][i] = 1e30f;
                ref[i] = kEmpty4;
            }
        }
        if (!quantize_node(lo, hi, ref, out[w.slot])) {
            c->last_error = "scene coordinates too large for the quantised BVH nodes";
            return EAR_ERR_INVALID_GEOMETRY;
        }
    }
    if (3 * depth + 1 >= (uint32_t)kStackDepth) {
        c->last_error = "4-wide BVH deeper than the traversal stack";
        return EAR_ERR_CAPACITY;
    }
    ear_status st = dev_alloc(c, (void**)&s->nodes4, sizeof(Bvh4Node) * out.size());
    if (st != EAR_OK) return st;
    s->n_nodes4 = s->n_nodes4_alloc = (uint32_t)out.size();
    EAR_HIP_TRY(c, hipMemcpy(s->nodes4, out.data(), sizeof(Bvh4Node) * out.size(), hipMemcpyHostToDevice));
    s->depth4 = depth;
    s->width = 4;
    s->collapse_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return EAR_OK;
}

// The SAH build of the 2-wide tree on the host, then the same device gather
// of the triangles into leaf order as the LBVH path.
static ear_status build_sah(Ctx* c, const float* v0, const float* v1, const float* v2, const uint16_t* mat_id,
                            uint32_t n, uint32_t leaf_size, DeviceScene* s) {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<HostBox> tb(n);
    std::vector<float> cen(3 * (size_t)n);
    HostBox scene;
    scene.reset();
    for (uint32_t i = 0; i < n; ++i) {
        HostBox& b = tb[i];
        for (int k = 0; k < 3; ++k) {
            const float a = v0[3 * (size_t)i + k], bb = v1[3 * (size_t)i + k], cc = v2[3 * (size_t)i + k];
            b.lo[k] = std::min(a, std::min(bb, cc));
            b.hi[k] = std::max(a, std::max(bb, cc));
            cen[3 * (size_t)i + k] = 0.5f * (b.lo[k] + b.hi[k]);
        }
        scene.grow(b);
    }
    float ext = 0.0f;
    for (int k = 0; k < 3; ++k) {
        s->bounds[k] = scene.lo[k];
        s->bounds[3 + k] = scene.hi[k];
        ext = std::max(ext, scene.hi[k] - scene.lo[k]);
    }
    const float pad = 1e-5f * ext + 1e-6f;
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::vector<BvhNode> nodes;
    nodes.reserve(n);
    SahBuild sb{tb, cen, order, nodes, leaf_size, pad};
    HostBox root;
    const int32_t r = sb.build(0, n, 1, &root);
    if (r < 0) {
        // the whole scene is one leaf: a root with that leaf and an empty child
        float pb[6];
        sb.padded(root, pb);
        BvhNode nd;
        nd.a = make_float4(pb[0], pb[1], pb[2], pb[3]);
        nd.b = make_float4(pb[4], pb[5], 1e30f, 1e30f);
        nd.c = make_float4(1e30f, 1e30f, 1e30f, 1e30f);
        nd.d = make_int4(r, kEmptyChild, 0, 0);
        nodes.push_back(nd);
        sb.max_depth = 1;
    }
    s->n_nodes = (uint32_t)nodes.size();
    s->max_depth = sb.max_depth;
    if (s->max_depth >= (uint32_t)kStackDepth) {
        c->last_error = "BVH deeper than the traversal stack";
        return EAR_ERR_CAPACITY;
    }
    hipStream_t st = c->stream;
    TRY(dev_alloc(c, (void**)&s->nodes, sizeof(BvhNode) * nodes.size()));
    TRY(dev_alloc(c, (void**)&s->tris, sizeof(float4) * 3 * (size_t)n));
    TRY(dev_alloc(c, (void**)&s->nm, sizeof(float4) * (size_t)n));
    Buf dv0(c), dv1(c), dv2(c), dmat(c), vals(c);
    const size_t n3 = 3 * (size_t)n;
    TRY(dv0.alloc(n3 * 4));
    TRY(dv1.alloc(n3 * 4));
    TRY(dv2.alloc(n3 * 4));
    TRY(dmat.alloc((size_t)n * 2));
    TRY(vals.alloc((size_t)n * 4));
    EAR_HIP_TRY(c, hipMemcpyAsync(s->nodes, nodes.data(), sizeof(BvhNode) * nodes.size(), hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(dv0.p, v0, n3 * 4, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(dv1.p, v1, n3 * 4, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(dv2.p, v2, n3 * 4, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(dmat.p, mat_id, (size_t)n * 2, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(vals.p, order.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
    k_gather<<<(n + 255) / 256, 256, 0, st>>>(n, vals.as<uint32_t>(), dv0.as<float>(), dv1.as<float>(),
                                              dv2.as<float>(), dmat.as<uint16_t>(), s->tris, s->nm);
    EAR_HIP_TRY(c, hipGetLastError());
    EAR_HIP_TRY(c, hipStreamSynchronize(st));
    s->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return EAR_OK;
}

// ---------------------------------------------------------------------------
// Device binned-SAH build: the host build's split rule (SahBuild::split, Wald
// 2007: kSahBins centroid bins per axis, the plane of least
// area(left) count(left) + area(right) count(right) in double, ties to the
// first axis and bin) evaluated on the device, level by level, one workgroup
// per node: centroid and box bounds, binning with LDS atomics (floats as
// order-preserving integers), the split, and a stable partition of the node's
// triangle indices into the next level's buffer.  Binning depends on the set
// of a node's triangles, not their order, so the nodes are the host build's
// (a split by count, taken only when every centroid coincides or below depth
// 32, depends on the order and may differ).  Ranges are split down to single
// triangles; the 4-wide collapse then makes every subtree of at most
// leaf_size triangles one leaf, as the host build's leaves.  Output: the
// binary tree in the PLOC arrays (nd_child, nd_box, nd_cnt; a leaf is
// kLeafFlag | triangle index).
// ---------------------------------------------------------------------------
namespace {

struct SahTask {
    uint32_t b, e, node, depth;
};

__device__ __forceinline__ uint32_t f2o(float x) {
    const uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
constexpr uint32_t kOrdInf = 0xFF800000u;     // f2o(+inf): the start of a minimum
constexpr uint32_t kOrdNegInf = 0x007FFFFFu;  // f2o(-inf): the start of a maximum
__device__ __forceinline__ float o2f(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}
// HostBox::area: max(0, hi - lo) per axis in double
__device__ __forceinline__ double sah_area(const float lo[3], const float hi[3]) {
    const double dx = fmax(0.0, (double)hi[0] - lo[0]), dy = fmax(0.0, (double)hi[1] - lo[1]),
                 dz = fmax(0.0, (double)hi[2] - lo[2]);
    return dx * dy + dy * dz + dz * dx;
}

template <int BLOCK>
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* sh) {
    constexpr int W = BLOCK / 64;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    __syncthreads();
    if ((threadIdx.x & 63u) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int w = 0; w < W; ++w) t += sh[w];
    return t;
}

// One workgroup per task of this level: splits its range [b, e) of `in` into
// `out` (same positions), records the node's box, count and children, and
// appends the children of two or more triangles to the next level's tasks.
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_sah_level(const SahTask* __restrict__ tasks, const uint32_t* __restrict__ in,
                                                     uint32_t* __restrict__ out, const float* __restrict__ cen,
                                                     const float* __restrict__ tri_box, SahTask* __restrict__ next,
                                                     uint32_t* __restrict__ n_next, uint32_t* __restrict__ node_ctr,
                                                     int2* __restrict__ nd_child, float* __restrict__ nd_box,
                                                     uint32_t* __restrict__ nd_cnt) {
    constexpr int NB = 32;  // kSahBins
    __shared__ uint32_t bnd[12];          // centroid lo xyz, hi xyz, box lo xyz, hi xyz (ordered ints)
    __shared__ uint32_t bins[3][NB][7];   // count, box lo xyz, hi xyz (ordered ints)
    __shared__ uint32_t sh[BLOCK / 64 + 1];
    __shared__ int s_axis, s_bin;
    __shared__ uint32_t s_base[2];
    __shared__ double s_ra[NB];  // the sweep's right-side areas and counts (thread 0)
    __shared__ uint32_t s_rc[NB];
    const SahTask T = tasks[blockIdx.x];
    const uint32_t b = T.b, e = T.e, n = e - b;
    if (threadIdx.x < 12) bnd[threadIdx.x] = (threadIdx.x % 6) < 3 ? kOrdInf : kOrdNegInf;
    for (uint32_t j = threadIdx.x; j < 3 * NB * 7; j += BLOCK) {
        const uint32_t f = j % 7;
        (&bins[0][0][0])[j] = f == 0 ? 0u : (f < 4 ? kOrdInf : kOrdNegInf);
    }
    __syncthreads();
    // bounds of the centroids and of the triangle boxes
    {
        uint32_t r[12];
        for (int k = 0; k < 12; ++k) r[k] = (k % 6) < 3 ? kOrdInf : kOrdNegInf;
        for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK) {
            const uint32_t t = in[i];
            for (int k = 0; k < 3; ++k) {
                const uint32_t c = f2o(cen[3 * (size_t)t + k]);
                r[k] = min(r[k], c);
                r[3 + k] = max(r[3 + k], c);
                r[6 + k] = min(r[6 + k], f2o(tri_box[6 * (size_t)t + k]));
                r[9 + k] = max(r[9 + k], f2o(tri_box[6 * (size_t)t + 3 + k]));
            }
        }
        for (int k = 0; k < 12; ++k) {
            if ((k % 6) < 3)
                atomicMin(&bnd[k], r[k]);
            else
                atomicMax(&bnd[k], r[k]);
        }
    }
    __syncthreads();
    float clo[3], chi[3], scale[3];
    for (int k = 0; k < 3; ++k) {
        clo[k] = o2f(bnd[k]);
        chi[k] = o2f(bnd[3 + k]);
        const float ext = chi[k] - clo[k];
        scale[k] = ext > 0.0f ? (float)NB / ext : 0.0f;
    }
    if (threadIdx.x == 0) {
        nd_cnt[T.node] = n;
        for (int k = 0; k < 3; ++k) {
            nd_box[6 * (size_t)T.node + k] = o2f(bnd[6 + k]);
            nd_box[6 * (size_t)T.node + 3 + k] = o2f(bnd[9 + k]);
        }
    }
    const bool by_count = T.depth >= 32;
    if (!by_count) {
        // binning
        for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK) {
            const uint32_t t = in[i];
            uint32_t tb[6];
            for (int k = 0; k < 6; ++k) tb[k] = f2o(tri_box[6 * (size_t)t + k]);
            for (int k = 0; k < 3; ++k) {
                if (!(scale[k] > 0.0f)) continue;
                int j = (int)((cen[3 * (size_t)t + k] - clo[k]) * scale[k]);
                j = min(max(j, 0), NB - 1);
                atomicAdd(&bins[k][j][0], 1u);
                for (int a = 0; a < 3; ++a) {
                    atomicMin(&bins[k][j][1 + a], tb[a]);
                    atomicMax(&bins[k][j][4 + a], tb[3 + a]);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // SahBuild::split's sweep, in its order
            double best = INFINITY;
            int best_axis = -1, best_bin = 0;
            for (int k = 0; k < 3; ++k) {
                if (!(scale[k] > 0.0f)) continue;
                double* ra = s_ra;
                uint32_t* rc = s_rc;
                float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
                uint32_t cnt = 0;
                for (int j = NB - 1; j > 0; --j) {
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = fminf(lo[a], o2f(bins[k][j][1 + a]));
                        hi[a] = fmaxf(hi[a], o2f(bins[k][j][4 + a]));
                    }
                    cnt += bins[k][j][0];
                    ra[j] = sah_area(lo, hi);
                    rc[j] = cnt;
                }
                for (int a = 0; a < 3; ++a) {
                    lo[a] = INFINITY;
                    hi[a] = -INFINITY;
                }
                cnt = 0;
                for (int j = 0; j < NB - 1; ++j) {
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = fminf(lo[a], o2f(bins[k][j][1 + a]));
                        hi[a] = fmaxf(hi[a], o2f(bins[k][j][4 + a]));
                    }
                    cnt += bins[k][j][0];
                    if (cnt == 0 || rc[j + 1] == 0) continue;
                    const double cost = sah_area(lo, hi) * cnt + ra[j + 1] * rc[j + 1];
                    if (cost < best) {
                        best = cost;
                        best_axis = k;
                        best_bin = j;
                    }
                }
            }
            s_axis = best_axis;
            s_bin = best_bin;
        }
        __syncthreads();
    } else if (threadIdx.x == 0) {
        s_axis = -1;
    }
    __syncthreads();
    const int axis = s_axis, sbin = s_bin;
    // left: the triangles of bins <= sbin (or the first half of the range)
    auto goes_left = [&](uint32_t i, uint32_t t) -> bool {
        if (axis < 0) return i - b < n / 2;
        int j = (int)((cen[3 * (size_t)t + axis] - clo[axis]) * scale[axis]);
        j = min(max(j, 0), NB - 1);
        return j <= sbin;
    };
    uint32_t nl_mine = 0;
    for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK) nl_mine += goes_left(i, in[i]) ? 1u : 0u;
    uint32_t nl = block_sum<BLOCK>(nl_mine, sh);
    bool count_split = axis < 0 || nl == 0 || nl == n;
    if (count_split) nl = n / 2;
    // stable partition, BLOCK positions at a time
    if (threadIdx.x == 0) {
        s_base[0] = 0;
        s_base[1] = 0;
    }
    __syncthreads();
    for (uint32_t i0 = b; i0 < e; i0 += BLOCK) {
        const uint32_t i = i0 + threadIdx.x;
        const bool valid = i < e;
        const uint32_t t = valid ? in[i] : 0u;
        const bool left = valid && (count_split ? (i - b < n / 2) : goes_left(i, t));
        const uint64_t ml = __ballot(left), mr = __ballot(valid && !left);
        const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
        const uint64_t below = (1ull << lane) - 1ull;
        __syncthreads();
        if (lane == 0) sh[wave] = (uint32_t)__popcll(ml) | ((uint32_t)__popcll(mr) << 16);
        __syncthreads();
        uint32_t pl = 0, pr = 0;
        for (uint32_t w = 0; w < wave; ++w) {
            pl += sh[w] & 0xFFFFu;
            pr += sh[w] >> 16;
        }
        if (left) out[b + s_base[0] + pl + (uint32_t)__popcll(ml & below)] = t;
        else if (valid) out[b + nl + s_base[1] + pr + (uint32_t)__popcll(mr & below)] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tl = 0, tr = 0;
            for (uint32_t w = 0; w < BLOCK / 64; ++w) {
                tl += sh[w] & 0xFFFFu;
                tr += sh[w] >> 16;
            }
            s_base[0] += tl;
            s_base[1] += tr;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t ref[2];
        const uint32_t rb[2] = {b, b + nl}, re[2] = {b + nl, e};
        for (int c = 0; c < 2; ++c) {
            if (re[c] - rb[c] == 1) {
                ref[c] = (int32_t)(kLeafFlag | out[rb[c]]);
            } else {
                const uint32_t id = atomicAdd(node_ctr, 1u);
                ref[c] = (int32_t)id;
                next[atomicAdd(n_next, 1u)] = SahTask{rb[c], re[c], id, T.depth + 1};
            }
        }
        nd_child[T.node] = make_int2(ref[0], ref[1]);
    }
}

}  // namespace

// The device top-down pass shared by the PLOC and device-SAH builds: the
// binary tree (nd_child / nd_box / nd_cnt, leaves kLeafFlag | p with vals[p]
// the triangle) into s->nodes4, breadth first, and the triangles' leaf order
// into `order`.  hctr: 8 bytes of pinned host memory.
static ear_status collapse_device(Ctx* c, DeviceScene* s, uint32_t n, uint32_t root, const int2* nd_child,
                                  const float* nd_box, const uint32_t* nd_cnt, const float* tri_box,
                                  const uint32_t* vals, uint32_t leaf_size, float pad, uint32_t* order,
                                  uint32_t* hctr, uint32_t* levels_out) {
    hipStream_t st = c->stream;
    const uint32_t n_int = n - 1;
    Buf fa(c), fb(c), fc(c), ga(c), gb(c), gc(c), kcnt(c), flags(c);
    TRY(dev_alloc(c, (void**)&s->nodes4, sizeof(Bvh4Node) * (size_t)n_int));
    s->n_nodes4_alloc = n_int;
    TRY(fa.alloc((size_t)n * 4));
    TRY(fb.alloc((size_t)n * 4));
    TRY(fc.alloc((size_t)n * 4));
    TRY(ga.alloc((size_t)n * 4));
    TRY(gb.alloc((size_t)n * 4));
    TRY(gc.alloc((size_t)n * 4));
    TRY(kcnt.alloc((size_t)n * 4));
    TRY(flags.alloc(8));  // [0] next frontier size, [1] quantisation error
    EAR_HIP_TRY(c, hipMemsetAsync(flags.p, 0, 8, st));
    const uint32_t first[3] = {root, 0u, 0u};
    EAR_HIP_TRY(c, hipMemcpyAsync(fa.p, &first[0], 4, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(fb.p, &first[1], 4, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(fc.p, &first[2], 4, hipMemcpyHostToDevice, st));
    uint32_t *fn = fa.as<uint32_t>(), *fs = fb.as<uint32_t>(), *fo = fc.as<uint32_t>();
    uint32_t *gn = ga.as<uint32_t>(), *gs = gb.as<uint32_t>(), *go = gc.as<uint32_t>();
    uint32_t F = 1, n4 = 1, levels = 0;
    while (F > 0) {
        const uint32_t g = (F + 255) / 256;
        k_c4_count<<<g, 256, 0, st>>>(F, fn, nd_child, nd_box, nd_cnt, leaf_size, kcnt.as<uint32_t>());
        k_scan_u32<<<1, 1024, 0, st>>>(F, kcnt.as<uint32_t>());
        k_c4_emit<<<g, 256, 0, st>>>(F, fn, fs, fo, kcnt.as<uint32_t>(), n4, nd_child, nd_box, nd_cnt, tri_box, vals,
                                      leaf_size, pad, s->nodes4, order, gn, gs, go, flags.as<uint32_t>(),
                                      flags.as<uint32_t>() + 1);
        EAR_HIP_TRY(c, hipGetLastError());
        EAR_HIP_TRY(c, hipMemcpyAsync(hctr, flags.p, 8, hipMemcpyDeviceToHost, st));
        EAR_HIP_TRY(c, hipStreamSynchronize(st));
        if (hctr[1]) {
            c->last_error = "scene coordinates too large for the quantised BVH nodes";
            return EAR_ERR_INVALID_GEOMETRY;
        }
        ++levels;
        F = hctr[0];
        if (n4 + F > n_int) {
            c->last_error = "the 4-wide collapse produced more nodes than the binary tree";
            return EAR_ERR_INTERNAL;
        }
        n4 += F;
        std::swap(fn, gn);
        std::swap(fs, gs);
        std::swap(fo, go);
    }
    s->n_nodes4 = n4;
    *levels_out = levels;
    return EAR_OK;
}

// The device builds of the 4-wide tree (EAR_BVH_PLOC or EAR_BVH_SAH_GPU, see
// above) and the triangles in leaf order.  build_ms: device time from the
// uploaded vertices to the gathered triangles; collapse_ms: the top-down
// pass's share.
static ear_status build_device_tree(Ctx* c, const float* v0, const float* v1, const float* v2,
                                    const uint16_t* mat_id, uint32_t n, uint32_t leaf_size, uint32_t builder,
                                    DeviceScene* s) {
    hipStream_t st = c->stream;
    hipEvent_t e0, e1, e2;
    EAR_HIP_TRY(c, hipEventCreate(&e0));
    EAR_HIP_TRY(c, hipEventCreate(&e1));
    EAR_HIP_TRY(c, hipEventCreate(&e2));
    struct EvGuard {
        hipEvent_t a, b, c;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            (void)hipEventDestroy(c);
        }
    } eg{e0, e1, e2};

    Buf dv0(c), dv1(c), dv2(c), dmat(c), box(c), cen(c), part(c), bnd(c), keys(c), vals(c), keys2(c), vals2(c),
        counts(c), cref(c), cbox(c), ccnt(c), tref(c), tbox(c), tcnt(c), nn(c), keep(c), pre(c), bsum(c),
        nd_child(c), nd_box(c), nd_cnt(c), ctr(c), order(c);
    const size_t n3 = 3 * (size_t)n;
    TRY(dv0.alloc(n3 * 4));
    TRY(dv1.alloc(n3 * 4));
    TRY(dv2.alloc(n3 * 4));
    TRY(dmat.alloc((size_t)n * 2));
    EAR_HIP_TRY(c, hipMemcpyAsync(dv0.p, v0, n3 * 4, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(dv1.p, v1, n3 * 4, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(dv2.p, v2, n3 * 4, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(dmat.p, mat_id, (size_t)n * 2, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipEventRecord(e0, st));

    // boxes, bounds, Morton order (as the LBVH build)
    const int nb256 = (int)((n + 255) / 256);
    TRY(box.alloc((size_t)n * 6 * 4));
    TRY(cen.alloc(n3 * 4));
    k_prep<<<nb256, 256, 0, st>>>(n, dv0.as<float>(), dv1.as<float>(), dv2.as<float>(), box.as<float>(),
                                  cen.as<float>());
    const int npart = std::min(nb256, 1024);
    TRY(part.alloc((size_t)npart * 12 * 4));
    TRY(bnd.alloc(12 * 4));
    k_bounds_partial<<<npart, 256, 0, st>>>(n, box.as<float>(), cen.as<float>(), part.as<float>());
    std::vector<float> hpart((size_t)npart * 12);
    EAR_HIP_TRY(c, hipMemcpyAsync(hpart.data(), part.p, hpart.size() * 4, hipMemcpyDeviceToHost, st));
    EAR_HIP_TRY(c, hipStreamSynchronize(st));
    float hb[12];
    for (int k = 0; k < 3; ++k) {
        hb[k] = INFINITY;
        hb[3 + k] = -INFINITY;
        hb[6 + k] = INFINITY;
        hb[9 + k] = -INFINITY;
    }
    for (int p = 0; p < npart; ++p)
        for (int k = 0; k < 3; ++k) {
            hb[k] = std::min(hb[k], hpart[12 * p + k]);
            hb[3 + k] = std::max(hb[3 + k], hpart[12 * p + 3 + k]);
            hb[6 + k] = std::min(hb[6 + k], hpart[12 * p + 6 + k]);
            hb[9 + k] = std::max(hb[9 + k], hpart[12 * p + 9 + k]);
        }
    EAR_HIP_TRY(c, hipMemcpyAsync(bnd.p, hb, 12 * 4, hipMemcpyHostToDevice, st));
    float ext = 0.0f;
    for (int k = 0; k < 3; ++k) {
        s->bounds[k] = hb[6 + k];
        s->bounds[3 + k] = hb[9 + k];
        ext = std::max(ext, hb[9 + k] - hb[6 + k]);
    }
    const float pad = 1e-5f * ext + 1e-6f;
    TRY(keys.alloc((size_t)n * 4));
    TRY(vals.alloc((size_t)n * 4));
    TRY(keys2.alloc((size_t)n * 4));
    TRY(vals2.alloc((size_t)n * 4));
    // va: the leaves' triangles (PLOC: the Morton order; the SAH build's
    // leaves carry the triangle index itself, va the identity)
    uint32_t *ka = keys.as<uint32_t>(), *va = vals.as<uint32_t>(), *kb = keys2.as<uint32_t>(),
             *vb = vals2.as<uint32_t>();
    k_morton<<<nb256, 256, 0, st>>>(n, cen.as<float>(), bnd.as<float>(), keys.as<uint32_t>(), vals.as<uint32_t>());
    if (builder == EAR_BVH_PLOC) {
        const uint32_t nsb = (n + kSortTile - 1) / kSortTile;
        TRY(counts.alloc((size_t)nsb * 16 * 4));
        for (uint32_t shift = 0; shift < 32; shift += 4) {
            k_radix_count<<<nsb, kSortThreads, 0, st>>>(n, ka, shift, counts.as<uint32_t>(), nsb);
            k_scan_u32<<<1, 1024, 0, st>>>(nsb * 16, counts.as<uint32_t>());
            k_radix_scatter<<<nsb, kSortThreads, 0, st>>>(n, ka, va, kb, vb, shift, counts.as<uint32_t>(), nsb);
            std::swap(ka, kb);
            std::swap(va, vb);
        }
    }

    TRY(dev_alloc(c, (void**)&s->tris, sizeof(float4) * 3 * (size_t)n));
    TRY(dev_alloc(c, (void**)&s->nm, sizeof(float4) * (size_t)n));
    TRY(order.alloc((size_t)n * 4));
    uint32_t levels = 1;
    if (n == 1) {
        // one triangle: a root with one leaf child
        TRY(dev_alloc(c, (void**)&s->nodes4, sizeof(Bvh4Node)));
        s->n_nodes4 = s->n_nodes4_alloc = 1;
        float tb[6];
        EAR_HIP_TRY(c, hipMemcpyAsync(tb, box.p, 24, hipMemcpyDeviceToHost, st));
        EAR_HIP_TRY(c, hipStreamSynchronize(st));
        float lo[3][4], hi[3][4];
        int32_t ref[4] = {(int32_t)kLeafFlag, kEmpty4, kEmpty4, kEmpty4};
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 4; ++i) lo[k][i] = hi[k][i] = 1e30f;
        for (int k = 0; k < 3; ++k) {
            lo[k][0] = tb[k] - (pad + std::fabs(tb[k]) * 1e-6f);
            hi[k][0] = tb[3 + k] + (pad + std::fabs(tb[3 + k]) * 1e-6f);
        }
        Bvh4Node nd;
        if (!quantize_node(lo, hi, ref, nd)) {
            c->last_error = "scene coordinates too large for the quantised BVH nodes";
            return EAR_ERR_INVALID_GEOMETRY;
        }
        EAR_HIP_TRY(c, hipMemcpyAsync(s->nodes4, &nd, sizeof(nd), hipMemcpyHostToDevice, st));
        const uint32_t zero = 0;
        EAR_HIP_TRY(c, hipMemcpyAsync(order.p, &zero, 4, hipMemcpyHostToDevice, st));
        EAR_HIP_TRY(c, hipEventRecord(e1, st));
    } else if (builder == EAR_BVH_SAH_GPU) {
        const uint32_t n_int = n - 1;
        TRY(nd_child.alloc((size_t)n_int * 8));
        TRY(nd_box.alloc((size_t)n_int * 24));
        TRY(nd_cnt.alloc((size_t)n_int * 4));
        TRY(ctr.alloc(8));  // [0] next node id, [1] next level's tasks
        Buf ta(c), tb(c), pa(c), pb(c);
        TRY(ta.alloc((size_t)n * sizeof(SahTask)));
        TRY(tb.alloc((size_t)n * sizeof(SahTask)));
        TRY(pa.alloc((size_t)n * 4));
        TRY(pb.alloc((size_t)n * 4));
        // the triangles in input order (k_morton's vals before any sort)
        EAR_HIP_TRY(c, hipMemcpyAsync(pa.p, va, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
        const SahTask root_task{0u, n, 0u, 1u};
        EAR_HIP_TRY(c, hipMemcpyAsync(ta.p, &root_task, sizeof(root_task), hipMemcpyHostToDevice, st));
        uint32_t* hctr = nullptr;
        EAR_HIP_TRY(c, hipHostMalloc((void**)&hctr, 8, hipHostMallocDefault));
        struct PinGuard {
            uint32_t* p;
            ~PinGuard() {
                if (p) (void)hipHostFree(p);
            }
        } pg{hctr};
        SahTask *tin = ta.as<SahTask>(), *tout = tb.as<SahTask>();
        uint32_t *pin = pa.as<uint32_t>(), *pout = pb.as<uint32_t>();
        uint32_t T = 1, lv = 0;
        const uint32_t init[2] = {1u, 0u};  // node 0 is the root
        EAR_HIP_TRY(c, hipMemcpyAsync(ctr.p, init, 8, hipMemcpyHostToDevice, st));
        while (T > 0) {
            EAR_HIP_TRY(c, hipMemsetAsync(ctr.as<uint32_t>() + 1, 0, 4, st));
            // wide workgroups while the tasks are few and large
            if (T < 64)
                k_sah_level<1024><<<T, 1024, 0, st>>>(tin, pin, pout, cen.as<float>(), box.as<float>(), tout,
                                                       ctr.as<uint32_t>() + 1, ctr.as<uint32_t>(), nd_child.as<int2>(),
                                                       nd_box.as<float>(), nd_cnt.as<uint32_t>());
            else if (T < 2048)
                k_sah_level<256><<<T, 256, 0, st>>>(tin, pin, pout, cen.as<float>(), box.as<float>(), tout,
                                                     ctr.as<uint32_t>() + 1, ctr.as<uint32_t>(), nd_child.as<int2>(),
                                                     nd_box.as<float>(), nd_cnt.as<uint32_t>());
            else
                k_sah_level<64><<<T, 64, 0, st>>>(tin, pin, pout, cen.as<float>(), box.as<float>(), tout,
                                                   ctr.as<uint32_t>() + 1, ctr.as<uint32_t>(), nd_child.as<int2>(),
                                                   nd_box.as<float>(), nd_cnt.as<uint32_t>());
            EAR_HIP_TRY(c, hipGetLastError());
            EAR_HIP_TRY(c, hipMemcpyAsync(hctr, ctr.p, 8, hipMemcpyDeviceToHost, st));
            EAR_HIP_TRY(c, hipStreamSynchronize(st));
            T = hctr[1];
            std::swap(tin, tout);
            std::swap(pin, pout);
            if (++lv > n || hctr[0] > n_int) {
                c->last_error = "device SAH build did not converge";
                return EAR_ERR_INTERNAL;
            }
        }
        if (hctr[0] != n_int) {
            c->last_error = "device SAH build left nodes unmade";
            return EAR_ERR_INTERNAL;
        }
        EAR_HIP_TRY(c, hipEventRecord(e1, st));
        // leaves carry triangle indices: the collapse maps them through va,
        // still the identity k_morton wrote (no sort in this build)
        TRY(collapse_device(c, s, n, 0u, nd_child.as<int2>(), nd_box.as<float>(), nd_cnt.as<uint32_t>(),
                            box.as<float>(), va, leaf_size, pad, order.as<uint32_t>(), hctr, &levels));
    } else {
        // clustering
        const uint32_t n_int = n - 1;
        const uint32_t nbk = (n + kPlocBlock - 1) / kPlocBlock;
        TRY(cref.alloc((size_t)n * 4));
        TRY(cbox.alloc((size_t)n * 24));
        TRY(ccnt.alloc((size_t)n * 4));
        TRY(tref.alloc((size_t)n * 4));
        TRY(tbox.alloc((size_t)n * 24));
        TRY(tcnt.alloc((size_t)n * 4));
        TRY(nn.alloc((size_t)n * 4));
        TRY(keep.alloc((size_t)n * 4));
        TRY(pre.alloc((size_t)n * 4));
        TRY(bsum.alloc((size_t)nbk * 4));
        TRY(nd_child.alloc((size_t)n_int * 8));
        TRY(nd_box.alloc((size_t)n_int * 24));
        TRY(nd_cnt.alloc((size_t)n_int * 4));
        TRY(ctr.alloc(8));  // [0] next node id, [1] clusters left
        EAR_HIP_TRY(c, hipMemsetAsync(ctr.p, 0, 8, st));
        k_ploc_init<<<nb256, 256, 0, st>>>(n, va, box.as<float>(), cref.as<uint32_t>(), cbox.as<float>(),
                                           ccnt.as<uint32_t>());
        uint32_t C = n, iters = 0;
        uint32_t* hctr = nullptr;
        EAR_HIP_TRY(c, hipHostMalloc((void**)&hctr, 8, hipHostMallocDefault));
        struct PinGuard {
            uint32_t* p;
            ~PinGuard() {
                if (p) (void)hipHostFree(p);
            }
        } pg{hctr};
        while (C > 1) {
            const uint32_t nb = (C + kPlocBlock - 1) / kPlocBlock;
            k_ploc_nn<<<nb, kPlocBlock, 0, st>>>(C, cbox.as<float>(), nn.as<uint32_t>());
            k_ploc_merge<<<nb, kPlocBlock, 0, st>>>(C, nn.as<uint32_t>(), cref.as<uint32_t>(), cbox.as<float>(),
                                                    ccnt.as<uint32_t>(), ctr.as<uint32_t>(), nd_child.as<int2>(),
                                                    nd_box.as<float>(), nd_cnt.as<uint32_t>(), tref.as<uint32_t>(),
                                                    tbox.as<float>(), tcnt.as<uint32_t>(), keep.as<uint32_t>(),
                                                    pre.as<uint32_t>(), bsum.as<uint32_t>());
            k_scan_u32<<<1, 1024, 0, st>>>(nb, bsum.as<uint32_t>());
            k_ploc_compact<<<nb, kPlocBlock, 0, st>>>(C, keep.as<uint32_t>(), pre.as<uint32_t>(), bsum.as<uint32_t>(),
                                                      tref.as<uint32_t>(), tbox.as<float>(), tcnt.as<uint32_t>(),
                                                      cref.as<uint32_t>(), cbox.as<float>(), ccnt.as<uint32_t>(),
                                                      ctr.as<uint32_t>() + 1);
            EAR_HIP_TRY(c, hipGetLastError());
            EAR_HIP_TRY(c, hipMemcpyAsync(hctr, ctr.p, 8, hipMemcpyDeviceToHost, st));
            EAR_HIP_TRY(c, hipStreamSynchronize(st));
            // a strict order on the candidate pairs makes the closest pair
            // mutual, so every iteration merges at least one
            if (hctr[1] >= C || ++iters > n) {
                c->last_error = "PLOC clustering made no progress";
                return EAR_ERR_INTERNAL;
            }
            C = hctr[1];
        }
        if (hctr[0] != n_int) {
            c->last_error = "PLOC clustering left a forest";
            return EAR_ERR_INTERNAL;
        }
        uint32_t root = 0;
        EAR_HIP_TRY(c, hipMemcpy(&root, cref.p, 4, hipMemcpyDeviceToHost));
        EAR_HIP_TRY(c, hipEventRecord(e1, st));

        TRY(collapse_device(c, s, n, root, nd_child.as<int2>(), nd_box.as<float>(), nd_cnt.as<uint32_t>(),
                            box.as<float>(), va, leaf_size, pad, order.as<uint32_t>(), hctr, &levels));
    }
    k_gather<<<nb256, 256, 0, st>>>(n, order.as<uint32_t>(), dv0.as<float>(), dv1.as<float>(), dv2.as<float>(),
                                    dmat.as<uint16_t>(), s->tris, s->nm);
    EAR_HIP_TRY(c, hipGetLastError());
    EAR_HIP_TRY(c, hipEventRecord(e2, st));
    EAR_HIP_TRY(c, hipStreamSynchronize(st));
    float ms_all = 0, ms_cl = 0;
    (void)hipEventElapsedTime(&ms_all, e0, e2);
    (void)hipEventElapsedTime(&ms_cl, e0, e1);
    s->build_ms = ms_all;
    s->collapse_ms = ms_all - ms_cl;
    if (3 * levels + 1 >= (uint32_t)kStackDepth) {
        c->last_error = "4-wide BVH deeper than the traversal stack";
        return EAR_ERR_CAPACITY;
    }
    s->depth4 = levels;
    s->max_depth = levels;
    s->width = 4;
    s->n_nodes = 0;
    return EAR_OK;
}

// Materials: the Material ctor (src/Material.cpp:26-72) and Material::Bounce
// (src/Material.cpp:76-83) folded into one row per band:
// {refl / (refl + refr) or -1, spec, retention, refr}.
ear_status material_table(Ctx* c, const ear_material* mats, uint32_t n_mat, std::vector<float4>* out) {
    std::vector<float4>& mb = *out;
    mb.assign(3 * (size_t)std::max<uint32_t>(1, n_mat), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (uint32_t m = 0; m < n_mat; ++m) {
        for (int b = 0; b < 3; ++b) {
            float fl = mats[m].refl[b];
            float fr = (mats[m].flags & EAR_MAT_HAS_REFR) ? mats[m].refr[b] : 0.0f;
            float sp = (mats[m].flags & EAR_MAT_HAS_SPEC) ? mats[m].spec[b] : 0.0f;
            float a = 1.0f;
            a -= fl - 1e-9f;
            if (mats[m].flags & EAR_MAT_HAS_REFR) a -= fr - 1e-9f;
            if (!(a >= 0.0f)) {
                c->last_error = "Invalid material settings (refl + refr > 1)";
                return EAR_ERR_INVALID_MATERIAL;
            }
            float ret = 1.0f - a;
            float fl2 = ((double)fl < 0.0001 && (double)fr < 0.0001) ? -1.0f : fl / (fl + fr);
            mb[3 * m + b] = make_float4(fl2, sp, ret, fr);
        }
    }
    return EAR_OK;
}

ear_status set_scene_materials(Ctx* c, DeviceScene* s, const ear_material* mats, uint32_t n_mat) {
    if (!mats || n_mat != s->n_mat) {
        c->last_error = "the material count must equal the scene's";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    std::vector<float4> mb;
    TRY(material_table(c, mats, n_mat, &mb));
    s->mat_band_host = mb;
    hipStream_t st = c->stream;
    EAR_HIP_TRY(c, hipMemcpyAsync(s->mat_band, mb.data(), sizeof(float4) * mb.size(), hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipStreamSynchronize(st));
    return EAR_OK;
}

ear_status build_scene(Ctx* c, const float* v0, const float* v1, const float* v2, const uint16_t* mat_id,
                       uint32_t n, const ear_material* mats, uint32_t n_mat, uint32_t leaf_size, uint32_t width,
                       uint32_t builder, DeviceScene* out) {
    if (width == 0) width = 4;
    if (builder == EAR_BVH_DEFAULT) builder = width == 4 ? EAR_BVH_SAH_GPU : EAR_BVH_SAH;
    if (builder != EAR_BVH_LBVH && builder != EAR_BVH_SAH && builder != EAR_BVH_PLOC && builder != EAR_BVH_SAH_GPU) {
        c->last_error = "unknown BVH builder";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if (width != 2 && width != 4) {
        c->last_error = "BVH width must be 2 or 4";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if ((builder == EAR_BVH_PLOC || builder == EAR_BVH_SAH_GPU) && width != 4) {
        c->last_error = "the device PLOC and SAH builds make 4-wide trees";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if (n > 0 && (!v0 || !v1 || !v2 || !mat_id)) {
        c->last_error = "null triangle array";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if (n_mat > 0 && !mats) {
        c->last_error = "null material array";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if (n >= kMaxTris) {
        c->last_error = "more than 2^24 - 1 triangles";
        return EAR_ERR_CAPACITY;
    }
    if (leaf_size == 0) leaf_size = 2;  // fastest on the 200k-triangle hall with the SAH tree
    if (leaf_size > 16) {
        c->last_error = "leaf_size must be 1..16";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    std::vector<float4> mb;
    TRY(material_table(c, mats, n_mat, &mb));
    for (uint32_t i = 0; i < n; ++i) {
        if (mat_id[i] >= n_mat) {
            c->last_error = "material id out of range";
            return EAR_ERR_INVALID_GEOMETRY;
        }
    }
    for (size_t i = 0; i < 3 * (size_t)n; ++i) {
        if (!std::isfinite(v0[i]) || !std::isfinite(v1[i]) || !std::isfinite(v2[i])) {
            c->last_error = "non-finite vertex coordinate";
            return EAR_ERR_INVALID_GEOMETRY;
        }
    }

    DeviceScene s;
    s.n_tris = n;
    s.n_mat = n_mat;
    s.leaf_size = leaf_size;
    s.builder = builder;
    s.mat_band_host = mb;
    TRY(dev_alloc(c, (void**)&s.mat_band, sizeof(float4) * mb.size()));
    hipStream_t st = c->stream;
    EAR_HIP_TRY(c, hipMemcpyAsync(s.mat_band, mb.data(), sizeof(float4) * mb.size(), hipMemcpyHostToDevice, st));
    if (n == 0) {
        EAR_HIP_TRY(c, hipStreamSynchronize(st));
        *out = s;
        return EAR_OK;
    }

    if (builder == EAR_BVH_PLOC || builder == EAR_BVH_SAH_GPU) {
        ear_status bs = build_device_tree(c, v0, v1, v2, mat_id, n, leaf_size, builder, &s);
        if (bs != EAR_OK) {
            free_scene(c, &s);
            return bs;
        }
        *out = s;
        return EAR_OK;
    }
    if (builder == EAR_BVH_SAH) {
        ear_status bs = build_sah(c, v0, v1, v2, mat_id, n, leaf_size, &s);
        if (bs != EAR_OK) {
            free_scene(c, &s);
            return bs;
        }
        if (width == 4) {
            ear_status cs = collapse_to_bvh4(c, &s);
            if (cs != EAR_OK) {
                free_scene(c, &s);
                return cs;
            }
        }
        *out = s;
        return EAR_OK;
    }

    hipEvent_t e0, e1;
    EAR_HIP_TRY(c, hipEventCreate(&e0));
    EAR_HIP_TRY(c, hipEventCreate(&e1));

    Buf dv0(c), dv1(c), dv2(c), dmat(c), box(c), cen(c), part(c), bnd(c), keys(c), vals(c), keys2(c), vals2(c),
        counts(c), left(c), right(c), first(c), last(c), pint(c), pleaf(c), nbox(c), arrive(c), depth(c);
    size_t n3 = 3 * (size_t)n;
    TRY(dv0.alloc(n3 * 4));
    TRY(dv1.alloc(n3 * 4));
    TRY(dv2.alloc(n3 * 4));
    TRY(dmat.alloc((size_t)n * 2));
    EAR_HIP_TRY(c, hipMemcpyAsync(dv0.p, v0, n3 * 4, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(dv1.p, v1, n3 * 4, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(dv2.p, v2, n3 * 4, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipMemcpyAsync(dmat.p, mat_id, (size_t)n * 2, hipMemcpyHostToDevice, st));
    EAR_HIP_TRY(c, hipEventRecord(e0, st));

    const int nb256 = (int)((n + 255) / 256);
    TRY(box.alloc((size_t)n * 6 * 4));
    TRY(cen.alloc(n3 * 4));
    k_prep<<<nb256, 256, 0, st>>>(n, dv0.as<float>(), dv1.as<float>(), dv2.as<float>(), box.as<float>(),
                                  cen.as<float>());
    int npart = std::min(nb256, 1024);
    TRY(part.alloc((size_t)npart * 12 * 4));
    TRY(bnd.alloc(12 * 4));
    k_bounds_partial<<<npart, 256, 0, st>>>(n, box.as<float>(), cen.as<float>(), part.as<float>());
    // final min/max over the (at most 1024) per-block partials on the host
    std::vector<float> hpart((size_t)npart * 12);
    EAR_HIP_TRY(c, hipMemcpyAsync(hpart.data(), part.p, hpart.size() * 4, hipMemcpyDeviceToHost, st));
    EAR_HIP_TRY(c, hipStreamSynchronize(st));
    float hb[12];
    for (int k = 0; k < 3; ++k) {
        hb[k] = INFINITY;
        hb[3 + k] = -INFINITY;
        hb[6 + k] = INFINITY;
        hb[9 + k] = -INFINITY;
    }
    for (int p = 0; p < npart; ++p)
        for (int k = 0; k < 3; ++k) {
            hb[k] = std::min(hb[k], hpart[12 * p + k]);
            hb[3 + k] = std::max(hb[3 + k], hpart[12 * p + 3 + k]);
            hb[6 + k] = std::min(hb[6 + k], hpart[12 * p + 6 + k]);
            hb[9 + k] = std::max(hb[9 + k], hpart[12 * p + 9 + k]);
        }
    EAR_HIP_TRY(c, hipMemcpyAsync(bnd.p, hb, 12 * 4, hipMemcpyHostToDevice, st));
    for (int k = 0; k < 3; ++k) {
        s.bounds[k] = hb[6 + k];
        s.bounds[3 + k] = hb[9 + k];
    }
    float ext = 0.0f;
    for (int k = 0; k < 3; ++k) ext = std::max(ext, hb[9 + k] - hb[6 + k]);
    float pad = 1e-5f * ext + 1e-6f;

    TRY(keys.alloc((size_t)n * 4));
    TRY(vals.alloc((size_t)n * 4));
    TRY(keys2.alloc((size_t)n * 4));
    TRY(vals2.alloc((size_t)n * 4));
    k_morton<<<nb256, 256, 0, st>>>(n, cen.as<float>(), bnd.as<float>(), keys.as<uint32_t>(), vals.as<uint32_t>());
    uint32_t nsb = (n + kSortTile - 1) / kSortTile;
    TRY(counts.alloc((size_t)nsb * 16 * 4));
    uint32_t *ka = keys.as<uint32_t>(), *va = vals.as<uint32_t>(), *kb = keys2.as<uint32_t>(),
             *vb = vals2.as<uint32_t>();
    for (uint32_t shift = 0; shift < 32; shift += 4) {
        k_radix_count<<<nsb, kSortThreads, 0, st>>>(n, ka, shift, counts.as<uint32_t>(), nsb);
        k_scan_u32<<<1, 1024, 0, st>>>(nsb * 16, counts.as<uint32_t>());
        k_radix_scatter<<<nsb, kSortThreads, 0, st>>>(n, ka, va, kb, vb, shift, counts.as<uint32_t>(), nsb);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    // after 8 passes the sorted data is back in keys/vals (ka, va)

    uint32_t n_int = n > 1 ? n - 1 : 1;
    TRY(left.alloc((size_t)n_int * 4));
    TRY(right.alloc((size_t)n_int * 4));
    TRY(first.alloc((size_t)n_int * 4));
    TRY(last.alloc((size_t)n_int * 4));
    TRY(pint.alloc((size_t)n_int * 4));
    TRY(pleaf.alloc((size_t)n * 4));
    TRY(nbox.alloc((size_t)n_int * 6 * 4));
    TRY(arrive.alloc((size_t)n_int * 4));
    TRY(depth.alloc(4));
    EAR_HIP_TRY(c, hipMemsetAsync(arrive.p, 0, (size_t)n_int * 4, st));
    EAR_HIP_TRY(c, hipMemsetAsync(depth.p, 0, 4, st));

    s.n_nodes = n_int;
    TRY(dev_alloc(c, (void**)&s.nodes, sizeof(BvhNode) * n_int));
    TRY(dev_alloc(c, (void**)&s.tris, sizeof(float4) * 3 * (size_t)n));
    TRY(dev_alloc(c, (void**)&s.nm, sizeof(float4) * (size_t)n));

    if (n == 1) {
        // one triangle: a root whose second child is a far-away empty box
        std::vector<float> tb(6);
        EAR_HIP_TRY(c, hipMemcpyAsync(tb.data(), box.p, 24, hipMemcpyDeviceToHost, st));
        EAR_HIP_TRY(c, hipStreamSynchronize(st));
        BvhNode nd;
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = tb[k] - (pad + std::fabs(tb[k]) * 1e-6f);
            hi[k] = tb[3 + k] + (pad + std::fabs(tb[3 + k]) * 1e-6f);
        }
        nd.a = make_float4(lo[0], lo[1], lo[2], hi[0]);
        nd.b = make_float4(hi[1], hi[2], 1e30f, 1e30f);
        nd.c = make_float4(1e30f, 1e30f, 1e30f, 1e30f);
        nd.d = make_int4((int32_t)kLeafFlag, kEmptyChild, 0, 0);
        EAR_HIP_TRY(c, hipMemcpyAsync(s.nodes, &nd, sizeof(nd), hipMemcpyHostToDevice, st));
        s.max_depth = 1;
    } else {
        int nbi = (int)((n_int + 255) / 256);
        k_karras<<<nbi, 256, 0, st>>>((int)n, ka, left.as<uint32_t>(), right.as<uint32_t>(), first.as<int>(),
                                      last.as<int>(), pint.as<int>(), pleaf.as<int>());
        k_refit<<<nb256, 256, 0, st>>>((int)n, va, box.as<float>(), left.as<uint32_t>(), right.as<uint32_t>(),
                                       pint.as<int>(), pleaf.as<int>(), nbox.as<float>(), arrive.as<uint32_t>());
        k_depth<<<nb256, 256, 0, st>>>((int)n, pint.as<int>(), pleaf.as<int>(), depth.as<uint32_t>());
        k_emit<<<nbi, 256, 0, st>>>((int)n, leaf_size, pad, va, box.as<float>(), left.as<uint32_t>(),
                                    right.as<uint32_t>(), first.as<int>(), last.as<int>(), nbox.as<float>(),
                                    s.nodes);
        EAR_HIP_TRY(c, hipMemcpyAsync(&s.max_depth, depth.p, 4, hipMemcpyDeviceToHost, st));
    }
    k_gather<<<nb256, 256, 0, st>>>(n, va, dv0.as<float>(), dv1.as<float>(), dv2.as<float>(), dmat.as<uint16_t>(),
                                    s.tris, s.nm);
    EAR_HIP_TRY(c, hipGetLastError());
    EAR_HIP_TRY(c, hipEventRecord(e1, st));
    EAR_HIP_TRY(c, hipStreamSynchronize(st));
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    s.build_ms = ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (s.max_depth >= (uint32_t)kStackDepth) {
        free_scene(c, &s);
        c->last_error = "BVH deeper than the traversal stack";
        return EAR_ERR_CAPACITY;
    }
    if (width == 4) {
        ear_status cs = collapse_to_bvh4(c, &s);
        if (cs != EAR_OK) {
            free_scene(c, &s);
            return cs;
        }
    }
    *out = s;
    return EAR_OK;
}

}  // namespace ear
