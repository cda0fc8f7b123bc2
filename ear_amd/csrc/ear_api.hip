// ear_api.hip -- extern "C" entry points of libear_hip.so (declared in include/ear.h).
//
// Status codes cross the boundary, never exceptions; the host C++ layer turns
// them back into the reference's exceptions (DatatypeException /
// std::runtime_error, src/EAR.cpp:404-408).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "ear_internal.hpp"

namespace ear {

// The no-allocation render's device memory (ear_render_noalloc): one fixed
// region per GPU, held from the library's first use of that GPU until the
// process ends.  In host code `scratch` is only a stand-in for the symbol;
// its device address on each GPU is looked up once, with that GPU current,
// when the first context on it is created (hipGetSymbolAddress), and only
// that address is ever passed on.
static __device__ __align__(256) unsigned char scratch[EAR_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
struct ScratchSlot {
    std::once_flag once;
    void* ptr = nullptr;          // device address of `scratch` on this GPU
    hipError_t err = hipSuccess;  // the lookup's result
    std::mutex busy;              // held for the whole of a no-allocation render
};
ScratchSlot g_scratch[kMaxDevices];

// Called with `device` current.
void lookup_scratch(int device) {
    ScratchSlot& s = g_scratch[device];
    std::call_once(s.once, [&] {
        void* p = nullptr;
        s.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (s.err == hipSuccess)
            s.ptr = p;
        else
            (void)hipGetLastError();
    });
}
}  // namespace

ear_status hip_status(Ctx* c, hipError_t e, const char* what) {
    if (c) c->last_error = std::string(what) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? EAR_ERR_OUT_OF_MEMORY : EAR_ERR_HIP;
}

ear_status dev_alloc(Ctx* c, void** p, size_t bytes) {
    if (c->mem_cap && c->mem_used + bytes > c->mem_cap) {
        c->last_error = "device memory cap exceeded";
        *p = nullptr;
        return EAR_ERR_OUT_OF_MEMORY;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_status(c, e, "hipMalloc");
    }
    c->mem_used += bytes;
    return EAR_OK;
}

void dev_free(Ctx* c, void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    c->mem_used = c->mem_used >= bytes ? c->mem_used - bytes : 0;
}

}  // namespace ear

using namespace ear;

struct ear_ctx : Ctx {};
struct ear_scene : SceneHandle {};

extern "C" {

int ear_abi_version(void) { return EAR_ABI_VERSION; }

const char* ear_status_string(ear_status s) {
    switch (s) {
        case EAR_OK: return "ok";
        case EAR_ERR_INVALID_ARGUMENT: return "invalid argument";
        case EAR_ERR_INVALID_MATERIAL: return "Invalid material settings";
        case EAR_ERR_INVALID_GEOMETRY: return "invalid geometry";
        case EAR_ERR_NO_DEVICE: return "no HIP device";
        case EAR_ERR_OUT_OF_MEMORY: return "out of device memory";
        case EAR_ERR_HIP: return "HIP runtime error";
        case EAR_ERR_CAPACITY: return "capacity exceeded";
        case EAR_ERR_INTERNAL: return "internal error";
    }
    return "unknown status";
}

ear_status ear_ctx_create(int device, ear_ctx** out) {
    if (!out) return EAR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return EAR_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) return EAR_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return EAR_ERR_NO_DEVICE;
    ear_ctx* c = new (std::nothrow) ear_ctx();
    if (!c) return EAR_ERR_OUT_OF_MEMORY;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return EAR_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    if (device < kMaxDevices) lookup_scratch(device);
    *out = c;
    return EAR_OK;
}

ear_status ear_ctx_destroy(ear_ctx* c) {
    if (!c) return EAR_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return EAR_OK;
}

ear_status ear_ctx_set_memory_cap(ear_ctx* c, uint64_t bytes) {
    if (!c) return EAR_ERR_INVALID_ARGUMENT;
    c->mem_cap = bytes;
    return EAR_OK;
}

const char* ear_ctx_last_error(const ear_ctx* c) { return c ? c->last_error.c_str() : ""; }

ear_status ear_ctx_set_tuning(ear_ctx* c, const ear_tuning* t) {
    if (!c || !t) return EAR_ERR_INVALID_ARGUMENT;
    if (t->traversal > EAR_TRAVERSAL_FETCH_ALL) {
        c->last_error = "unknown traversal mode";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if (t->refill > 64 || t->fetch_chunk > (1u << 20) || t->fetch_segments > 64) {
        c->last_error = "refill must be 0..64, fetch_chunk 0..2^20, fetch_segments 0..64";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    c->tuning = *t;
    return EAR_OK;
}

ear_status ear_ctx_get_tuning(const ear_ctx* c, ear_tuning* t) {
    if (!c || !t) return EAR_ERR_INVALID_ARGUMENT;
    *t = c->tuning;
    return EAR_OK;
}

ear_status ear_scene_upload(ear_ctx* c, const float* v0, const float* v1, const float* v2, const uint16_t* mat_id,
                            uint32_t n_tris, const ear_material* mats, uint32_t n_mat, const ear_bvh_opts* opts,
                            ear_scene** out) {
    if (!c || !out) return EAR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (hipSetDevice(c->device) != hipSuccess) return EAR_ERR_NO_DEVICE;
    ear_scene* s = new (std::nothrow) ear_scene();
    if (!s) return EAR_ERR_OUT_OF_MEMORY;
    s->ctx = c;
    ear_status st = build_scene(c, v0, v1, v2, mat_id, n_tris, mats, n_mat, opts ? opts->leaf_size : 0,
                                 opts ? opts->width : 0, opts ? opts->builder : 0, &s->ds);
    if (st != EAR_OK) {
        delete s;
        return st;
    }
    *out = s;
    return EAR_OK;
}

ear_status ear_scene_destroy(ear_scene* s) {
    if (!s) return EAR_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(s->ctx->device);
    free_scene(s->ctx, &s->ds);
    delete s;
    return EAR_OK;
}

ear_status ear_scene_set_materials(ear_scene* s, const ear_material* mats, uint32_t n_mat) {
    if (!s) return EAR_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(s->ctx->device) != hipSuccess) return EAR_ERR_NO_DEVICE;
    return set_scene_materials(s->ctx, &s->ds, mats, n_mat);
}

ear_status ear_scene_get_info(const ear_scene* s, ear_scene_info* info) {
    if (!s || !info) return EAR_ERR_INVALID_ARGUMENT;
    std::memset(info, 0, sizeof(*info));
    info->n_tris = s->ds.n_tris;
    info->n_nodes = s->ds.n_tris > 0 ? s->ds.n_nodes : 0;
    info->max_depth = s->ds.max_depth;
    info->leaf_size = s->ds.leaf_size;
    for (int k = 0; k < 6; ++k) info->bounds[k] = s->ds.bounds[k];
    info->build_ms = s->ds.build_ms;
    info->width = s->ds.width;
    info->node_bytes = s->ds.width == 4 ? 64 : 56;
    info->collapse_ms = s->ds.collapse_ms;
    info->builder = s->ds.builder;
    if (s->ds.width == 4) {
        info->n_nodes = s->ds.n_nodes4;
        info->max_depth = s->ds.depth4 + 1;
    }
    return EAR_OK;
}

ear_status ear_query_closest(ear_scene* s, uint32_t n, const float* o, const float* d, float* t, int32_t* tri) {
    if (!s) return EAR_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(s->ctx->device) != hipSuccess) return EAR_ERR_NO_DEVICE;
    return run_query_closest(s->ctx, &s->ds, n, o, d, t, tri);
}

ear_status ear_query_occluded(ear_scene* s, uint32_t n, const float* p, const float* q, uint8_t* occ) {
    if (!s) return EAR_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(s->ctx->device) != hipSuccess) return EAR_ERR_NO_DEVICE;
    return run_query_occluded(s->ctx, &s->ds, n, p, q, occ);
}

ear_status ear_render(ear_ctx* c, ear_scene* s, const ear_render_desc* d, ear_hist* out, ear_stats* stats) {
    if (!c || !s || !d) return EAR_ERR_INVALID_ARGUMENT;
    if (s->ctx != c) {
        c->last_error = "scene belongs to another context";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if (hipSetDevice(c->device) != hipSuccess) return EAR_ERR_NO_DEVICE;
    try {
        return run_render(c, &s->ds, d, out, stats);
    } catch (const std::bad_alloc&) {
        c->last_error = "host allocation failed";
        return EAR_ERR_OUT_OF_MEMORY;
    } catch (...) {
        c->last_error = "internal exception";
        return EAR_ERR_INTERNAL;
    }
}

ear_status ear_render_noalloc(ear_ctx* c, ear_scene* s, const ear_render_desc* d, ear_hist* out,
                              ear_stats* stats) {
    if (!c || !s || !d) return EAR_ERR_INVALID_ARGUMENT;
    if (s->ctx != c) {
        c->last_error = "scene belongs to another context";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if (c->device >= kMaxDevices) {
        c->last_error = "no scratch region for device ordinals >= 64";
        return EAR_ERR_CAPACITY;
    }
    ScratchSlot& slot = g_scratch[c->device];
    if (!slot.ptr) {
        c->last_error = std::string("scratch region lookup (hipGetSymbolAddress): ") + hipGetErrorString(slot.err);
        return EAR_ERR_HIP;
    }
    if (hipSetDevice(c->device) != hipSuccess) return EAR_ERR_NO_DEVICE;
    try {
        // one region per GPU: renders on it take turns
        std::lock_guard<std::mutex> lock(slot.busy);
        Arena arena;
        arena.base = (char*)slot.ptr;
        arena.size = EAR_NOALLOC_SCRATCH_BYTES;
        DeviceScene* sc[1] = {&s->ds};
        const ear_render_desc* dd[1] = {d};
        ear_hist* oo[1] = {out};
        return run_batch(c, 1, sc, dd, oo, stats, &arena);
    } catch (const std::bad_alloc&) {
        c->last_error = "host allocation failed";
        return EAR_ERR_OUT_OF_MEMORY;
    } catch (...) {
        c->last_error = "internal exception";
        return EAR_ERR_INTERNAL;
    }
}

ear_status ear_render_batch(ear_ctx* c, const ear_batch_item* items, uint32_t n, ear_stats* stats) {
    if (!c || !items || n == 0) return EAR_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(c->device) != hipSuccess) return EAR_ERR_NO_DEVICE;
    try {
        std::vector<DeviceScene*> sc(n);
        std::vector<const ear_render_desc*> ds(n);
        std::vector<ear_hist*> os(n);
        for (uint32_t i = 0; i < n; ++i) {
            if (!items[i].scene || !items[i].desc) return EAR_ERR_INVALID_ARGUMENT;
            if (items[i].scene->ctx != c) {
                c->last_error = "scene belongs to another context";
                return EAR_ERR_INVALID_ARGUMENT;
            }
            sc[i] = &items[i].scene->ds;
            ds[i] = items[i].desc;
            os[i] = items[i].out;
        }
        return run_batch(c, n, sc.data(), ds.data(), os.data(), stats);
    } catch (const std::bad_alloc&) {
        c->last_error = "host allocation failed";
        return EAR_ERR_OUT_OF_MEMORY;
    } catch (...) {
        c->last_error = "internal exception";
        return EAR_ERR_INTERNAL;
    }
}

ear_status ear_render_backward(ear_ctx* c, ear_scene* s, const ear_render_desc* d, const double* adj_value,
                               const double* adj_abs, double* grad_refl, double* grad_spec) {
    if (!c || !s || !d) return EAR_ERR_INVALID_ARGUMENT;
    if (s->ctx != c) {
        c->last_error = "scene belongs to another context";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if (d->estimator != EAR_EST_NEE) {
        c->last_error = "the backward pass covers the NEE estimator only";
        return EAR_ERR_INVALID_ARGUMENT;
    }
    if (hipSetDevice(c->device) != hipSuccess) return EAR_ERR_NO_DEVICE;
    try {
        return run_backward(c, &s->ds, d, adj_value, adj_abs, grad_refl, grad_spec);
    } catch (const std::bad_alloc&) {
        c->last_error = "host allocation failed";
        return EAR_ERR_OUT_OF_MEMORY;
    } catch (...) {
        c->last_error = "internal exception";
        return EAR_ERR_INTERNAL;
    }
}

ear_status ear_render_backward_batch(ear_ctx* c, const ear_backward_item* items, uint32_t n_items) {
    if (!c || !items || n_items == 0) return EAR_ERR_INVALID_ARGUMENT;
    std::vector<DeviceScene*> sc(n_items);
    std::vector<const ear_render_desc*> ds(n_items);
    std::vector<const double*> av(n_items), aa(n_items);
    std::vector<double*> gr(n_items), gsp(n_items);
    for (uint32_t i = 0; i < n_items; ++i) {
        const ear_backward_item& it = items[i];
        if (!it.scene || !it.desc) return EAR_ERR_INVALID_ARGUMENT;
        if (it.scene->ctx != c) {
            c->last_error = "scene belongs to another context";
            return EAR_ERR_INVALID_ARGUMENT;
        }
        if (it.desc->estimator != EAR_EST_NEE) {
            c->last_error = "the backward pass covers the NEE estimator only";
            return EAR_ERR_INVALID_ARGUMENT;
        }
        sc[i] = &it.scene->ds;
        ds[i] = it.desc;
        av[i] = it.adj_value;
        aa[i] = it.adj_abs;
        gr[i] = it.grad_refl;
        gsp[i] = it.grad_spec;
    }
    if (hipSetDevice(c->device) != hipSuccess) return EAR_ERR_NO_DEVICE;
    try {
        return run_backward_batch(c, n_items, sc.data(), ds.data(), av.data(), aa.data(), gr.data(), gsp.data());
    } catch (const std::bad_alloc&) {
        c->last_error = "host allocation failed";
        return EAR_ERR_OUT_OF_MEMORY;
    } catch (...) {
        c->last_error = "internal exception";
        return EAR_ERR_INTERNAL;
    }
}

ear_status ear_convolve(ear_ctx* c, const float* ir, uint32_t n_ir, const float* ir2, uint32_t n_ir2,
                        const float* signal, uint32_t n_sig, float* out) {
    if (!c || !ir || !signal || !out || (n_ir2 && !ir2)) return EAR_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(c->device) != hipSuccess) return EAR_ERR_NO_DEVICE;
    const uint64_t n_out64 = (uint64_t)std::max(n_ir, ir2 ? n_ir2 : 0u) + n_sig + 1;
    if (n_out64 > (1u << 30)) {
        c->last_error = "convolution longer than 2^30 samples";
        return EAR_ERR_CAPACITY;
    }
    const uint32_t n_out = (uint32_t)n_out64;
    try {
        if (!ir2) {
            const float* tr[1] = {ir};
            const uint32_t nt[1] = {n_ir};
            const float* sg[1] = {signal};
            return run_convolve(c, tr, nt, sg, 1, n_sig, out, n_out);
        }
        // the reference's fades: a running float factor from 1 down by 1/N
        // (fade out) and from 0 up by 1/N (fade in), src/Recorder.cpp:172-181
        std::vector<float> s_out(n_sig), s_in(n_sig);
        const float df = 1.0f / (float)std::max<uint32_t>(1, n_sig);
        float f_out = 1.0f, f_in = 0.0f;
        for (uint32_t i = 0; i < n_sig; ++i) {
            s_out[i] = signal[i] * f_out;
            s_in[i] = signal[i] * f_in;
            f_out += -df;
            f_in += df;
        }
        const float* tr[2] = {ir, ir2};
        const uint32_t nt[2] = {n_ir, n_ir2};
        const float* sg[2] = {s_out.data(), s_in.data()};
        return run_convolve(c, tr, nt, sg, 2, n_sig, out, n_out);
    } catch (const std::bad_alloc&) {
        c->last_error = "host allocation failed";
        return EAR_ERR_OUT_OF_MEMORY;
    } catch (...) {
        c->last_error = "internal exception";
        return EAR_ERR_INTERNAL;
    }
}

}  // extern "C"
