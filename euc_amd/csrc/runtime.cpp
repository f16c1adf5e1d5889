// runtime.cpp -- host side of the C ABI declared in include/euchip.h.
//
// Mirrors Pipeline::render (src/pipeline.rs:248-300): target-size selection,
// the size assertion, the early-outs, the MSAA level clamp and render_par's
// band arithmetic (src/pipeline.rs:303-366), then enqueues the frame's kernels
// on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/euchip.h"
#include "internal.hpp"

using namespace euchip;

namespace {

struct Pending {
    const char *name;
    hipEvent_t a, b;
};

struct Stat {
    uint64_t launches = 0;
    double total_ms = 0.0;
};

template <class T>
struct DevArray {
    T *ptr = nullptr;
    size_t cap = 0;
};

}  // namespace

struct euchip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    bool profiling = false;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, Stat> stats;
    // per-frame scratch, grown on demand
    DevArray<float4> clip;
    DevArray<float> vary;
    DevArray<uint32_t> tile_count, bins;
    DevArray<uint2> ovf;
    uint32_t tile_cap = 128;    // bin slots per tile; raised when a frame needed more
    uint32_t frame = 0;
    uint32_t max_groups = 2048; // resident raster workgroups: CUs x 8
    Counters *counters = nullptr;
    Counters *host_counters = nullptr;
};

struct euchip_buffer {
    euchip_ctx *ctx;
    euchip_format fmt;
    uint32_t w, h;
    void *ptr;
    bool owned;
};

struct euchip_mesh {
    euchip_ctx *ctx;
    float *pos = nullptr, *nrm = nullptr, *col = nullptr, *uv = nullptr, *models = nullptr;
    uint32_t *rgba8 = nullptr, *idx = nullptr;
    uint32_t pos_comps = 3, n_vertices = 0, count = 0, n_instances = 1;
};

namespace {

euchip_status fail(euchip_ctx *ctx, euchip_status s, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return s;
}

euchip_status hip_fail(euchip_ctx *ctx, hipError_t e, const char *what) {
    return fail(ctx, e == hipErrorOutOfMemory ? EUCHIP_OOM : EUCHIP_HIP_ERROR, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                   \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return hip_fail(ctx, _e, #expr); \
    } while (0)

template <class T>
euchip_status ensure(euchip_ctx *ctx, DevArray<T> &a, size_t n, const char *what) {
    if (a.cap >= n && a.ptr) return EUCHIP_OK;
    size_t cap = std::max<size_t>(n, a.cap + a.cap / 2);
    cap = std::max<size_t>(cap, 256);
    if (a.ptr) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        HIP_TRY(ctx, hipFree(a.ptr));
        a.ptr = nullptr;
        a.cap = 0;
    }
    hipError_t e = hipMalloc((void **)&a.ptr, cap * sizeof(T));
    if (e != hipSuccess) return hip_fail(ctx, e, what);
    a.cap = cap;
    return EUCHIP_OK;
}

hipEvent_t take_event(euchip_ctx *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Launch one kernel, bracketed by events when profiling.
template <class F>
euchip_status timed(euchip_ctx *ctx, const char *name, F launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->profiling) {
        a = take_event(ctx);
        b = take_event(ctx);
        if (a) (void)hipEventRecord(a, ctx->stream);
    }
    hipError_t e = launch();
    if (e != hipSuccess) return hip_fail(ctx, e, name);
    if (ctx->profiling && a && b) {
        (void)hipEventRecord(b, ctx->stream);
        ctx->pending.push_back({name, a, b});
    }
    return EUCHIP_OK;
}

void resolve_pending(euchip_ctx *ctx) {
    for (auto &p : ctx->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            Stat &s = ctx->stats[p.name];
            s.launches++;
            s.total_ms += ms;
        }
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    ctx->pending.clear();
}

bool is_rgba8(const euchip_buffer *b) { return b && b->fmt == EUCHIP_FORMAT_RGBA8; }
bool is_f32(const euchip_buffer *b) { return b && b->fmt == EUCHIP_FORMAT_F32; }

euchip_status clear_buffer(euchip_ctx *ctx, euchip_buffer *b, uint32_t bits) {
    size_t n = (size_t)b->w * b->h;
    return timed(ctx, "clear", [&] { return launch_fill_u32((uint32_t *)b->ptr, bits, n, ctx->stream); });
}

// The frame: target selection and early-outs of render (pipeline.rs:256-299),
// band arithmetic of render_par (pipeline.rs:323-330), kernel enqueue.
euchip_status render_impl(euchip_ctx *ctx, const euchip_pipeline_desc *d, const float *uniforms,
                          const euchip_mesh *mesh, euchip_buffer *pixel, euchip_buffer *depth, bool sync) {
    if (!ctx || !d || !uniforms || !mesh) return fail(ctx, EUCHIP_INVALID_ARG, "null argument");
    if (mesh->ctx != ctx) return fail(ctx, EUCHIP_INVALID_ARG, "mesh belongs to another context");
    if (d->shader > EUCHIP_SHADER_TEX_SHADOW) return fail(ctx, EUCHIP_INVALID_ARG, "unknown shader %u", d->shader);
    if (d->depth_test > EUCHIP_DEPTH_GREATER || d->cull > EUCHIP_CULL_FRONT || d->band_policy > EUCHIP_BANDS_SEQ ||
        d->xform > EUCHIP_XFORM_MVP)
        return fail(ctx, EUCHIP_INVALID_ARG, "enum field out of range");
    if (pixel && (!is_rgba8(pixel) || pixel->ctx != ctx)) return fail(ctx, EUCHIP_INVALID_ARG, "pixel target must be RGBA8");
    if (depth && (!is_f32(depth) || depth->ctx != ctx)) return fail(ctx, EUCHIP_INVALID_ARG, "depth target must be f32");
    const uint32_t sh = d->shader;
    if ((sh == EUCHIP_SHADER_TEXTURED || sh == EUCHIP_SHADER_TEX_SHADOW) &&
        (!is_rgba8(d->texture) || d->texture->w == 0 || d->texture->h == 0))
        return fail(ctx, EUCHIP_INVALID_ARG, "Linearly-interpolated texture cannot have no size");
    if ((sh == EUCHIP_SHADER_PHONG_SHADOW || sh == EUCHIP_SHADER_TEX_SHADOW) &&
        (!is_f32(d->shadow) || d->shadow->w == 0 || d->shadow->h == 0))
        return fail(ctx, EUCHIP_INVALID_ARG, "Linearly-interpolated texture cannot have no size");
    if ((sh == EUCHIP_SHADER_FLAT && !mesh->rgba8) || (sh == EUCHIP_SHADER_COLOR && !mesh->col) ||
        ((sh == EUCHIP_SHADER_GOURAUD || sh == EUCHIP_SHADER_PHONG || sh == EUCHIP_SHADER_PHONG_SHADOW ||
          sh == EUCHIP_SHADER_TEX_SHADOW) && !mesh->nrm) ||
        ((sh == EUCHIP_SHADER_TEXTURED || sh == EUCHIP_SHADER_TEX_SHADOW) && !mesh->uv))
        return fail(ctx, EUCHIP_INVALID_ARG, "mesh lacks an attribute the shader reads");

    const bool uses_depth = d->depth_test != EUCHIP_DEPTH_NONE || d->depth_write;
    uint32_t pw = pixel ? pixel->w : 0, ph = pixel ? pixel->h : 0;
    uint32_t dw = depth ? depth->w : 0, dh = depth ? depth->h : 0;

    // Clears requested through the load ops happen first, whatever render does.
    bool fuse_color = false, fuse_depth = false;
    uint32_t W = 0, H = 0;
    bool render_nothing = false;
    if (!d->pixel_write && !uses_depth) {
        render_nothing = true;
    } else if (d->pixel_write && !uses_depth) {
        W = pw, H = ph;
    } else if (!d->pixel_write && uses_depth) {
        W = dw, H = dh;
    } else {
        if (pw != dw || ph != dh) {
            if (d->clear_color_enable && pixel) {
                euchip_status s = clear_buffer(ctx, pixel, d->clear_color);
                if (s) return s;
            }
            if (d->clear_depth_enable && depth) {
                uint32_t bits;
                memcpy(&bits, &d->clear_depth, 4);
                euchip_status s = clear_buffer(ctx, depth, bits);
                if (s) return s;
            }
            return fail(ctx, EUCHIP_SIZE_MISMATCH, "Pixel target size is compatible with depth target size");
        }
        W = pw, H = ph;
    }
    const uint32_t L = d->aa_msaa ? std::min<uint32_t>(d->aa_level, 6u) : 0u;
    uint32_t G = H;
    euchip_status band_status = EUCHIP_OK;
    if (!render_nothing && d->band_policy == EUCHIP_BANDS_PAR) {
        G = (uint32_t)((20000ull << L) / std::max<uint32_t>(W, 1u));
        if (G == 0) {
            band_status = fail(ctx, EUCHIP_BAND_ZERO, "attempt to divide by zero (render_par group_rows == 0)");
            render_nothing = true;
        } else {
            uint32_t threads = std::max<uint32_t>(d->threads, 1u);
            if (std::min<uint32_t>(H / G, threads) == 0) render_nothing = true;  // no worker threads spawned
        }
    }
    const uint64_t total = (uint64_t)mesh->count * mesh->n_instances;
    const uint64_t n_tris = total / 3;
    if (W == 0 || H == 0 || n_tris == 0) render_nothing = true;
    if (!render_nothing) {
        fuse_color = d->clear_color_enable && pixel && pixel->w == W && pixel->h == H;
        fuse_depth = d->clear_depth_enable && depth && depth->w == W && depth->h == H;
    }
    if (d->clear_color_enable && pixel && !fuse_color) {
        euchip_status s = clear_buffer(ctx, pixel, d->clear_color);
        if (s) return s;
    }
    if (d->clear_depth_enable && depth && !fuse_depth) {
        uint32_t bits;
        memcpy(&bits, &d->clear_depth, 4);
        euchip_status s = clear_buffer(ctx, depth, bits);
        if (s) return s;
    }
    if (render_nothing) {
        if (sync) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return band_status;
    }
    if (W >= (1u << 24) || H >= (1u << 24)) return fail(ctx, EUCHIP_INVALID_ARG, "target larger than 2^24");

    FrameParams p;
    memset(&p, 0, sizeof(p));
    p.W = W, p.H = H;
    p.pixel = pixel ? (uint32_t *)pixel->ptr : nullptr;
    p.depth = depth ? (float *)depth->ptr : nullptr;
    // Targets that are not the render size are never touched (Empty-like use).
    if (pixel && (pixel->w != W || pixel->h != H)) p.pixel = nullptr;
    if (depth && (depth->w != W || depth->h != H)) p.depth = nullptr;
    p.shader = sh, p.xform = d->xform, p.pixel_write = d->pixel_write ? 1u : 0u;
    p.depth_test = d->depth_test, p.depth_write = d->depth_write ? 1u : 0u, p.y_up = d->y_up ? 1u : 0u;
    p.z_clip = d->z_clip ? 1u : 0u, p.z_min = d->z_min, p.z_max = d->z_max;
    p.cull = d->cull, p.msaa_level = L, p.band_par = d->band_policy == EUCHIP_BANDS_PAR ? 1u : 0u, p.G = G;
    p.clear_color_en = fuse_color ? 1u : 0u, p.clear_color = d->clear_color;
    p.clear_depth_en = fuse_depth ? 1u : 0u, p.clear_depth = d->clear_depth;
    if (uses_depth && !p.depth) return fail(ctx, EUCHIP_INVALID_ARG, "depth test without a depth target");
    p.pos = mesh->pos, p.nrm = mesh->nrm, p.col = mesh->col, p.uv = mesh->uv, p.rgba8 = mesh->rgba8;
    p.idx = mesh->idx, p.models = mesh->models;
    p.pos_comps = mesh->pos_comps, p.count = mesh->count, p.n_vertices = mesh->n_vertices;
    p.n_instances = mesh->n_instances, p.n_tris = (uint32_t)n_tris;
    p.ndc_x = 2.0f / (float)W;
    p.ndc_y = -2.0f / (float)H;
    if (d->texture) p.tex = (const uint32_t *)d->texture->ptr, p.tex_w = d->texture->w, p.tex_h = d->texture->h;
    if (d->shadow) p.shadow = (const float *)d->shadow->ptr, p.sh_w = d->shadow->w, p.sh_h = d->shadow->h;
    memcpy(p.u, uniforms, sizeof(p.u));
    p.ntx = (W + kTileW - 1) / kTileW;
    p.nty = (H + kTileH - 1) / kTileH;
    const size_t ntiles = (size_t)p.ntx * p.nty;

    euchip_status s;
    const size_t nsv = (size_t)mesh->n_instances * mesh->n_vertices;
    if ((s = ensure(ctx, ctx->clip, std::max<size_t>(nsv, 1), "clip positions"))) return s;
    if ((s = ensure(ctx, ctx->vary, std::max<size_t>(nsv * vd_count(sh), 1), "vertex data"))) return s;
    const bool fresh_tiles = ctx->tile_count.cap < ntiles;
    if ((s = ensure(ctx, ctx->tile_count, ntiles, "tile counts"))) return s;
    if (fresh_tiles) HIP_TRY(ctx, hipMemsetAsync(ctx->tile_count.ptr, 0, ctx->tile_count.cap * 4, ctx->stream));
    if ((s = ensure(ctx, ctx->ovf, std::max<size_t>((size_t)1 << 20, (size_t)n_tris), "overflow list"))) return s;

    for (int attempt = 0;; attempt++) {
        if ((s = ensure(ctx, ctx->bins, ntiles * ctx->tile_cap, "bins"))) return s;
        p.clip = ctx->clip.ptr;
        p.vary = ctx->vary.ptr;
        p.ctr = ctx->counters;
        p.tile_count = ctx->tile_count.ptr;
        p.bins = ctx->bins.ptr;
        p.tile_cap = ctx->tile_cap;
        p.ovf = ctx->ovf.ptr;
        p.ovf_cap = (uint32_t)std::min<size_t>(ctx->ovf.cap, 0xFFFFFFF0ull);
        p.frame = ctx->frame++;
        if ((s = timed(ctx, "setup", [&] { return launch_setup(p, ctx->stream); }))) return s;
        if ((s = timed(ctx, "raster", [&] { return launch_raster(p, ctx->stream, ctx->max_groups); }))) return s;
        if (!sync) return EUCHIP_OK;
        HIP_TRY(ctx, hipMemcpyAsync(ctx->host_counters, ctx->counters, sizeof(Counters), hipMemcpyDeviceToHost,
                                    ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->profiling) resolve_pending(ctx);
        const Counters &hc = *ctx->host_counters;
        if (hc.max_count > ctx->tile_cap) {  // grow the bins for the frames that follow
            uint32_t c = ctx->tile_cap;
            while (c < hc.max_count && c < (1u << 24)) c <<= 1;
            ctx->tile_cap = c;
        }
        if (!(hc.flags & FLAG_BIN_OVERFLOW)) return EUCHIP_OK;
        // The overflow list filled up: nothing was written.  Reset and retry larger.
        HIP_TRY(ctx, hipMemsetAsync(ctx->counters, 0, sizeof(Counters), ctx->stream));
        HIP_TRY(ctx, hipMemsetAsync(ctx->tile_count.ptr, 0, ctx->tile_count.cap * 4, ctx->stream));
        if (attempt >= 3) return fail(ctx, EUCHIP_OVER_BUDGET, "bin storage overflow");
        if ((s = ensure(ctx, ctx->ovf, ctx->ovf.cap * 4, "overflow list"))) return s;
    }
}

}  // namespace

extern "C" {

const char *euchip_status_string(euchip_status s) {
    switch (s) {
    case EUCHIP_OK: return "ok";
    case EUCHIP_SIZE_MISMATCH: return "size mismatch";
    case EUCHIP_BAND_ZERO: return "band height is zero";
    case EUCHIP_INVALID_ARG: return "invalid argument";
    case EUCHIP_OOM: return "out of device memory";
    case EUCHIP_OVER_BUDGET: return "over budget";
    case EUCHIP_HIP_ERROR: return "HIP error";
    case EUCHIP_NO_DEVICE: return "no device";
    case EUCHIP_NOT_READY: return "not ready";
    }
    return "unknown status";
}

euchip_status euchip_ctx_create(int device, euchip_ctx **out) {
    if (!out) return EUCHIP_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return EUCHIP_NO_DEVICE;
    if (device < 0 || device >= n) return EUCHIP_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return EUCHIP_HIP_ERROR;
    euchip_ctx *ctx = new euchip_ctx();
    ctx->device = device;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->max_groups = (uint32_t)cus * 8u;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&ctx->counters, sizeof(Counters)) != hipSuccess ||
        hipHostMalloc((void **)&ctx->host_counters, sizeof(Counters), hipHostMallocDefault) != hipSuccess ||
        hipMemset(ctx->counters, 0, sizeof(Counters)) != hipSuccess) {
        euchip_ctx_destroy(ctx);
        return EUCHIP_HIP_ERROR;
    }
    *out = ctx;
    return EUCHIP_OK;
}

euchip_status euchip_ctx_destroy(euchip_ctx *ctx) {
    if (!ctx) return EUCHIP_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    (void)hipFree(ctx->clip.ptr);
    (void)hipFree(ctx->vary.ptr);
    (void)hipFree(ctx->tile_count.ptr);
    (void)hipFree(ctx->bins.ptr);
    (void)hipFree(ctx->ovf.ptr);
    (void)hipFree(ctx->counters);
    if (ctx->host_counters) (void)hipHostFree(ctx->host_counters);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return EUCHIP_OK;
}

const char *euchip_last_error(const euchip_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

void *euchip_ctx_stream(euchip_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

euchip_status euchip_synchronize(euchip_ctx *ctx) {
    if (!ctx) return EUCHIP_INVALID_ARG;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->host_counters, ctx->counters, sizeof(Counters), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->profiling) resolve_pending(ctx);
    const Counters &hc = *ctx->host_counters;
    if (hc.max_count > ctx->tile_cap) {
        uint32_t c = ctx->tile_cap;
        while (c < hc.max_count && c < (1u << 24)) c <<= 1;
        ctx->tile_cap = c;
    }
    if (hc.flags & FLAG_BIN_OVERFLOW) {
        HIP_TRY(ctx, hipMemset(ctx->counters, 0, sizeof(Counters)));
        HIP_TRY(ctx, hipMemset(ctx->tile_count.ptr, 0, ctx->tile_count.cap * 4));
        return fail(ctx, EUCHIP_OVER_BUDGET, "bin storage overflow in an asynchronous frame; it was dropped");
    }
    return EUCHIP_OK;
}

euchip_status euchip_set_profiling(euchip_ctx *ctx, int enable) {
    if (!ctx) return EUCHIP_INVALID_ARG;
    if (!enable && ctx->profiling) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        resolve_pending(ctx);
    }
    ctx->profiling = enable != 0;
    return EUCHIP_OK;
}

int euchip_kernel_stats(euchip_ctx *ctx, euchip_kernel_stat *out, int max) {
    if (!ctx) return 0;
    int i = 0;
    for (auto &kv : ctx->stats) {
        if (out && i < max) {
            memset(&out[i], 0, sizeof(out[i]));
            strncpy(out[i].name, kv.first.c_str(), sizeof(out[i].name) - 1);
            out[i].launches = kv.second.launches;
            out[i].total_ms = kv.second.total_ms;
        }
        i++;
    }
    return i;
}

void euchip_reset_stats(euchip_ctx *ctx) {
    if (ctx) ctx->stats.clear();
}

euchip_status euchip_buffer_create(euchip_ctx *ctx, euchip_format fmt, uint32_t w, uint32_t h, euchip_buffer **out) {
    if (!ctx || !out || (fmt != EUCHIP_FORMAT_RGBA8 && fmt != EUCHIP_FORMAT_F32))
        return fail(ctx, EUCHIP_INVALID_ARG, "bad buffer arguments");
    *out = nullptr;
    void *ptr = nullptr;
    size_t bytes = (size_t)w * h * 4;
    if (bytes) {
        hipError_t e = hipMalloc(&ptr, bytes);
        if (e != hipSuccess) return hip_fail(ctx, e, "buffer");
    }
    *out = new euchip_buffer{ctx, fmt, w, h, ptr, true};
    return EUCHIP_OK;
}

euchip_status euchip_buffer_wrap(euchip_ctx *ctx, euchip_format fmt, uint32_t w, uint32_t h, void *device_ptr,
                                 euchip_buffer **out) {
    if (!ctx || !out || (!device_ptr && (size_t)w * h) || (fmt != EUCHIP_FORMAT_RGBA8 && fmt != EUCHIP_FORMAT_F32))
        return fail(ctx, EUCHIP_INVALID_ARG, "bad buffer arguments");
    *out = new euchip_buffer{ctx, fmt, w, h, device_ptr, false};
    return EUCHIP_OK;
}

euchip_status euchip_buffer_destroy(euchip_buffer *b) {
    if (!b) return EUCHIP_INVALID_ARG;
    if (b->owned && b->ptr) {
        (void)hipStreamSynchronize(b->ctx->stream);
        (void)hipFree(b->ptr);
    }
    delete b;
    return EUCHIP_OK;
}

euchip_status euchip_buffer_upload(euchip_buffer *b, const void *host, size_t bytes) {
    if (!b || !host || bytes != (size_t)b->w * b->h * 4) return fail(b ? b->ctx : nullptr, EUCHIP_INVALID_ARG, "upload size");
    HIP_TRY(b->ctx, hipMemcpyAsync(b->ptr, host, bytes, hipMemcpyHostToDevice, b->ctx->stream));
    HIP_TRY(b->ctx, hipStreamSynchronize(b->ctx->stream));
    return EUCHIP_OK;
}

euchip_status euchip_buffer_download(const euchip_buffer *b, void *host, size_t bytes) {
    if (!b || !host || bytes != (size_t)b->w * b->h * 4) return fail(b ? b->ctx : nullptr, EUCHIP_INVALID_ARG, "download size");
    HIP_TRY(b->ctx, hipMemcpyAsync(host, b->ptr, bytes, hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_TRY(b->ctx, hipStreamSynchronize(b->ctx->stream));
    return EUCHIP_OK;
}

euchip_status euchip_buffer_clear(euchip_buffer *b, const void *texel) {
    if (!b || !texel) return EUCHIP_INVALID_ARG;
    uint32_t bits;
    memcpy(&bits, texel, 4);
    return clear_buffer(b->ctx, b, bits);
}

void *euchip_buffer_device_ptr(const euchip_buffer *b) { return b ? b->ptr : nullptr; }

euchip_status euchip_mesh_create(euchip_ctx *ctx, const euchip_mesh_desc *d, euchip_mesh **out) {
    if (!ctx || !d || !out || !d->pos || (d->pos_comps != 3 && d->pos_comps != 4))
        return fail(ctx, EUCHIP_INVALID_ARG, "bad mesh arguments");
    *out = nullptr;
    if (d->n_vertices >= (1ull << 31) || d->count >= (1ull << 31) || d->n_instances == 0 ||
        (uint64_t)d->n_instances * d->n_vertices >= (1ull << 32) || (uint64_t)d->n_instances * d->count >= (1ull << 32))
        return fail(ctx, EUCHIP_INVALID_ARG, "mesh too large for 32-bit stream ids");
    // Every index must address a vertex (IndexedVertices would panic on a slice OOB).
    if (d->idx) {
        for (uint64_t i = 0; i < d->count; i++)
            if (d->idx[i] >= d->n_vertices)
                return fail(ctx, EUCHIP_INVALID_ARG, "index %llu out of range (%u >= %llu)", (unsigned long long)i,
                            d->idx[i], (unsigned long long)d->n_vertices);
    } else if (d->count > d->n_vertices) {
        return fail(ctx, EUCHIP_INVALID_ARG, "stream longer than the vertex array");
    }
    euchip_mesh *m = new euchip_mesh();
    m->ctx = ctx;
    m->pos_comps = d->pos_comps;
    m->n_vertices = (uint32_t)d->n_vertices;
    m->count = (uint32_t)d->count;
    m->n_instances = d->n_instances;
    auto up = [&](void **dst, const void *src, size_t bytes) -> hipError_t {
        if (!src || !bytes) return hipSuccess;
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) return e;
        return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    };
    const size_t nv = d->n_vertices;
    std::vector<float> ident;
    const float *models = d->models;
    if (!models) {
        ident.assign((size_t)16 * d->n_instances, 0.0f);
        for (uint32_t i = 0; i < d->n_instances; i++)
            for (int k = 0; k < 4; k++) ident[16 * i + 5 * k] = 1.0f;
        models = ident.data();
    }
    hipError_t e;
    if ((e = up((void **)&m->pos, d->pos, nv * d->pos_comps * 4)) != hipSuccess ||
        (e = up((void **)&m->nrm, d->nrm, nv * 12)) != hipSuccess ||
        (e = up((void **)&m->col, d->col, nv * 16)) != hipSuccess ||
        (e = up((void **)&m->uv, d->uv, nv * 8)) != hipSuccess ||
        (e = up((void **)&m->rgba8, d->rgba8, nv * 4)) != hipSuccess ||
        (e = up((void **)&m->idx, d->idx, d->count * 4)) != hipSuccess ||
        (e = up((void **)&m->models, models, (size_t)16 * 4 * d->n_instances)) != hipSuccess) {
        euchip_mesh_destroy(m);
        return hip_fail(ctx, e, "mesh upload");
    }
    *out = m;
    return EUCHIP_OK;
}

euchip_status euchip_mesh_destroy(euchip_mesh *m) {
    if (!m) return EUCHIP_INVALID_ARG;
    (void)hipStreamSynchronize(m->ctx->stream);
    (void)hipFree(m->pos);
    (void)hipFree(m->nrm);
    (void)hipFree(m->col);
    (void)hipFree(m->uv);
    (void)hipFree(m->rgba8);
    (void)hipFree(m->idx);
    (void)hipFree(m->models);
    delete m;
    return EUCHIP_OK;
}

euchip_status euchip_mesh_set_models(euchip_mesh *m, const float *models, uint32_t n_instances) {
    if (!m || !models || n_instances != m->n_instances) return fail(m ? m->ctx : nullptr, EUCHIP_INVALID_ARG, "models");
    HIP_TRY(m->ctx, hipMemcpyAsync(m->models, models, (size_t)16 * 4 * n_instances, hipMemcpyHostToDevice, m->ctx->stream));
    HIP_TRY(m->ctx, hipStreamSynchronize(m->ctx->stream));
    return EUCHIP_OK;
}

euchip_status euchip_render(euchip_ctx *ctx, const euchip_pipeline_desc *desc, const float *uniforms,
                            const euchip_mesh *mesh, euchip_buffer *pixel, euchip_buffer *depth) {
    if (ctx) (void)hipSetDevice(ctx->device);
    return render_impl(ctx, desc, uniforms, mesh, pixel, depth, true);
}

euchip_status euchip_render_async(euchip_ctx *ctx, const euchip_pipeline_desc *desc, const float *uniforms,
                                  const euchip_mesh *mesh, euchip_buffer *pixel, euchip_buffer *depth) {
    if (ctx) (void)hipSetDevice(ctx->device);
    return render_impl(ctx, desc, uniforms, mesh, pixel, depth, false);
}

}  // extern "C"
