// runtime.cpp -- host side of the C ABI declared in include/euchip.h.
//
// Mirrors Pipeline::render (src/pipeline.rs:248-300): target-size selection,
// the size assertion, the early-outs, the MSAA level clamp and render_par's
// band arithmetic (src/pipeline.rs:303-366), then enqueues the frame's kernels
// on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
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

// Device memory replaced while work already enqueued may still read it: freed
// once `done` (recorded after that work) has completed.
struct Deferred {
    void *ptr;
    size_t bytes;
    hipEvent_t done;
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
    // per-frame scratch, grown on demand (old storage is freed once the frames
    // enqueued on it have finished)
    DevArray<uint4> rec0, rec1;
    DevArray<float> vary;
    DevArray<TileBin> tile_count;
    DevArray<uint32_t> bins;
    DevArray<uint32_t> pool;    // overflow segments
    DevArray<uint32_t> heavy;   // heavy-tile list
    DevArray<uint32_t> busy;    // busy-tile lists (one per XCD tile range)
    DevArray<uint32_t> big;     // big-triangle list
    DevArray<uint4> big_rec;    // per triangle of a checkpoint frame: what its row tasks read
    DevArray<uint4> ck;         // row checkpoints of big triangles
    DevArray<uint2> cktab;     // per tile: its list of (triangle, row mask; key-order bound)
    DevArray<uint2> ck_count;  // the lists' lengths and occluder bounds, two halves like tile_count
    uint32_t ck_dirty[2] = {0, 0};  // counts a checkpoint frame left in each half, zeroed by the next frame
    DevArray<uint32_t> lrec;    // line records
    DevArray<uint64_t> lkey;    // per-pixel winner keys of a lines frame
    std::vector<Deferred> deferred;
    uint32_t counts_used[2] = {0, 0};  // tiles of the last frame that counted into each half of tile_count
    uint32_t tile_cap = 128;    // bin slots per tile; raised when a frame needed more (within the bin budget)
    size_t pool_want = 0;       // pool words the next frames should have (from the reports)
    uint32_t frame = 0;
    uint32_t max_groups = 2048; // resident raster workgroups: CUs x 8
    uint32_t big_tile_sx = 5;   // tiles of big-triangle frames: (1 << sx) x (256 >> sx) pixels
    Counters *counters = nullptr;
    Report *report = nullptr;   // host memory the raster pass writes each frame's counts to
    Report *report_dev = nullptr;  // its device address
    uint32_t seen_frame = ~0u;  // the last report taken into account
    uint32_t seen_dropped = 0;  // frames reported dropped before the last synchronize
    size_t bytes_in_use = 0;    // device bytes held through this context
    size_t budget = 0;          // 0: no limit
    void *noalloc = nullptr;    // device address of the no-allocation scratch on this device
    hipError_t noalloc_err = hipSuccess;
    uint32_t *na_readback = nullptr;  // pinned host words: flags and pool use of a no-allocation frame
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
    size_t bytes = 0;
};

namespace {

// Host half of render() for one request (defined below).
struct Plan {
    View v;
    bool render = false;
    euchip_status status = EUCHIP_OK;
    bool clear_pixel = false, clear_depth = false;  // clears left to fill kernels
    uint64_t n_tris = 0, n_sv = 0;
};

}  // namespace

struct euchip_batch {
    euchip_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    euchip_pipeline_desc desc{};
    std::vector<euchip_request> reqs;
    std::vector<std::array<float, kUniforms>> uniforms;
    std::vector<Plan> plans;
    std::vector<uint32_t> active;       // requests whose render() runs kernels
    uint32_t tile_sx = 5;
    std::vector<uint32_t> caps[3];      // bin slots per tile, per tile width (sx - 3) and active request
    bool tables_dirty = true;
    // device tables and scratch
    DevArray<uint4> tables;             // views, workgroup map and tile map, one block
    uint8_t *stage = nullptr;           // pinned staging of the tables
    size_t stage_bytes = 0;
    size_t n_tiles = 0;                 // tiles of the launch (their counts are zeroed before setup)
    DevArray<uint4> rec0, rec1;
    DevArray<float> vary;
    DevArray<TileBin> tile_count;
    DevArray<uint32_t> bins, ctrbuf;              // ctrbuf: Counters, then req_max per active request
    DevArray<uint32_t> busy;                      // busy-tile lists (one slot per tile)
    DevArray<uint32_t> pool;                      // overflow segments
    size_t pool_want = 0;                         // pool words the last launch asked for, plus a quarter
    uint32_t *host_ctr = nullptr;       // pinned copy of ctrbuf
    size_t host_ctr_words = 0;
    FrameParams p{};
    uint32_t setup_blocks = 0;
    uint64_t tiles_total = 0, bins_total = 0, sv_total = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool launched = false;
    // euchip_batch_set_multi: one part per distinct desc (this batch's own
    // requests are then empty)
    std::vector<euchip_batch *> parts;
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

// Device allocations of a context, checked against its budget.
euchip_status dev_alloc(euchip_ctx *ctx, void **ptr, size_t bytes, const char *what) {
    *ptr = nullptr;
    if (bytes == 0) return EUCHIP_OK;
    if (ctx->budget && ctx->bytes_in_use + bytes > ctx->budget)
        return fail(ctx, EUCHIP_OVER_BUDGET, "%s: %zu bytes would exceed the device budget (%zu of %zu in use)", what,
                    bytes, ctx->bytes_in_use, ctx->budget);
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return hip_fail(ctx, e, what);
    }
    ctx->bytes_in_use += bytes;
    return EUCHIP_OK;
}

void dev_free(euchip_ctx *ctx, void *ptr, size_t bytes) {
    if (!ptr) return;
    (void)hipFree(ptr);
    ctx->bytes_in_use -= std::min(bytes, ctx->bytes_in_use);
}

// Free deferred storage whose frames have finished (all of it with `wait`).
void reap_deferred(euchip_ctx *ctx, bool wait) {
    size_t k = 0;
    for (Deferred &d : ctx->deferred) {
        if (wait) (void)hipEventSynchronize(d.done);
        const bool done = wait || hipEventQuery(d.done) != hipErrorNotReady;
        if (done) {
            dev_free(ctx, d.ptr, d.bytes);
            (void)hipEventDestroy(d.done);
        } else {
            ctx->deferred[k++] = d;
        }
    }
    ctx->deferred.resize(k);
}

// Grow `a` to at least n elements (contents are not kept).  `s` is the stream
// whose enqueued work may still use the old storage: it is freed once that
// work has finished, without waiting for it here.
template <class T>
euchip_status ensure(euchip_ctx *ctx, DevArray<T> &a, size_t n, const char *what, hipStream_t s) {
    if (a.cap >= n && a.ptr) return EUCHIP_OK;
    size_t cap = std::max<size_t>(n, a.cap + a.cap / 2);
    cap = std::max<size_t>(cap, 256);
    if (a.ptr) {
        hipEvent_t e = nullptr;
        HIP_TRY(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        const hipError_t er = hipEventRecord(e, s);
        if (er != hipSuccess) {
            (void)hipEventDestroy(e);
            return hip_fail(ctx, er, "hipEventRecord");
        }
        ctx->deferred.push_back({a.ptr, a.cap * sizeof(T), e});
        a.ptr = nullptr;
        a.cap = 0;
    }
    euchip_status st = dev_alloc(ctx, (void **)&a.ptr, cap * sizeof(T), what);
    if (st == EUCHIP_OVER_BUDGET && !ctx->deferred.empty()) {  // storage on its way out counts too
        reap_deferred(ctx, true);
        st = dev_alloc(ctx, (void **)&a.ptr, cap * sizeof(T), what);
    }
    if (st) return st;
    a.cap = cap;
    return EUCHIP_OK;
}

template <class T>
void release(euchip_ctx *ctx, DevArray<T> &a) {
    dev_free(ctx, a.ptr, a.cap * sizeof(T));
    a.ptr = nullptr;
    a.cap = 0;
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

// Launch one kernel on `s`, bracketed by events on `s` when profiling.
template <class F>
euchip_status timed_on(euchip_ctx *ctx, hipStream_t s, const char *name, F launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->profiling) {
        a = take_event(ctx);
        b = take_event(ctx);
        if (a) (void)hipEventRecord(a, s);
    }
    hipError_t e = launch();
    if (e != hipSuccess) return hip_fail(ctx, e, name);
    if (ctx->profiling && a && b) {
        (void)hipEventRecord(b, s);
        ctx->pending.push_back({name, a, b});
    }
    return EUCHIP_OK;
}

template <class F>
euchip_status timed(euchip_ctx *ctx, const char *name, F launch) {
    return timed_on(ctx, ctx->stream, name, launch);
}

void resolve_pending(euchip_ctx *ctx) {
    std::vector<Pending> keep;  // events of launches still in flight (another stream)
    for (auto &p : ctx->pending) {
        if (hipEventQuery(p.b) == hipErrorNotReady) {
            keep.push_back(p);
            continue;
        }
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            Stat &s = ctx->stats[p.name];
            s.launches++;
            s.total_ms += ms;
        }
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    ctx->pending.swap(keep);
}

bool is_rgba8(const euchip_buffer *b) { return b && b->fmt == EUCHIP_FORMAT_RGBA8; }
bool is_f32(const euchip_buffer *b) { return b && b->fmt == EUCHIP_FORMAT_F32; }

euchip_status clear_buffer(euchip_ctx *ctx, euchip_buffer *b, uint32_t bits, hipStream_t s) {
    size_t n = (size_t)b->w * b->h;
    return timed(ctx, "clear", [&] { return launch_fill_u32((uint32_t *)b->ptr, bits, n, s); });
}

uint32_t f32_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

// Pipeline state shared by every request of a launch.
void fill_pipe(FrameParams &p, const euchip_pipeline_desc *d) {
    p.shader = d->shader, p.xform = d->xform, p.pixel_write = d->pixel_write ? 1u : 0u;
    p.depth_test = d->depth_test, p.depth_write = d->depth_write ? 1u : 0u, p.y_up = d->y_up ? 1u : 0u;
    p.z_clip = d->z_clip ? 1u : 0u, p.z_min = d->z_min, p.z_max = d->z_max;
    p.cull = d->cull, p.msaa_level = d->aa_msaa ? std::min<uint32_t>(d->aa_level, 6u) : 0u;
    p.band_par = d->band_policy == EUCHIP_BANDS_PAR ? 1u : 0u;
    p.bgra = d->pixel_format == EUCHIP_PIXEL_BGRA8 ? 1u : 0u;
    p.sampler = d->sampler != EUCHIP_SAMPLER_DEFAULT ? d->sampler
              : d->shader == EUCHIP_SHADER_TEX_SHADOW ? (uint32_t)EUCHIP_SAMPLER_LINEAR_CLAMPED
                                                      : (uint32_t)EUCHIP_SAMPLER_LINEAR;
    p.geometry = d->geometry;
}

// Triangles a stream of `count` elements per instance yields after the geometry
// stage; for the line kinds, the lines (LineTriangleList: three per triangle).
uint64_t stream_triangles(const euchip_pipeline_desc *d, uint64_t count, uint64_t instances) {
    if (d->primitives == EUCHIP_PRIMS_LINE_LIST) return count * instances / 2;
    const uint64_t tris = count * instances / 3 * (d->geometry == EUCHIP_GEOMETRY_BOTH_SIDES ? 2 : 1);
    return d->primitives == EUCHIP_PRIMS_LINE_TRIANGLE_LIST ? 3 * tris : tris;
}

// Batches, servers and the no-allocation render draw triangle lists only.
euchip_status check_triangles(euchip_ctx *ctx, const euchip_pipeline_desc *d) {
    if (d->primitives != EUCHIP_PRIMS_TRIANGLE_LIST)
        return fail(ctx, EUCHIP_INVALID_ARG, "line primitives are drawn by euchip_render / euchip_render_async only");
    return EUCHIP_OK;
}

euchip_status check_desc(euchip_ctx *ctx, const euchip_pipeline_desc *d) {
    if (d->shader > EUCHIP_SHADER_TEX_SHADOW) return fail(ctx, EUCHIP_INVALID_ARG, "unknown shader %u", d->shader);
    if (d->depth_test > EUCHIP_DEPTH_GREATER || d->cull > EUCHIP_CULL_FRONT || d->band_policy > EUCHIP_BANDS_SEQ ||
        d->xform > EUCHIP_XFORM_MVP || d->pixel_format > EUCHIP_PIXEL_BGRA8 || d->sampler > EUCHIP_SAMPLER_NEAREST_MIRRORED ||
        d->geometry > EUCHIP_GEOMETRY_BOTH_SIDES || d->primitives > EUCHIP_PRIMS_LINE_TRIANGLE_LIST)
        return fail(ctx, EUCHIP_INVALID_ARG, "enum field out of range");
    if (d->primitives == EUCHIP_PRIMS_LINE_LIST && d->geometry != EUCHIP_GEOMETRY_IDENTITY)
        return fail(ctx, EUCHIP_INVALID_ARG, "the geometry stages act on triangles; a LineList takes the identity");
    return EUCHIP_OK;
}

// The host half of render() for one request (pipeline.rs:256-299, 323-330):
// argument checks, target selection, the size assertion, the early-outs and
// render_par's band height.  `status` carries the error render() reports after
// its clears (size mismatch, zero band height); `render` says whether kernels
// run; clears that cannot be fused into the raster pass are flagged.
euchip_status plan_request(euchip_ctx *ctx, const euchip_pipeline_desc *d, const float *uniforms,
                           const euchip_mesh *mesh, euchip_buffer *pixel, euchip_buffer *depth,
                           const euchip_buffer *texture, const euchip_buffer *shadow, Plan &P) {
    P = Plan();
    memset(&P.v, 0, sizeof(P.v));
    if (!uniforms || !mesh) return fail(ctx, EUCHIP_INVALID_ARG, "null argument");
    if (mesh->ctx != ctx) return fail(ctx, EUCHIP_INVALID_ARG, "mesh belongs to another context");
    if (pixel && (!is_rgba8(pixel) || pixel->ctx != ctx)) return fail(ctx, EUCHIP_INVALID_ARG, "pixel target must be RGBA8");
    if (depth && (!is_f32(depth) || depth->ctx != ctx)) return fail(ctx, EUCHIP_INVALID_ARG, "depth target must be f32");
    const uint32_t sh = d->shader;
    if ((sh == EUCHIP_SHADER_TEXTURED || sh == EUCHIP_SHADER_TEX_SHADOW) &&
        (!is_rgba8(texture) || texture->w == 0 || texture->h == 0))
        return fail(ctx, EUCHIP_INVALID_ARG, "Linearly-interpolated texture cannot have no size");
    if ((sh == EUCHIP_SHADER_PHONG_SHADOW || sh == EUCHIP_SHADER_TEX_SHADOW) &&
        (!is_f32(shadow) || shadow->w == 0 || shadow->h == 0))
        return fail(ctx, EUCHIP_INVALID_ARG, "Linearly-interpolated texture cannot have no size");
    if ((sh == EUCHIP_SHADER_FLAT && !mesh->rgba8) || (sh == EUCHIP_SHADER_COLOR && !mesh->col) ||
        ((sh == EUCHIP_SHADER_GOURAUD || sh == EUCHIP_SHADER_PHONG || sh == EUCHIP_SHADER_PHONG_SHADOW ||
          sh == EUCHIP_SHADER_TEX_SHADOW) && !mesh->nrm) ||
        ((sh == EUCHIP_SHADER_TEXTURED || sh == EUCHIP_SHADER_TEX_SHADOW) && !mesh->uv))
        return fail(ctx, EUCHIP_INVALID_ARG, "mesh lacks an attribute the shader reads");

    const bool uses_depth = d->depth_test != EUCHIP_DEPTH_NONE || d->depth_write;
    const uint32_t pw = pixel ? pixel->w : 0, ph = pixel ? pixel->h : 0;
    const uint32_t dw = depth ? depth->w : 0, dh = depth ? depth->h : 0;
    uint32_t W = 0, H = 0;
    bool nothing = false;
    if (!d->pixel_write && !uses_depth) {
        nothing = true;
    } else if (d->pixel_write && !uses_depth) {
        W = pw, H = ph;
    } else if (!d->pixel_write && uses_depth) {
        W = dw, H = dh;
    } else if (pw != dw || ph != dh) {
        P.status = fail(ctx, EUCHIP_SIZE_MISMATCH, "Pixel target size is compatible with depth target size");
        nothing = true;
    } else {
        W = pw, H = ph;
    }
    const uint32_t L = d->aa_msaa ? std::min<uint32_t>(d->aa_level, 6u) : 0u;
    uint32_t G = H;
    if (!nothing && d->band_policy == EUCHIP_BANDS_PAR) {
        G = (uint32_t)((20000ull << L) / std::max<uint32_t>(W, 1u));
        if (G == 0) {
            P.status = fail(ctx, EUCHIP_BAND_ZERO, "attempt to divide by zero (render_par group_rows == 0)");
            nothing = true;
        } else if (std::min<uint32_t>(H / G, std::max<uint32_t>(d->threads, 1u)) == 0) {
            nothing = true;  // no worker threads spawned
        }
    }
    P.n_tris = stream_triangles(d, mesh->count, mesh->n_instances);
    P.n_sv = (uint64_t)mesh->n_instances * mesh->n_vertices;
    if (W == 0 || H == 0 || P.n_tris == 0) nothing = true;
    if (!nothing && (W >= (1u << 24) || H >= (1u << 24))) return fail(ctx, EUCHIP_INVALID_ARG, "target larger than 2^24");
    const bool fuse_color = !nothing && d->clear_color_enable && pixel && pixel->w == W && pixel->h == H;
    const bool fuse_depth = !nothing && d->clear_depth_enable && depth && depth->w == W && depth->h == H;
    P.clear_pixel = d->clear_color_enable && pixel && !fuse_color;
    P.clear_depth = d->clear_depth_enable && depth && !fuse_depth;
    P.render = !nothing;
    if (!P.render) return EUCHIP_OK;

    View &v = P.v;
    v.W = W, v.H = H, v.G = G;
    // Targets that are not the render size are never touched (Empty-like use).
    v.pixel = pixel && pixel->w == W && pixel->h == H ? (uint32_t *)pixel->ptr : nullptr;
    v.depth = depth && depth->w == W && depth->h == H ? (float *)depth->ptr : nullptr;
    if (uses_depth && !v.depth) return fail(ctx, EUCHIP_INVALID_ARG, "depth test without a depth target");
    v.clear_color_en = fuse_color ? 1u : 0u, v.clear_color = d->clear_color;
    v.clear_depth_en = fuse_depth ? 1u : 0u, v.clear_depth = d->clear_depth;
    v.ndc_x = 2.0f / (float)W;
    v.ndc_y = -2.0f / (float)H;
    v.pos = mesh->pos, v.nrm = mesh->nrm, v.col = mesh->col, v.uv = mesh->uv, v.rgba8 = mesh->rgba8;
    v.idx = mesh->idx, v.models = mesh->models;
    v.pos_comps = mesh->pos_comps, v.count = mesh->count, v.n_vertices = mesh->n_vertices;
    v.n_instances = mesh->n_instances, v.n_tris = (uint32_t)P.n_tris;
    if (texture) v.tex = (const uint32_t *)texture->ptr, v.tex_w = texture->w, v.tex_h = texture->h;
    if (shadow) v.shadow = (const float *)shadow->ptr, v.sh_w = shadow->w, v.sh_h = shadow->h;
    memcpy(v.u, uniforms, sizeof(v.u));
    return EUCHIP_OK;
}

// Tiles of a W x H target with tile_sx (tiles of 1 << sx by 256 >> sx pixels).
void tile_grid(View &v, uint32_t sx) {
    const uint32_t tw = 1u << sx, th = 256u >> sx;
    v.ntx = (v.W + tw - 1) / tw;
    v.nty = (v.H + th - 1) / th;
}

euchip_status enqueue_clears(euchip_ctx *ctx, const euchip_pipeline_desc *d, const Plan &P, euchip_buffer *pixel,
                             euchip_buffer *depth, hipStream_t s) {
    euchip_status st;
    if (P.clear_pixel && (st = clear_buffer(ctx, pixel, d->clear_color, s))) return st;
    if (P.clear_depth && (st = clear_buffer(ctx, depth, f32_bits(d->clear_depth), s))) return st;
    return EUCHIP_OK;
}

// Bin slots per tile of a frame: the capacity the reports asked for, within
// the frame's bin budget of max(64 MiB, 64 bytes per triangle); the entries of
// denser tiles go to their overflow segments.
uint32_t frame_tile_cap(uint32_t want, uint64_t ntiles, uint64_t n_tris) {
    const uint64_t budget = std::max<uint64_t>((uint64_t)64 << 20, 64 * n_tris);
    uint32_t c = 16;
    while (c < want && c < (1u << 24) && (uint64_t)(2 * c) * ntiles * 4 <= budget) c <<= 1;
    return c;
}

// Pool words a frame starts with before any report: two per triangle, at least 2 Mi.
size_t initial_pool_words(uint64_t n_tris) { return std::max<size_t>((size_t)2 << 20, 2 * n_tris); }

// Take the raster pass's latest report into account (host memory it writes;
// nothing waits for the stream): the bin capacity and the pool of the frames
// that follow grow to what that frame needed.
void absorb_report(euchip_ctx *ctx) {
    const volatile Report *r = ctx->report;
    const uint32_t f = r->frame;
    if (f == ctx->seen_frame) return;
    ctx->seen_frame = f;
    const uint32_t mc = r->max_count, used = r->pool_used;
    if (mc > ctx->tile_cap) {
        uint32_t c = ctx->tile_cap;
        while (c < mc && c < (1u << 24)) c <<= 1;
        ctx->tile_cap = c;
    }
    ctx->pool_want = std::max<size_t>(ctx->pool_want, (size_t)used + used / 4);
}

// render() of one request of lines (lines.rs:15-121) on the context's stream.
euchip_status render_lines(euchip_ctx *ctx, const euchip_pipeline_desc *d, const Plan &P, bool sync) {
    euchip_status s;
    FrameParams p;
    memset(&p, 0, sizeof(p));
    fill_pipe(p, d);
    p.v0 = P.v;
    p.n_views = 1;
    p.primitives = d->primitives;
    if (P.n_tris >= (1ull << 31)) return fail(ctx, EUCHIP_INVALID_ARG, "too many lines");
    p.n_lines = (uint32_t)P.n_tris;
    const size_t npx = (size_t)P.v.W * P.v.H;
    if ((s = ensure(ctx, ctx->vary, std::max<size_t>(P.n_sv * vd_count(d->shader), 1), "vertex data", ctx->stream)))
        return s;
    if ((s = ensure(ctx, ctx->lrec, std::max<size_t>((size_t)p.n_lines * kLineRecWords, 1), "line records",
                    ctx->stream)))
        return s;
    if ((s = ensure(ctx, ctx->lkey, npx, "line keys", ctx->stream))) return s;
    p.vary = ctx->vary.ptr;
    p.lrec = ctx->lrec.ptr;
    p.lkey = ctx->lkey.ptr;
    if ((s = timed(ctx, "lines", [&] { return launch_lines(p, ctx->stream); }))) return s;
    if (sync) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->profiling) resolve_pending(ctx);
    }
    return EUCHIP_OK;
}

// render() of one request on the context's stream.
euchip_status render_impl(euchip_ctx *ctx, const euchip_pipeline_desc *d, const float *uniforms,
                          const euchip_mesh *mesh, euchip_buffer *pixel, euchip_buffer *depth, bool sync) {
    if (!ctx || !d || !uniforms || !mesh) return fail(ctx, EUCHIP_INVALID_ARG, "null argument");
    euchip_status s = check_desc(ctx, d);
    if (s) return s;
    Plan P;
    if ((s = plan_request(ctx, d, uniforms, mesh, pixel, depth, d->texture, d->shadow, P))) return s;
    if ((s = enqueue_clears(ctx, d, P, pixel, depth, ctx->stream))) return s;
    if (!P.render) {
        if (sync) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return P.status;
    }
    reap_deferred(ctx, false);
    if (d->primitives != EUCHIP_PRIMS_TRIANGLE_LIST) return render_lines(ctx, d, P, sync);
    absorb_report(ctx);
    FrameParams p;
    memset(&p, 0, sizeof(p));
    fill_pipe(p, d);
    p.tile_sx = 5;
    p.v0 = P.v;
    tile_grid(p.v0, p.tile_sx);
    // big-triangle mode when the mesh has no more triangles than the target has
    // 32 x 8 tiles (triangles then tend to cover many tiles each); its tiles are
    // wider (fewer tiles per row walk the same row start)
    const bool big_mode = P.n_tris <= (uint64_t)p.v0.ntx * p.v0.nty;
    if (big_mode && ctx->big_tile_sx != p.tile_sx) {
        p.tile_sx = ctx->big_tile_sx;
        tile_grid(p.v0, p.tile_sx);
    }
    const size_t ntiles = (size_t)p.v0.ntx * p.v0.nty;
    p.ntiles = (uint32_t)ntiles;
    p.n_views = 1;

    const size_t nsv = P.n_sv;
    if ((s = ensure(ctx, ctx->rec0, std::max<size_t>(P.n_tris * kRec0Quads, 1), "setup records", ctx->stream))) return s;
    if ((s = ensure(ctx, ctx->rec1, std::max<size_t>(P.n_tris * kRec1Quads, 1), "setup records", ctx->stream))) return s;
    if ((s = ensure(ctx, ctx->vary, std::max<size_t>(nsv * vd_count(d->shader), 1), "vertex data", ctx->stream)))
        return s;
    // two halves: setup counts into one while the raster pass zeroes the other
    const bool fresh_tiles = ctx->tile_count.cap < 2 * ntiles;
    if ((s = ensure(ctx, ctx->tile_count, 2 * ntiles, "tile counts", ctx->stream))) return s;
    if (fresh_tiles) {  // new halves: the checkpoint list counts start over with them
        HIP_TRY(ctx, hipMemsetAsync(ctx->tile_count.ptr, 0, ctx->tile_count.cap * sizeof(TileBin), ctx->stream));
        ctx->counts_used[0] = ctx->counts_used[1] = 0;
        if (ctx->ck_count.cap)
            HIP_TRY(ctx, hipMemsetAsync(ctx->ck_count.ptr, 0, ctx->ck_count.cap * sizeof(uint2), ctx->stream));
        ctx->ck_dirty[0] = ctx->ck_dirty[1] = 0;
    }
    const size_t half = ctx->tile_count.cap / 2;
    if ((s = ensure(ctx, ctx->heavy, ntiles, "heavy tiles", ctx->stream))) return s;
    if ((s = ensure(ctx, ctx->busy, 8 * ntiles, "busy tiles", ctx->stream))) return s;
    const size_t big_cap = std::min<size_t>(std::max<size_t>(P.n_tris, 1), kBigCap);
    if ((s = ensure(ctx, ctx->big, big_cap, "big triangles", ctx->stream))) return s;
    // Row checkpoints (big-triangle frames): optional, so a pool the budget
    // cannot hold leaves the big triangles on the big list, walked from the row
    // starts.  Tiles at most 8 rows high (an 8-bit row mask per tile and
    // triangle), at most 64 MB of tile entries and 1 GiB of checkpoints (one
    // slot per triangle, tile column and row), and the span blocks' LDS
    // tables (4 waves, span_table_bytes each) within 64 KB.
    bool ck_on = false;
    const size_t mstride = ((size_t)P.n_tris + 3) & ~(size_t)3;
    const size_t ck_slots = (size_t)P.n_tris * p.v0.ntx * p.v0.H;
    if (big_mode && p.tile_sx >= 5 && ntiles * mstride <= ((size_t)8 << 20) && ck_slots <= kCkMaxSlots &&
        4u * (size_t)span_table_bytes(8u - p.tile_sx, p.v0.ntx) <= ((size_t)64 << 10)) {
        const size_t counts_had = ctx->ck_count.cap;
        ck_on = ensure(ctx, ctx->big_rec, (size_t)P.n_tris * kBigRecQuads, "big triangles", ctx->stream) == EUCHIP_OK &&
                ensure(ctx, ctx->ck, ck_slots, "row checkpoints", ctx->stream) == EUCHIP_OK &&
                ensure(ctx, ctx->cktab, ntiles * mstride, "checkpoint lists", ctx->stream) == EUCHIP_OK &&
                ensure(ctx, ctx->ck_count, ctx->tile_count.cap, "checkpoint lists", ctx->stream) == EUCHIP_OK;
        if (ck_on && ctx->ck_count.cap != counts_had) {  // both halves start at zero
            HIP_TRY(ctx, hipMemsetAsync(ctx->ck_count.ptr, 0, ctx->ck_count.cap * sizeof(uint2), ctx->stream));
            ctx->ck_dirty[0] = ctx->ck_dirty[1] = 0;
        }
    }

    for (int attempt = 0;; attempt++) {
        const uint32_t cap = frame_tile_cap(ctx->tile_cap, ntiles, P.n_tris);
        if ((s = ensure(ctx, ctx->bins, ntiles * cap, "bins", ctx->stream))) return s;
        const size_t pool_words = std::min<size_t>(std::max(ctx->pool_want, initial_pool_words(P.n_tris)), 0xFFFFFFF0ull);
        if ((s = ensure(ctx, ctx->pool, pool_words, "overflow pool", ctx->stream))) return s;
        p.rec0 = ctx->rec0.ptr;
        p.rec1 = ctx->rec1.ptr;
        p.vary = ctx->vary.ptr;
        p.ctr = ctx->counters;
        p.report = ctx->report_dev;
        const uint32_t par = ctx->frame & 1u, npar = par ^ 1u;
        p.tile_count = ctx->tile_count.ptr + par * half;
        p.count_next = ctx->tile_count.ptr + npar * half;
        p.zero_n = ctx->counts_used[npar];
        // Checkpoint list counts: a checkpoint frame counts into its half, which
        // must be zero.  The raster of a big-triangle frame zeroes the other half
        // (left by the frame before); a half still left after other frames is
        // cleared here first.
        if (ck_on) {
            p.ck_count = ctx->ck_count.ptr + par * half;
            if (ctx->ck_dirty[par])
                HIP_TRY(ctx, launch_fill_u32((uint32_t *)p.ck_count, 0u, 2u * (size_t)ctx->ck_dirty[par], ctx->stream));
            ctx->ck_dirty[par] = 0;
        }
        p.ck_count_next = nullptr, p.ck_zero_n = 0u;
        if (big_mode && ctx->ck_dirty[npar]) {  // the raster is the big-triangle kernel
            p.ck_count_next = ctx->ck_count.ptr + npar * half;
            p.ck_zero_n = ctx->ck_dirty[npar];
        }
        p.busy = ctx->busy.ptr;
        p.busy_map = BUSY_BLOCKS;
        p.big = ctx->big.ptr;
        p.big_cap = big_mode ? (uint32_t)std::min<size_t>(ctx->big.cap, kBigCap) : 0u;
        if (ck_on) {
            p.big_rec = ctx->big_rec.ptr;
            p.ck = ctx->ck.ptr;
            p.ck_on = 1u;
            p.cktab = ctx->cktab.ptr;
            p.mstride = (uint32_t)mstride;
        }
        p.bins = ctx->bins.ptr;
        p.v0.tile_cap = cap;
        p.v0.bin_base = 0;
        p.pool = ctx->pool.ptr;
        p.pool_cap = (uint32_t)std::min<size_t>(ctx->pool.cap, 0xFFFFFFF0ull);
        p.heavy = ctx->heavy.ptr;
        p.heavy_cap = (uint32_t)std::min<size_t>(ctx->heavy.cap, 0xFFFFFFF0ull);
        p.frame = ctx->frame++;
        if ((s = timed(ctx, "setup", [&] { return launch_setup(p, ctx->stream, false, 0); }))) return s;
        if ((s = timed(ctx, "raster", [&] { return launch_raster(p, ctx->stream, false, ctx->max_groups); }))) return s;
        ctx->counts_used[npar] = 0;
        ctx->counts_used[par] = (uint32_t)ntiles;
        if (p.ck_zero_n) ctx->ck_dirty[npar] = 0;
        if (ck_on) ctx->ck_dirty[par] = (uint32_t)ntiles;
        if (!sync) return EUCHIP_OK;
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->profiling) resolve_pending(ctx);
        absorb_report(ctx);  // this frame's
        const volatile Report *r = ctx->report;
        if (!(r->flags & FLAG_BIN_OVERFLOW)) return EUCHIP_OK;
        // The pool ran out: nothing was written.  Re-run with the pool the
        // frame asked for (absorb_report raised pool_want to it).
        ctx->seen_dropped = r->dropped;
        if (attempt >= 3) return fail(ctx, EUCHIP_OVER_BUDGET, "bin storage overflow");
    }
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------

uint32_t pow2_at_least(uint64_t n) {
    uint32_t c = 16;
    while (c < n && c < (1u << 24)) c <<= 1;
    return c;
}

// First guess of a request's bin slots per tile: eight times its triangles
// spread evenly over its tiles (an object covers a fraction of the target, and
// entries past the capacity go to the slow overflow list), at least 16, at most
// all of them, and no more than 8 Mi slots for the request.
uint32_t initial_cap(uint64_t n_tris, uint64_t ntiles) {
    const uint64_t even = ntiles ? (8 * n_tris + ntiles - 1) / ntiles : n_tris;
    uint32_t cap = pow2_at_least(std::min<uint64_t>(std::max<uint64_t>(even, 16), std::max<uint64_t>(n_tris, 16)));
    while (cap > 16 && (uint64_t)cap * ntiles > ((uint64_t)8 << 20)) cap >>= 1;
    return cap;
}

void drop_graph(euchip_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Lay the active requests out in the launch's tile, vertex and bin spaces for
// the current tile width and upload the tables.
euchip_status build_tables(euchip_batch *b) {
    euchip_ctx *ctx = b->ctx;
    const uint32_t sx = b->tile_sx;
    std::vector<uint32_t> &caps = b->caps[sx - 3];
    const size_t na = b->active.size();
    if (caps.size() < na) caps.resize(na, 0);
    // sizes first: the tables are written straight into pinned staging memory
    uint64_t tiles = 0, nblk = 0, bins = 0, sv = 0, tris = 0;
    for (size_t k = 0; k < na; k++) {
        const Plan &P = b->plans[b->active[k]];
        View v = P.v;
        tile_grid(v, sx);
        const uint64_t nt = (uint64_t)v.ntx * v.nty;
        if (caps[k] == 0) caps[k] = initial_cap(P.n_tris, nt);
        tiles += nt;
        nblk += (std::max<uint64_t>(P.n_sv, P.n_tris) + 255) / 256;
        bins += nt * caps[k];
        sv += P.n_sv;
        tris += P.n_tris;
    }
    if (tiles >= (1ull << 32) || sv >= (1ull << 32) || tris >= (1ull << 32) || nblk >= (1ull << 31))
        return fail(ctx, EUCHIP_INVALID_ARG, "batch too large for 32-bit tile / vertex ids");
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t off_blk = align(na * sizeof(View));
    const size_t off_tile = off_blk + align(nblk * sizeof(uint2));
    const size_t table_bytes = std::max<size_t>(off_tile + tiles * 4, 16);
    if (b->stage_bytes < table_bytes) {
        if (b->stage) (void)hipHostFree(b->stage);
        b->stage = nullptr;
        b->stage_bytes = 0;
        const size_t want = table_bytes + table_bytes / 2;
        HIP_TRY(ctx, hipHostMalloc((void **)&b->stage, want, hipHostMallocDefault));
        b->stage_bytes = want;
    }
    View *hv = reinterpret_cast<View *>(b->stage);
    uint2 *hblk = reinterpret_cast<uint2 *>(b->stage + off_blk);
    uint32_t *htile = reinterpret_cast<uint32_t *>(b->stage + off_tile);
    tiles = bins = sv = tris = 0;
    uint64_t blk = 0;
    for (size_t k = 0; k < na; k++) {
        const Plan &P = b->plans[b->active[k]];
        View v = P.v;
        tile_grid(v, sx);
        const uint64_t nt = (uint64_t)v.ntx * v.nty;
        v.tile_base = (uint32_t)tiles;
        v.sv_base = (uint32_t)sv;
        v.tri_base = (uint32_t)tris;
        v.tile_cap = caps[k];
        v.bin_base = bins;
        hv[k] = v;
        const uint64_t nb = (std::max<uint64_t>(P.n_sv, P.n_tris) + 255) / 256;
        for (uint64_t j = 0; j < nb; j++) hblk[blk++] = make_uint2((uint32_t)k, (uint32_t)j);
        std::fill(htile + tiles, htile + tiles + nt, (uint32_t)k);
        tiles += nt;
        bins += nt * caps[k];
        sv += P.n_sv;
        tris += P.n_tris;
    }
    euchip_status s;
    const uint32_t vdc = vd_count(b->desc.shader);
    if ((s = ensure(ctx, b->tables, (table_bytes + 15) / 16, "batch tables", b->stream))) return s;
    if ((s = ensure(ctx, b->rec0, std::max<size_t>(tris * kRec0Quads, 1), "batch setup records", b->stream))) return s;
    if ((s = ensure(ctx, b->rec1, std::max<size_t>(tris * kRec1Quads, 1), "batch setup records", b->stream))) return s;
    if ((s = ensure(ctx, b->vary, std::max<size_t>(sv * vdc, 1), "batch vertex data", b->stream))) return s;
    if ((s = ensure(ctx, b->tile_count, std::max<size_t>(tiles, 1), "batch tile counts", b->stream))) return s;
    if ((s = ensure(ctx, b->busy, std::max<size_t>(tiles, 1), "batch busy tiles", b->stream))) return s;
    b->n_tiles = tiles;
    if ((s = ensure(ctx, b->bins, std::max<size_t>(bins, 1), "batch bins", b->stream))) return s;
    if ((s = ensure(ctx, b->pool, std::max(b->pool_want, initial_pool_words(tris)), "batch overflow pool", b->stream)))
        return s;
    const size_t words = sizeof(Counters) / 4 + na;
    if ((s = ensure(ctx, b->ctrbuf, words, "batch counters", b->stream))) return s;
    if (b->host_ctr_words < words) {
        if (b->host_ctr) (void)hipHostFree(b->host_ctr);
        b->host_ctr = nullptr;
        b->host_ctr_words = 0;
        HIP_TRY(ctx, hipHostMalloc((void **)&b->host_ctr, words * 4, hipHostMallocDefault));
        b->host_ctr_words = words;
    }
    memset(b->host_ctr, 0, words * 4);
    // One copy from pinned memory; the staging area is rewritten only when the
    // batch's stream is idle again (set, set_tile_width and wait synchronise it).
    HIP_TRY(ctx, hipMemcpyAsync(b->tables.ptr, b->stage, table_bytes, hipMemcpyHostToDevice, b->stream));
    const uint8_t *dev = reinterpret_cast<const uint8_t *>(b->tables.ptr);
    const View *d_views = reinterpret_cast<const View *>(dev);
    const uint2 *d_blk = reinterpret_cast<const uint2 *>(dev + off_blk);
    const uint32_t *d_tile = reinterpret_cast<const uint32_t *>(dev + off_tile);

    FrameParams &p = b->p;
    memset(&p, 0, sizeof(p));
    fill_pipe(p, &b->desc);
    p.rec0 = b->rec0.ptr;
    p.rec1 = b->rec1.ptr;
    p.vary = b->vary.ptr;
    p.ctr = (Counters *)b->ctrbuf.ptr;
    p.req_max = b->ctrbuf.ptr + sizeof(Counters) / 4;
    p.tile_count = b->tile_count.ptr;
    p.bins = b->bins.ptr;
    p.pool = b->pool.ptr;
    p.pool_cap = (uint32_t)std::min<size_t>(b->pool.cap, 0xFFFFFFF0ull);
    p.report = nullptr;
    p.heavy = nullptr;  // batches keep their tile order: most of their tiles would be heavy
    p.heavy_cap = 0;
    p.busy = b->busy.ptr;
    p.busy_map = BUSY_RANGES;
    p.big = nullptr;
    p.big_cap = 0;
    p.count_next = nullptr;
    p.zero_n = 0;

    p.frame = 0;
    p.ntiles = (uint32_t)tiles;
    p.tile_sx = sx;
    p.n_views = (uint32_t)na;
    p.views = d_views;
    p.blk = d_blk;
    p.tile_req = d_tile;
    b->setup_blocks = (uint32_t)nblk;
    b->tiles_total = tiles;
    b->bins_total = bins;
    b->sv_total = sv;
    b->tables_dirty = false;
    drop_graph(b);
    return EUCHIP_OK;
}

// The work of one launch, enqueued on the batch's stream (also what a graph
// captures): leftover clears, counter reset, setup, raster, counter read-back.
// Eager launches of a profiling context time setup and raster with events.
euchip_status enqueue_batch(euchip_batch *b, bool capturing = false) {
    euchip_ctx *ctx = b->ctx;
    for (size_t i = 0; i < b->reqs.size(); i++) {
        const Plan &P = b->plans[i];
        if (P.clear_pixel) HIP_TRY(ctx, launch_fill_u32((uint32_t *)b->reqs[i].pixel->ptr, b->desc.clear_color,
                                                        (size_t)b->reqs[i].pixel->w * b->reqs[i].pixel->h, b->stream));
        if (P.clear_depth) HIP_TRY(ctx, launch_fill_u32((uint32_t *)b->reqs[i].depth->ptr, f32_bits(b->desc.clear_depth),
                                                        (size_t)b->reqs[i].depth->w * b->reqs[i].depth->h, b->stream));
    }
    if (b->active.empty()) return EUCHIP_OK;
    const size_t words = sizeof(Counters) / 4 + b->active.size();
    HIP_TRY(ctx, launch_fill_u32(b->ctrbuf.ptr, 0u, words, b->stream));
    HIP_TRY(ctx, launch_fill_u32(reinterpret_cast<uint32_t *>(b->tile_count.ptr), 0u, 2 * b->n_tiles,
                                 b->stream));  // raster leaves them as it read them
    euchip_status s;
    if (ctx->profiling && !capturing) {
        if ((s = timed_on(ctx, b->stream, "batch_setup", [&] { return launch_setup(b->p, b->stream, true, b->setup_blocks); })))
            return s;
        if ((s = timed_on(ctx, b->stream, "batch_raster", [&] { return launch_raster(b->p, b->stream, true, ctx->max_groups); })))
            return s;
    } else {
        HIP_TRY(ctx, launch_setup(b->p, b->stream, true, b->setup_blocks));
        HIP_TRY(ctx, launch_raster(b->p, b->stream, true, ctx->max_groups));
    }
    HIP_TRY(ctx, hipMemcpyAsync(b->host_ctr, b->ctrbuf.ptr, words * 4, hipMemcpyDeviceToHost, b->stream));
    return EUCHIP_OK;
}

euchip_status launch_batch(euchip_batch *b, bool use_graph) {
    euchip_ctx *ctx = b->ctx;
    euchip_status s;
    if (b->tables_dirty && (s = build_tables(b))) return s;
    if (use_graph) {
        if (!b->exec) {
            HIP_TRY(ctx, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            s = enqueue_batch(b, true);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (s) {
                if (g) (void)hipGraphDestroy(g);
                return s;
            }
            if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
            b->graph = g;
            HIP_TRY(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        HIP_TRY(ctx, hipGraphLaunch(b->exec, b->stream));
    } else if ((s = enqueue_batch(b))) {
        return s;
    }
    HIP_TRY(ctx, hipEventRecord(b->done, b->stream));
    b->launched = true;
    return EUCHIP_OK;
}

// Wait for the last launch; grow bins that overflowed; re-run a launch whose
// overflow list filled up (its targets were not written).
euchip_status wait_batch(euchip_batch *b) {
    euchip_ctx *ctx = b->ctx;
    for (int attempt = 0;; attempt++) {
        HIP_TRY(ctx, hipStreamSynchronize(b->stream));
        if (ctx->profiling) {
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            resolve_pending(ctx);
        }
        if (!b->launched || b->active.empty()) return EUCHIP_OK;
        const Counters &hc = *(const Counters *)b->host_ctr;
        const uint32_t *req_max = b->host_ctr + sizeof(Counters) / 4;
        std::vector<uint32_t> &caps = b->caps[b->tile_sx - 3];
        for (size_t k = 0; k < b->active.size(); k++) {
            if (req_max[k] > caps[k]) {
                const Plan &P = b->plans[b->active[k]];
                View v = P.v;
                tile_grid(v, b->tile_sx);
                const uint32_t c = frame_tile_cap(req_max[k], (uint64_t)v.ntx * v.nty, P.n_tris);
                if (c > caps[k]) caps[k] = c, b->tables_dirty = true;
            }
        }
        const size_t asked = hc.pool[0][0];
        b->pool_want = std::max<size_t>(b->pool_want, asked + asked / 4);
        if (!(hc.flags[0][0] & FLAG_BIN_OVERFLOW)) return EUCHIP_OK;
        // the pool ran out: nothing was written; re-run with what it asked for
        if (attempt >= 3) return fail(ctx, EUCHIP_OVER_BUDGET, "batch bin storage overflow");
        b->tables_dirty = true;
        euchip_status s;
        if ((s = launch_batch(b, false))) return s;
    }
}

void free_batch_memory(euchip_batch *b) {
    euchip_ctx *ctx = b->ctx;
    drop_graph(b);
    release(ctx, b->tables);
    if (b->stage) (void)hipHostFree(b->stage);
    b->stage = nullptr;
    b->stage_bytes = 0;
    release(ctx, b->rec0);
    release(ctx, b->rec1);
    release(ctx, b->vary);
    release(ctx, b->tile_count);
    release(ctx, b->busy);
    release(ctx, b->bins);
    release(ctx, b->ctrbuf);
    release(ctx, b->pool);
    if (b->host_ctr) (void)hipHostFree(b->host_ctr);
    b->host_ctr = nullptr;
    b->host_ctr_words = 0;
}

}  // namespace

namespace euchip {

size_t request_scratch_estimate(const euchip_pipeline_desc *d, const euchip_mesh *m, uint32_t W, uint32_t H) {
    const uint64_t n_tris = stream_triangles(d, m->count, m->n_instances);
    const uint64_t n_sv = (uint64_t)m->n_instances * m->n_vertices;
    const uint64_t nt = (uint64_t)((W + 31) / 32) * ((H + 7) / 8);
    const uint64_t nb = (std::max<uint64_t>(n_sv, n_tris) + 255) / 256;
    return sizeof(View) + nb * sizeof(uint2) + nt * 16 + n_sv * 4 * vd_count(d->shader) +
           n_tris * 16 * (kRec0Quads + kRec1Quads) + nt * initial_cap(n_tris, nt) * 4 + 4 + 8 * n_tris;
}

size_t batch_fixed_bytes() { return sizeof(Counters) + 8 * ((size_t)1 << 20); }

void set_last_error(euchip_ctx *ctx, const char *msg) {
    if (ctx) ctx->last_error = msg;
}

}  // namespace euchip

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
    if (const char *e = getenv("EUCHIP_BIG_TILE_SX")) {  // tuning: tile shape of big-triangle frames
        const int v = atoi(e);
        if (v >= 3 && v <= 8) ctx->big_tile_sx = (uint32_t)v;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->max_groups = (uint32_t)cus * 8u;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&ctx->counters, sizeof(Counters)) != hipSuccess ||
        hipHostMalloc((void **)&ctx->report, sizeof(Report), hipHostMallocDefault) != hipSuccess ||
        hipHostGetDevicePointer((void **)&ctx->report_dev, ctx->report, 0) != hipSuccess ||
        hipMemset(ctx->counters, 0, sizeof(Counters)) != hipSuccess) {
        euchip_ctx_destroy(ctx);
        return EUCHIP_HIP_ERROR;
    }
    memset(ctx->report, 0, sizeof(Report));
    ctx->report->frame = ~0u;  // no frame reported yet
    if (hipHostMalloc((void **)&ctx->na_readback, 256, hipHostMallocDefault) != hipSuccess) {
        euchip_ctx_destroy(ctx);
        return EUCHIP_HIP_ERROR;
    }
    // the no-allocation scratch of this device (current above); a failed lookup
    // only disables euchip_render_noalloc
    ctx->noalloc_err = noalloc_scratch(device, &ctx->noalloc);
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
    release(ctx, ctx->rec0);
    release(ctx, ctx->rec1);
    release(ctx, ctx->vary);
    release(ctx, ctx->tile_count);
    release(ctx, ctx->bins);
    release(ctx, ctx->pool);
    release(ctx, ctx->heavy);
    release(ctx, ctx->busy);
    release(ctx, ctx->big);
    release(ctx, ctx->big_rec);
    release(ctx, ctx->ck);
    release(ctx, ctx->cktab);
    release(ctx, ctx->ck_count);
    release(ctx, ctx->lrec);
    release(ctx, ctx->lkey);

    reap_deferred(ctx, true);
    (void)hipFree(ctx->counters);
    if (ctx->report) (void)hipHostFree(ctx->report);
    if (ctx->na_readback) (void)hipHostFree(ctx->na_readback);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return EUCHIP_OK;
}

const char *euchip_last_error(const euchip_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

void *euchip_ctx_stream(euchip_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

euchip_status euchip_synchronize(euchip_ctx *ctx) {
    if (!ctx) return EUCHIP_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->profiling) resolve_pending(ctx);
    reap_deferred(ctx, false);
    absorb_report(ctx);
    const uint32_t dropped = ((const volatile Report *)ctx->report)->dropped;
    if (dropped != ctx->seen_dropped) {
        const uint32_t n = dropped - ctx->seen_dropped;
        ctx->seen_dropped = dropped;
        return fail(ctx, EUCHIP_OVER_BUDGET,
                    "%u asynchronous frame(s) ran out of bin overflow storage and were not written; the frames "
                    "enqueued after this call use storage of the size they asked for", n);
    }
    return EUCHIP_OK;
}

size_t euchip_ctx_bytes_in_use(const euchip_ctx *ctx) { return ctx ? ctx->bytes_in_use : 0; }

euchip_status euchip_ctx_set_budget(euchip_ctx *ctx, size_t bytes) {
    if (!ctx) return EUCHIP_INVALID_ARG;
    ctx->budget = bytes;
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
    euchip_status s = dev_alloc(ctx, &ptr, (size_t)w * h * 4, "buffer");
    if (s) return s;
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
        (void)hipDeviceSynchronize();  // batches on other streams may still use it
        dev_free(b->ctx, b->ptr, (size_t)b->w * b->h * 4);
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
    return clear_buffer(b->ctx, b, bits, b->ctx->stream);
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
    auto up = [&](void **dst, const void *src, size_t bytes) -> euchip_status {
        if (!src || !bytes) return EUCHIP_OK;
        euchip_status st = dev_alloc(ctx, dst, bytes, "mesh");
        if (st) return st;
        m->bytes += bytes;
        hipError_t e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        return e == hipSuccess ? EUCHIP_OK : hip_fail(ctx, e, "mesh upload");
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
    euchip_status e;
    if ((e = up((void **)&m->pos, d->pos, nv * d->pos_comps * 4)) || (e = up((void **)&m->nrm, d->nrm, nv * 12)) ||
        (e = up((void **)&m->col, d->col, nv * 16)) || (e = up((void **)&m->uv, d->uv, nv * 8)) ||
        (e = up((void **)&m->rgba8, d->rgba8, nv * 4)) || (e = up((void **)&m->idx, d->idx, d->count * 4)) ||
        (e = up((void **)&m->models, models, (size_t)16 * 4 * d->n_instances))) {
        std::string msg = ctx->last_error;
        euchip_mesh_destroy(m);
        ctx->last_error = msg;
        return e;
    }
    *out = m;
    return EUCHIP_OK;
}

euchip_status euchip_mesh_destroy(euchip_mesh *m) {
    if (!m) return EUCHIP_INVALID_ARG;
    (void)hipDeviceSynchronize();  // batches on other streams may still use it
    for (void *ptr : {(void *)m->pos, (void *)m->nrm, (void *)m->col, (void *)m->uv, (void *)m->rgba8, (void *)m->idx,
                      (void *)m->models})
        if (ptr) (void)hipFree(ptr);
    m->ctx->bytes_in_use -= std::min(m->bytes, m->ctx->bytes_in_use);
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

euchip_status euchip_render_noalloc(euchip_ctx *ctx, const euchip_pipeline_desc *d, const float *uniforms,
                                    const euchip_mesh *mesh, euchip_buffer *pixel, euchip_buffer *depth) {
    if (!ctx || !d || !uniforms || !mesh) return fail(ctx, EUCHIP_INVALID_ARG, "null argument");
    (void)hipSetDevice(ctx->device);
    if (!ctx->noalloc)
        return fail(ctx, EUCHIP_HIP_ERROR, "no-allocation scratch unavailable: %s", hipGetErrorString(ctx->noalloc_err));
    euchip_status s = check_desc(ctx, d);
    if (!s) s = check_triangles(ctx, d);
    if (s) return s;
    Plan P;
    if ((s = plan_request(ctx, d, uniforms, mesh, pixel, depth, d->texture, d->shadow, P))) return s;
    FrameParams p;
    memset(&p, 0, sizeof(p));
    size_t ntiles = 0;
    if (P.render) {
        fill_pipe(p, d);
        p.tile_sx = 5;
        p.v0 = P.v;
        tile_grid(p.v0, p.tile_sx);
        ntiles = (size_t)p.v0.ntx * p.v0.nty;
        // the limits of the scratch layout, checked before anything is enqueued
        if (P.n_tris > kNaMaxTris)
            return fail(ctx, EUCHIP_OVER_BUDGET, "no-allocation render: %llu triangles (at most %u)",
                        (unsigned long long)P.n_tris, kNaMaxTris);
        if (P.n_sv * vd_count(d->shader) > kNaMaxVaryFloats)
            return fail(ctx, EUCHIP_OVER_BUDGET, "no-allocation render: %llu shaded vertices of %u floats (at most %u floats)",
                        (unsigned long long)P.n_sv, vd_count(d->shader), kNaMaxVaryFloats);
        if (ntiles > kNaMaxTiles)
            return fail(ctx, EUCHIP_OVER_BUDGET, "no-allocation render: %zu tiles of 32 x 8 pixels (at most %u)", ntiles,
                        kNaMaxTiles);
    }
    const hipStream_t st = ctx->stream;
    if (P.clear_pixel)
        HIP_TRY(ctx, launch_fill_u32((uint32_t *)pixel->ptr, d->clear_color, (size_t)pixel->w * pixel->h, st));
    if (P.clear_depth)
        HIP_TRY(ctx, launch_fill_u32((uint32_t *)depth->ptr, f32_bits(d->clear_depth), (size_t)depth->w * depth->h, st));
    if (!P.render) {
        HIP_TRY(ctx, hipStreamSynchronize(st));
        return P.status;
    }
    uint8_t *base = static_cast<uint8_t *>(ctx->noalloc);
    p.n_views = 1;
    p.ntiles = (uint32_t)ntiles;
    p.rec0 = reinterpret_cast<uint4 *>(base + kNaRec0);
    p.rec1 = reinterpret_cast<uint4 *>(base + kNaRec1);
    p.vary = reinterpret_cast<float *>(base + kNaVary);
    p.ctr = reinterpret_cast<Counters *>(base + kNaCtr);
    p.tile_count = reinterpret_cast<TileBin *>(base + kNaTiles);
    p.count_next = nullptr;
    p.zero_n = 0;
    p.busy = reinterpret_cast<uint32_t *>(base + kNaBusy);
    p.busy_map = BUSY_BLOCKS;
    p.heavy = reinterpret_cast<uint32_t *>(base + kNaHeavy);
    p.heavy_cap = kNaMaxTiles;
    p.big = reinterpret_cast<uint32_t *>(base + kNaBig);
    p.big_cap = P.n_tris <= ntiles ? kBigCap : 0u;
    p.bins = reinterpret_cast<uint32_t *>(base + kNaBins);
    p.v0.tile_cap = (uint32_t)std::min<size_t>(kNaBinSlots / ntiles, (size_t)1 << 24);
    p.v0.bin_base = 0;
    p.pool = reinterpret_cast<uint32_t *>(base + kNaPool);
    p.pool_cap = (uint32_t)kNaPoolWords;
    p.report = nullptr;
    p.frame = 0;
    // the array keeps the last call's contents: counters and tile bins start at
    // zero (fill kernels: no memset nodes)
    HIP_TRY(ctx, launch_fill_u32(reinterpret_cast<uint32_t *>(p.ctr), 0u, sizeof(Counters) / 4, st));
    HIP_TRY(ctx, launch_fill_u32(reinterpret_cast<uint32_t *>(p.tile_count), 0u, 2 * ntiles, st));
    HIP_TRY(ctx, launch_setup(p, st, false, 0));
    HIP_TRY(ctx, launch_raster(p, st, false, ctx->max_groups));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->na_readback, &p.ctr->flags[0][0], 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->na_readback + 1, &p.ctr->pool[0][0], 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    if (ctx->na_readback[0] & FLAG_BIN_OVERFLOW)
        return fail(ctx, EUCHIP_OVER_BUDGET,
                    "no-allocation render: the frame's bin overflow needs %u words, the scratch holds %zu; its "
                    "targets were not written", ctx->na_readback[1], kNaPoolWords);
    return EUCHIP_OK;
}

euchip_status euchip_batch_create(euchip_ctx *ctx, euchip_batch **out) {
    if (!ctx || !out) return fail(ctx, EUCHIP_INVALID_ARG, "null argument");
    (void)hipSetDevice(ctx->device);
    *out = nullptr;
    euchip_batch *b = new euchip_batch();
    b->ctx = ctx;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return fail(ctx, EUCHIP_HIP_ERROR, "batch stream");
    }
    *out = b;
    return EUCHIP_OK;
}

// Destroy the parts of a multi-desc batch.
void drop_parts(euchip_batch *b) {
    for (euchip_batch *q : b->parts) euchip_batch_destroy(q);
    b->parts.clear();
}

// Two descs that render identically in one launch (every field equal).
bool same_desc(const euchip_pipeline_desc &a, const euchip_pipeline_desc &b) {
    return a.shader == b.shader && a.xform == b.xform && a.pixel_write == b.pixel_write &&
           a.depth_test == b.depth_test && a.depth_write == b.depth_write && a.y_up == b.y_up &&
           a.z_clip == b.z_clip && (!a.z_clip || (a.z_min == b.z_min && a.z_max == b.z_max)) && a.cull == b.cull &&
           a.aa_msaa == b.aa_msaa && a.aa_level == b.aa_level && a.band_policy == b.band_policy &&
           a.threads == b.threads && a.clear_color_enable == b.clear_color_enable &&
           a.clear_color == b.clear_color && a.clear_depth_enable == b.clear_depth_enable &&
           (a.clear_depth == b.clear_depth || (a.clear_depth != a.clear_depth && b.clear_depth != b.clear_depth)) &&
           a.texture == b.texture && a.shadow == b.shadow && a.pixel_format == b.pixel_format &&
           a.sampler == b.sampler && a.geometry == b.geometry && a.primitives == b.primitives;
}

euchip_status euchip_batch_destroy(euchip_batch *b) {
    if (!b) return EUCHIP_INVALID_ARG;
    drop_parts(b);
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->stream);
    free_batch_memory(b);
    (void)hipEventDestroy(b->done);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return EUCHIP_OK;
}

euchip_status euchip_batch_set(euchip_batch *b, const euchip_pipeline_desc *desc, const euchip_request *reqs,
                               uint32_t n) {
    if (!b || !desc || (n && !reqs)) return fail(b ? b->ctx : nullptr, EUCHIP_INVALID_ARG, "null argument");
    euchip_ctx *ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    euchip_status s = check_desc(ctx, desc);
    if (!s) s = check_triangles(ctx, desc);
    if (s) return s;
    HIP_TRY(ctx, hipStreamSynchronize(b->stream));
    drop_parts(b);
    std::vector<Plan> plans(n);
    std::vector<std::array<float, kUniforms>> uni(n);
    std::vector<uint32_t> active;
    for (uint32_t i = 0; i < n; i++) {
        const euchip_request &r = reqs[i];
        if (!r.uniforms || !r.mesh) return fail(ctx, EUCHIP_INVALID_ARG, "request %u: null mesh or uniforms", i);
        memcpy(uni[i].data(), r.uniforms, sizeof(float) * kUniforms);
        const euchip_buffer *tex = r.texture ? r.texture : desc->texture;
        const euchip_buffer *shw = r.shadow ? r.shadow : desc->shadow;
        if ((s = plan_request(ctx, desc, uni[i].data(), r.mesh, r.pixel, r.depth, tex, shw, plans[i])) ||
            (s = plans[i].status)) {
            ctx->last_error = "request " + std::to_string(i) + ": " + ctx->last_error;
            return s;
        }
        if (plans[i].render) active.push_back(i);
    }
    b->desc = *desc;
    b->reqs.assign(reqs, reqs + n);
    b->uniforms = std::move(uni);
    b->plans = std::move(plans);
    b->active = std::move(active);
    for (auto &c : b->caps) c.clear();
    b->tables_dirty = true;
    b->launched = false;
    drop_graph(b);
    return EUCHIP_OK;
}

euchip_status euchip_batch_set_multi(euchip_batch *b, const euchip_pipeline_desc *descs, const euchip_request *reqs,
                                     uint32_t n) {
    if (!b || (n && (!descs || !reqs))) return fail(b ? b->ctx : nullptr, EUCHIP_INVALID_ARG, "null argument");
    euchip_ctx *ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    HIP_TRY(ctx, hipStreamSynchronize(b->stream));
    // this batch's own requests become empty; the parts replace them (existing
    // parts are reused, with the storage they grew)
    b->reqs.clear();
    b->uniforms.clear();
    b->plans.clear();
    b->active.clear();
    b->tables_dirty = true;
    b->launched = false;
    drop_graph(b);
    euchip_status s = EUCHIP_OK;
    std::vector<euchip_pipeline_desc> kinds;  // distinct descs, in order of first use
    std::vector<std::vector<uint32_t>> members;
    for (uint32_t i = 0; i < n; i++) {
        size_t g = 0;
        while (g < kinds.size() && !same_desc(kinds[g], descs[i])) g++;
        if (g == kinds.size()) kinds.push_back(descs[i]), members.emplace_back();
        members[g].push_back(i);
    }
    while (b->parts.size() > kinds.size()) {
        euchip_batch_destroy(b->parts.back());
        b->parts.pop_back();
    }
    for (size_t g = 0; g < kinds.size(); g++) {
        if (g == b->parts.size()) {
            euchip_batch *q = nullptr;
            if ((s = euchip_batch_create(ctx, &q))) break;
            q->tile_sx = b->tile_sx;
            b->parts.push_back(q);
        }
        euchip_batch *q = b->parts[g];
        std::vector<euchip_request> rs;
        for (uint32_t i : members[g]) rs.push_back(reqs[i]);
        if ((s = euchip_batch_set(q, &kinds[g], rs.data(), (uint32_t)rs.size()))) {
            // name the request by its index in this call
            const std::string msg = ctx->last_error;
            const size_t at = msg.find("request ");
            if (at == 0) {
                const uint32_t k = (uint32_t)std::strtoul(msg.c_str() + 8, nullptr, 10);
                if (k < members[g].size())
                    ctx->last_error = "request " + std::to_string(members[g][k]) + msg.substr(msg.find(':'));
            }
            break;
        }
    }
    if (s) drop_parts(b);
    return s;
}

euchip_status euchip_batch_set_tile_width(euchip_batch *b, uint32_t tile_w) {
    if (!b) return EUCHIP_INVALID_ARG;
    for (euchip_batch *q : b->parts) {
        const euchip_status s = euchip_batch_set_tile_width(q, tile_w);
        if (s) return s;
    }
    const uint32_t sx = tile_w == 8 ? 3u : tile_w == 16 ? 4u : tile_w == 32 ? 5u : 0u;
    if (!sx) return fail(b->ctx, EUCHIP_INVALID_ARG, "tile width %u (8, 16 or 32)", tile_w);
    if (sx != b->tile_sx) {
        HIP_TRY(b->ctx, hipStreamSynchronize(b->stream));
        b->tile_sx = sx;
        b->tables_dirty = true;
        drop_graph(b);
    }
    return EUCHIP_OK;
}

euchip_status euchip_batch_launch(euchip_batch *b, int use_graph) {
    if (!b) return EUCHIP_INVALID_ARG;
    (void)hipSetDevice(b->ctx->device);
    for (euchip_batch *q : b->parts) {
        const euchip_status s = launch_batch(q, use_graph != 0);
        if (s) return s;
    }
    return b->parts.empty() ? launch_batch(b, use_graph != 0) : EUCHIP_OK;
}

euchip_status euchip_batch_wait(euchip_batch *b) {
    if (!b) return EUCHIP_INVALID_ARG;
    (void)hipSetDevice(b->ctx->device);
    if (b->parts.empty()) return wait_batch(b);
    euchip_status s = EUCHIP_OK;
    for (euchip_batch *q : b->parts) {
        const euchip_status t = wait_batch(q);
        if (!s) s = t;
    }
    return s;
}

euchip_status euchip_batch_query(euchip_batch *b) {
    if (!b) return EUCHIP_INVALID_ARG;
    for (euchip_batch *q : b->parts) {
        const euchip_status s = euchip_batch_query(q);
        if (s) return s;
    }
    if (!b->parts.empty()) return EUCHIP_OK;
    const hipError_t e = hipEventQuery(b->done);
    if (e == hipSuccess) return EUCHIP_OK;
    if (e == hipErrorNotReady) return EUCHIP_NOT_READY;
    return hip_fail(b->ctx, e, "hipEventQuery");
}

euchip_status euchip_batch_autotune(euchip_batch *b, uint32_t reps, uint32_t *best_tile_w, float *times_ms) {
    if (!b) return EUCHIP_INVALID_ARG;
    for (size_t i = 0; i < b->parts.size(); i++) {
        const euchip_status s = euchip_batch_autotune(b->parts[i], reps, i ? nullptr : best_tile_w, i ? nullptr : times_ms);
        if (s) return s;
    }
    if (!b->parts.empty()) return EUCHIP_OK;
    euchip_ctx *ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    reps = std::max<uint32_t>(reps, 1u);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(ctx, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(ctx, EUCHIP_HIP_ERROR, "event");
    }
    float best = 0.0f;
    uint32_t best_w = 32;
    euchip_status s = EUCHIP_OK;
    for (uint32_t w : {8u, 16u, 32u}) {
        if ((s = euchip_batch_set_tile_width(b, w))) break;
        // two calibration launches size the bins for this width
        for (int i = 0; i < 2 && !s; i++) {
            if (!(s = launch_batch(b, false))) s = wait_batch(b);
        }
        if (s) break;
        if (b->tables_dirty && (s = build_tables(b))) break;
        (void)hipEventRecord(e0, b->stream);
        for (uint32_t i = 0; i < reps && !s; i++) s = enqueue_batch(b);
        (void)hipEventRecord(e1, b->stream);
        if (s) break;
        b->launched = true;
        if ((s = wait_batch(b))) break;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)reps;
        if (times_ms) times_ms[(w == 8) ? 0 : (w == 16) ? 1 : 2] = ms;
        if (w == 8 || ms < best) best = ms, best_w = w;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (s) return s;
    if (best_tile_w) *best_tile_w = best_w;
    return euchip_batch_set_tile_width(b, best_w);
}

void *euchip_batch_stream(euchip_batch *b) {
    if (!b) return nullptr;
    return (void *)(b->parts.empty() ? b->stream : b->parts[0]->stream);
}

size_t euchip_batch_scratch_bytes(const euchip_batch *b) {
    if (!b) return 0;
    size_t parts = 0;
    for (const euchip_batch *q : b->parts) parts += euchip_batch_scratch_bytes(q);
    return parts + b->tables.cap * 16 + b->rec0.cap * 16 +
           b->rec1.cap * 16 +
           b->vary.cap * 4 + b->tile_count.cap * sizeof(TileBin) + b->busy.cap * 4 + b->bins.cap * 4 + b->ctrbuf.cap * 4 +
           b->pool.cap * 4;
}

size_t euchip_batch_estimate_bytes(const euchip_pipeline_desc *desc, const euchip_request *reqs, uint32_t n) {
    if (!desc || (n && !reqs)) return 0;
    size_t total = sizeof(Counters) + 8 * ((size_t)1 << 20);
    for (uint32_t i = 0; i < n; i++) {
        const euchip_request &r = reqs[i];
        if (!r.mesh || !r.uniforms) continue;
        Plan P;
        const euchip_buffer *tex = r.texture ? r.texture : desc->texture;
        const euchip_buffer *shw = r.shadow ? r.shadow : desc->shadow;
        if (plan_request(r.mesh->ctx, desc, r.uniforms, r.mesh, r.pixel, r.depth, tex, shw, P) || !P.render) continue;
        total += euchip::request_scratch_estimate(desc, r.mesh, P.v.W, P.v.H);
    }
    return total;
}

}  // extern "C"
