/*
 * euchip.h -- C ABI of the MI355X (gfx950) triangle-rasterisation path.
 *
 * Drop-in boundary for euc 0.6.0's `Pipeline::render` (src/pipeline.rs:248-300):
 * a thin Rust FFI crate implements a `GpuPipeline` trait on top of these
 * entry points and keeps the reference's surface (TriangleList, Buffer2d,
 * Empty, DepthMode, PixelMode, CoordinateMode, CullMode, AaMode).  Rust
 * shader closures cannot run on the device, so a pipeline names one of the
 * fixed device shaders (EUCHIP_SHADER_*) and passes a 64-float uniform block.
 *
 * Conventions
 *   - Plain C types only; every pointer argument documents whether it is host
 *     or device memory.  Host memory passed in is never retained after return.
 *   - A context is externally synchronised: one context per host thread, or
 *     callers serialise their calls on it.
 *   - euchip_render() blocks until the targets hold the result, like the
 *     reference's render().  The *_async variants only enqueue on the
 *     context's stream; euchip_synchronize() waits and reports deferred errors.
 *   - Errors are status codes.  The reference panics in these cases; a Rust
 *     wrapper over this header re-panics with the reference's message.
 */
#ifndef EUCHIP_H
#define EUCHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum euchip_status {
    EUCHIP_OK = 0,
    /* assert_eq!(pixel.size(), depth.size()) in render (src/pipeline.rs:262-266) */
    EUCHIP_SIZE_MISMATCH = 1,
    /* render_par divides the target height by G = 20000*2^L/W, which is 0 when
     * W > 20000*2^L (src/pipeline.rs:328-330) */
    EUCHIP_BAND_ZERO = 2,
    EUCHIP_INVALID_ARG = 3,
    EUCHIP_OOM = 4,
    /* a request does not fit the server's HBM budget, or bin storage overflowed */
    EUCHIP_OVER_BUDGET = 5,
    EUCHIP_HIP_ERROR = 6,
    EUCHIP_NO_DEVICE = 7,
    EUCHIP_NOT_READY = 8
} euchip_status;

/* Fixed device shader set (vertex + fragment + blend). */
typedef enum euchip_shader {
    EUCHIP_SHADER_FLAT = 0,         /* provoking-vertex RGBA8, replace blend          */
    EUCHIP_SHADER_COLOR = 1,        /* interpolated Rgba<f32>, examples/triangle.rs    */
    EUCHIP_SHADER_GOURAUD = 2,      /* per-vertex lighting, 2 directional lights      */
    EUCHIP_SHADER_PHONG = 3,        /* per-fragment lighting, 2 directional lights    */
    EUCHIP_SHADER_DEPTH_ONLY = 4,   /* Unit fragment, examples/teapot.rs:9-50          */
    EUCHIP_SHADER_TEXTURED = 5,     /* bilinear texture, examples/texture_mapping.rs   */
    EUCHIP_SHADER_PHONG_SHADOW = 6, /* examples/teapot.rs:52-141                       */
    EUCHIP_SHADER_TEX_SHADOW = 7    /* texture x sun light x shadow map                */
} euchip_shader;

/* DepthMode::test (src/pipeline.rs:14-19); NONE = `test: None`. */
typedef enum { EUCHIP_DEPTH_NONE = 0, EUCHIP_DEPTH_LESS = 1, EUCHIP_DEPTH_EQUAL = 2, EUCHIP_DEPTH_GREATER = 3 } euchip_depth_test;
/* CullMode (src/rasterizer/mod.rs:9-18). */
typedef enum { EUCHIP_CULL_NONE = 0, EUCHIP_CULL_BACK = 1, EUCHIP_CULL_FRONT = 2 } euchip_cull;
/* render_par (default `par` feature) or render_seq (src/pipeline.rs:296-299). */
typedef enum { EUCHIP_BANDS_PAR = 0, EUCHIP_BANDS_SEQ = 1 } euchip_band_policy;
/* How the vertex shader forms the clip position. */
typedef enum { EUCHIP_XFORM_CLIP = 0, EUCHIP_XFORM_MVP = 1 } euchip_xform;
/* How the blend of the Rgba<f32> shaders packs a pixel into its u32 texel:
 * u32::from_le_bytes([r, g, b, a]), or [b, g, r, a] as the teapot example's
 * window framebuffer does (examples/teapot.rs:135-140). */
typedef enum { EUCHIP_PIXEL_RGBA8 = 0, EUCHIP_PIXEL_BGRA8 = 1 } euchip_pixel_format;
/* Texture sampler of TEXTURED / TEX_SHADOW: Linear (src/sampler/linear.rs:30-66)
 * or Nearest (src/sampler/nearest.rs:20-38, Denormalize src/math.rs:43-79),
 * bare or wrapped in Clamped / Tiled / Mirrored (src/sampler/mod.rs:98-179).
 * DEFAULT: Linear for TEXTURED, Clamped<Linear> for TEX_SHADOW. */
typedef enum {
    EUCHIP_SAMPLER_DEFAULT = 0,
    EUCHIP_SAMPLER_LINEAR = 1, EUCHIP_SAMPLER_LINEAR_CLAMPED = 2, EUCHIP_SAMPLER_LINEAR_TILED = 3,
    EUCHIP_SAMPLER_LINEAR_MIRRORED = 4,
    EUCHIP_SAMPLER_NEAREST = 5, EUCHIP_SAMPLER_NEAREST_CLAMPED = 6, EUCHIP_SAMPLER_NEAREST_TILED = 7,
    EUCHIP_SAMPLER_NEAREST_MIRRORED = 8
} euchip_sampler;
/* Fixed geometry stages (Pipeline::geometry, src/pipeline.rs:221-229) of each
 * triangle [a, b, c]: output it (the default); output [a, c, b]; output both,
 * in that order (each output is its own primitive). */
typedef enum { EUCHIP_GEOMETRY_IDENTITY = 0, EUCHIP_GEOMETRY_REVERSED = 1, EUCHIP_GEOMETRY_BOTH_SIDES = 2 } euchip_geometry;
/* PrimitiveKind of the vertex stream (src/primitives.rs:18-104). */
typedef enum {
    EUCHIP_PRIMS_TRIANGLE_LIST = 0, EUCHIP_PRIMS_LINE_LIST = 1, EUCHIP_PRIMS_LINE_TRIANGLE_LIST = 2
} euchip_primitives;
/* Texel formats: RGBA8 packed little-endian in a uint32 ([u8;4]), or f32. */
typedef enum { EUCHIP_FORMAT_RGBA8 = 0, EUCHIP_FORMAT_F32 = 1 } euchip_format;

typedef struct euchip_ctx euchip_ctx;
typedef struct euchip_buffer euchip_buffer;
typedef struct euchip_mesh euchip_mesh;

/* One `impl Pipeline` as data (src/pipeline.rs:171-243). */
typedef struct euchip_pipeline_desc {
    uint32_t shader;       /* euchip_shader                                         */
    uint32_t xform;        /* euchip_xform                                          */
    uint32_t pixel_write;  /* PixelMode::write (src/pipeline.rs:58-61)              */
    uint32_t depth_test;   /* euchip_depth_test                                     */
    uint32_t depth_write;  /* DepthMode::write                                      */
    uint32_t y_up;         /* CoordinateMode::y_axis_direction == Up                */
    uint32_t z_clip;       /* CoordinateMode::z_clip_range.is_some()                */
    float z_min, z_max;    /* z_clip_range (inclusive test, pipeline.rs:151-156)    */
    uint32_t cull;         /* euchip_cull (Triangles' rasterizer_config)            */
    uint32_t aa_msaa;      /* AaMode::Msaa selected                                 */
    uint32_t aa_level;     /* AaMode::Msaa { level }, clamped to [0, 6]             */
    uint32_t band_policy;  /* euchip_band_policy                                    */
    uint32_t threads;      /* available_parallelism() of the caller (PAR: T = min(H/G, threads)) */
    /* Load ops: 1 = the target is cleared to the value first, fused into the
     * raster pass (Target::clear + render, src/texture.rs:469-477).          */
    uint32_t clear_color_enable;
    uint32_t clear_color;  /* RGBA8 texel                                           */
    uint32_t clear_depth_enable;
    float clear_depth;
    const euchip_buffer *texture; /* RGBA8 texture for TEXTURED / TEX_SHADOW, else NULL */
    const euchip_buffer *shadow;  /* f32 shadow map for *_SHADOW, else NULL             */
    uint32_t pixel_format; /* euchip_pixel_format                                   */
    uint32_t sampler;      /* euchip_sampler                                        */
    uint32_t geometry;     /* euchip_geometry (TRIANGLE_LIST streams)               */
    uint32_t primitives;   /* euchip_primitives                                     */
} euchip_pipeline_desc;

/* Uniform block: 64 floats.
 *   [0,16)  view_proj (row-major)      [16,20) light 0 dir xyz, intensity
 *   [20,24) light 1 dir xyz, intensity [24,28) camera position xyz
 *   [28,32) surface colour rgba        [32,36) ambient, k_diffuse, k_specular, shadow_bias
 *   [36,40) point light position xyz   [40,56) light view_proj (row-major)   */
#define EUCHIP_UNIFORM_FLOATS 64

/* Vertex stream for a TriangleList (src/primitives.rs:18-44), optionally
 * indexed (src/index.rs:4-55).  Stream element e reads vertex
 * idx[e % count] (or e % count) of instance e / count; triangle t is elements
 * 3t, 3t+1, 3t+2 and a trailing partial triangle is dropped.  Host pointers;
 * euchip_mesh_create uploads them into device-resident storage. */
typedef struct euchip_mesh_desc {
    const float *pos;       /* pos_comps floats per vertex (3: object space, 4: clip) */
    uint32_t pos_comps;
    const float *nrm;       /* 3 floats per vertex, or NULL                          */
    const float *col;       /* 4 floats per vertex (Rgba<f32>), or NULL               */
    const float *uv;        /* 2 floats per vertex, or NULL                          */
    const uint32_t *rgba8;  /* packed RGBA8 per vertex (FLAT), or NULL                */
    uint64_t n_vertices;
    const uint32_t *idx;    /* count indices, or NULL for the identity               */
    uint64_t count;         /* stream elements per instance                          */
    uint32_t n_instances;
    const float *models;    /* n_instances * 16 floats (row-major), or NULL = identity */
} euchip_mesh_desc;

/* Per-kernel timing collected while profiling is on (hipEvents on the
 * context's stream around each launch). */
typedef struct euchip_kernel_stat {
    char name[32];
    uint64_t launches;
    double total_ms;
} euchip_kernel_stat;

/* ---- context -------------------------------------------------------------- */
const char *euchip_status_string(euchip_status s);
euchip_status euchip_ctx_create(int device, euchip_ctx **out);
euchip_status euchip_ctx_destroy(euchip_ctx *ctx);
const char *euchip_last_error(const euchip_ctx *ctx);
/* The hipStream_t all work of this context is enqueued on. */
void *euchip_ctx_stream(euchip_ctx *ctx);
euchip_status euchip_synchronize(euchip_ctx *ctx);
euchip_status euchip_set_profiling(euchip_ctx *ctx, int enable);
/* Copies up to max entries; returns the number of distinct kernels. */
int euchip_kernel_stats(euchip_ctx *ctx, euchip_kernel_stat *out, int max);
void euchip_reset_stats(euchip_ctx *ctx);

/* ---- targets and textures: Buffer2d<T> (src/buffer.rs:19-22, x + W*y) ------- */
euchip_status euchip_buffer_create(euchip_ctx *ctx, euchip_format fmt, uint32_t w, uint32_t h, euchip_buffer **out);
/* Wrap caller-owned device memory (w*h texels, row-major); not freed by destroy. */
euchip_status euchip_buffer_wrap(euchip_ctx *ctx, euchip_format fmt, uint32_t w, uint32_t h, void *device_ptr,
                                 euchip_buffer **out);
euchip_status euchip_buffer_destroy(euchip_buffer *buf);
euchip_status euchip_buffer_upload(euchip_buffer *buf, const void *host, size_t bytes);
euchip_status euchip_buffer_download(const euchip_buffer *buf, void *host, size_t bytes);
/* Target::clear (src/buffer.rs:213-218): texel points at one 4-byte texel (host). */
euchip_status euchip_buffer_clear(euchip_buffer *buf, const void *texel);
void *euchip_buffer_device_ptr(const euchip_buffer *buf);

/* ---- resident meshes --------------------------------------------------------- */
euchip_status euchip_mesh_create(euchip_ctx *ctx, const euchip_mesh_desc *desc, euchip_mesh **out);
euchip_status euchip_mesh_destroy(euchip_mesh *mesh);
/* Replace the per-instance model matrices (host, n_instances * 16 floats). */
euchip_status euchip_mesh_set_models(euchip_mesh *mesh, const float *models, uint32_t n_instances);

/* ---- rendering ------------------------------------------------------------------ */
/* Pipeline::render (src/pipeline.rs:248-300).  pixel / depth NULL = Empty.
 * uniforms: EUCHIP_UNIFORM_FLOATS host floats.  Blocks until done. */
euchip_status euchip_render(euchip_ctx *ctx, const euchip_pipeline_desc *desc, const float *uniforms,
                            const euchip_mesh *mesh, euchip_buffer *pixel, euchip_buffer *depth);
/* Same, enqueue only. */
euchip_status euchip_render_async(euchip_ctx *ctx, const euchip_pipeline_desc *desc, const float *uniforms,
                                  const euchip_mesh *mesh, euchip_buffer *pixel, euchip_buffer *depth);

/* Same as euchip_render (blocking, same results), but allocates no device memory:
 * every per-frame buffer comes from one array the library holds statically on
 * each device it runs on (64 MiB, held until the process ends).  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Limits: at most 262,144
 * triangles, 3 Mi VertexData floats (shaded vertices x the shader's floats),
 * 16,384 tiles of 32 x 8 pixels (e.g. 2048 x 2048) -- beyond them it returns
 * EUCHIP_OVER_BUDGET before enqueuing anything; a frame whose bin overflow does
 * not fit the array's pool (about 4 Mi entries) also returns EUCHIP_OVER_BUDGET,
 * its targets unwritten.  Every call on a device shares that device's array:
 * calls that overlap on one device (other host threads, other contexts) race. */
euchip_status euchip_render_noalloc(euchip_ctx *ctx, const euchip_pipeline_desc *desc, const float *uniforms,
                                    const euchip_mesh *mesh, euchip_buffer *pixel, euchip_buffer *depth);

/* ---- device memory budget ---------------------------------------------------------- */
/* Bytes of device memory the context holds (buffers, meshes, per-frame and batch
 * scratch).  With a budget set (0 = none), an allocation that would take the total
 * past it fails with EUCHIP_OVER_BUDGET and allocates nothing. */
size_t euchip_ctx_bytes_in_use(const euchip_ctx *ctx);
euchip_status euchip_ctx_set_budget(euchip_ctx *ctx, size_t bytes);

/* ---- batched rendering -------------------------------------------------------------
 * Many render() calls of one pipeline in one setup launch and one raster launch:
 * each request has its own mesh, targets and uniform block, and its results are
 * those of euchip_render() on that request alone (same targets, same uniforms).
 * A batch owns its stream (default priority) and its scratch memory, so batches
 * can be in flight together and with the context's own stream.  No counterpart
 * in euc: it is the serving form of Pipeline::render (src/pipeline.rs:248-300)
 * applied to many independent targets. */
typedef struct euchip_batch euchip_batch;

typedef struct euchip_request {
    const euchip_mesh *mesh;
    euchip_buffer *pixel;          /* RGBA8 target or NULL (Empty) */
    euchip_buffer *depth;          /* f32 target or NULL (Empty) */
    const float *uniforms;         /* EUCHIP_UNIFORM_FLOATS host floats, copied by batch_set */
    const euchip_buffer *texture;  /* NULL: the pipeline desc's texture */
    const euchip_buffer *shadow;   /* NULL: the pipeline desc's shadow map */
} euchip_request;

euchip_status euchip_batch_create(euchip_ctx *ctx, euchip_batch **out);
euchip_status euchip_batch_destroy(euchip_batch *b);
/* Replace the requests (desc and requests are copied; scratch grows, never
 * shrinks).  Fails with the first request's error (EUCHIP_SIZE_MISMATCH,
 * EUCHIP_BAND_ZERO, EUCHIP_INVALID_ARG) naming its index in last_error. */
euchip_status euchip_batch_set(euchip_batch *b, const euchip_pipeline_desc *desc, const euchip_request *reqs,
                               uint32_t n);
/* Replace the requests, request i rendered with descs[i] (no counterpart in
 * euc: a batch of render() calls of different pipelines).  Requests of
 * identical descs share one setup and one raster launch; each distinct desc
 * runs as a part of the batch on a stream of its own, so the parts overlap.
 * launch / wait / query / set_tile_width / autotune act on every part
 * (autotune tunes each part; best_tile_w and times_ms report the first);
 * euchip_batch_stream returns the first part's stream.  Errors as
 * euchip_batch_set, naming the request's index. */
euchip_status euchip_batch_set_multi(euchip_batch *b, const euchip_pipeline_desc *descs, const euchip_request *reqs,
                                     uint32_t n);
/* Raster tile shape: tile_w x (256 / tile_w) pixels, tile_w 8, 16 or 32. */
euchip_status euchip_batch_set_tile_width(euchip_batch *b, uint32_t tile_w);
/* Launch the batch `reps` times at each tile width, keep the fastest.
 * best_tile_w (may be NULL) gets it; times_ms (may be NULL) gets the mean launch
 * time of widths 8, 16, 32. */
euchip_status euchip_batch_autotune(euchip_batch *b, uint32_t reps, uint32_t *best_tile_w, float *times_ms);
/* Enqueue the batch on its stream.  use_graph: replay a hipGraph of the launch
 * (captured on first use, recaptured after set / set_tile_width / bin growth). */
euchip_status euchip_batch_launch(euchip_batch *b, int use_graph);
/* Wait for the last launch.  A launch whose bins overflowed their storage is
 * re-run with larger bins before this returns, so on EUCHIP_OK every target holds
 * its result. */
euchip_status euchip_batch_wait(euchip_batch *b);
/* EUCHIP_OK when the last launch has finished, EUCHIP_NOT_READY otherwise. */
euchip_status euchip_batch_query(euchip_batch *b);
void *euchip_batch_stream(euchip_batch *b);
/* Device bytes of the batch's scratch (vertices, bins, tables). */
size_t euchip_batch_scratch_bytes(const euchip_batch *b);
/* Device bytes a batch of these requests would add to an empty batch's scratch at
 * tile width 32: the admission estimate of a serving loop. */
size_t euchip_batch_estimate_bytes(const euchip_pipeline_desc *desc, const euchip_request *reqs, uint32_t n);

/* ---- serving loop ------------------------------------------------------------------
 * Continuous batching over resident meshes and textures: submitted requests wait
 * in a FIFO queue; each poll retires finished batches and admits queued requests
 * into free batch slots while the context's device memory (targets, scratch, the
 * resident data) stays within the budget, then launches them.  Slots run on their
 * own streams, so a batch admitted while another is on the GPU overlaps with it.
 * Each request's targets hold exactly what euchip_render of that request alone
 * would leave.  No counterpart in euc (a serving layer over Pipeline::render). */
typedef struct euchip_server euchip_server;

typedef struct euchip_serve_request {
    const euchip_mesh *mesh;       /* resident mesh */
    const euchip_buffer *texture;  /* NULL: the pipeline desc's texture */
    const euchip_buffer *shadow;   /* NULL: the pipeline desc's shadow map */
    const float *uniforms;         /* EUCHIP_UNIFORM_FLOATS host floats, copied */
    uint32_t width, height;        /* the server allocates an RGBA8 and an f32 target of this size */
} euchip_serve_request;

typedef struct euchip_completion {
    uint64_t ticket;
    euchip_status status;          /* EUCHIP_OVER_BUDGET: the request alone does not fit the budget */
    euchip_buffer *pixel;          /* valid until euchip_server_release(ticket) */
    euchip_buffer *depth;
    uint64_t done_ns;              /* euchip_clock_ns() when the server saw the request finish */
} euchip_completion;

typedef struct euchip_server_stats {
    uint64_t submitted, admitted, completed, rejected, batches;
    uint32_t queued, in_flight;
    size_t bytes_in_use, peak_bytes, budget;
} euchip_server_stats;

/* budget_bytes: device bytes the context may hold (0 = no limit; it becomes the
 * context's budget); max_batch: requests per batch; slots: batches in flight. */
euchip_status euchip_server_create(euchip_ctx *ctx, const euchip_pipeline_desc *desc, size_t budget_bytes,
                                   uint32_t max_batch, uint32_t slots, euchip_server **out);
euchip_status euchip_server_destroy(euchip_server *srv);
euchip_status euchip_server_submit(euchip_server *srv, const euchip_serve_request *req, uint64_t *ticket);
/* Submit a request rendered with its own pipeline desc (NULL: the server's),
 * which must draw triangle lists with both clears enabled (EUCHIP_INVALID_ARG
 * otherwise).  Requests of different descs admitted together run as parts of
 * one batch (euchip_batch_set_multi).  No counterpart in euc. */
euchip_status euchip_server_submit_desc(euchip_server *srv, const euchip_serve_request *req,
                                        const euchip_pipeline_desc *desc, uint64_t *ticket);
/* Submit n requests in order; tickets (may be NULL) receives their tickets. */
euchip_status euchip_server_submit_n(euchip_server *srv, const euchip_serve_request *reqs, uint32_t n,
                                     uint64_t *tickets);
/* Retire finished batches, admit and launch queued requests; copies up to max
 * completions (oldest first) to out and their number to *n_out. */
euchip_status euchip_server_poll(euchip_server *srv, euchip_completion *out, uint32_t max, uint32_t *n_out);
/* Hand a completed request's targets back to the server's pool. */
euchip_status euchip_server_release(euchip_server *srv, uint64_t ticket);
euchip_status euchip_server_release_n(euchip_server *srv, const uint64_t *tickets, uint32_t n);
/* The host clock of completion times (monotonic, nanoseconds). */
uint64_t euchip_clock_ns(void);
euchip_status euchip_server_get_stats(const euchip_server *srv, euchip_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif /* EUCHIP_H */
