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
 *   - Errors are status codes.  The reference panics in these cases; the Rust
 *     wrapper re-panics with the reference's message.
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

#ifdef __cplusplus
}
#endif
#endif /* EUCHIP_H */
