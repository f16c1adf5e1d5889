/*
 * euc_oracle.h -- CPU restatement of zesterer/euc 0.6.0's triangle path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (euc_amd/, include/) may
 * link, load or call this library.  It is the checker the HIP path is compared
 * against: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg are
 * its only users.
 *
 * Parameter blocks below are plain structs filled from Python (ctypes); the
 * float layout of the uniform block is the interface format shared with the
 * product's euchip_uniforms (include/euchip.h), nothing more.
 */
#ifndef EUC_ORACLE_H
#define EUC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fixed shader set (the same ids as the product's EUCHIP_SHADER_*). */
enum {
    ORC_SHADER_FLAT = 0,        /* provoking RGBA8 colour, replace blend             */
    ORC_SHADER_COLOR = 1,       /* per-vertex Rgba<f32>, interpolated (triangle.rs)  */
    ORC_SHADER_GOURAUD = 2,     /* per-vertex lighting, 2 directional lights         */
    ORC_SHADER_PHONG = 3,       /* per-fragment lighting, 2 directional lights       */
    ORC_SHADER_DEPTH_ONLY = 4,  /* Unit fragment, PixelMode::PASS pipelines          */
    ORC_SHADER_TEXTURED = 5,    /* bilinear RGBA8 texture, `e as u8` blend           */
    ORC_SHADER_PHONG_SHADOW = 6,/* teapot model: point light + shadow-map lookup     */
    ORC_SHADER_TEX_SHADOW = 7   /* texture x sun light x shadow-map lookup           */
};

/* u32 pixel layouts of the Rgba<f32> shaders' blend: [r, g, b, a] or, as the
 * teapot example's window framebuffer (examples/teapot.rs:135-140), [b, g, r, a]
 * little-endian. */
enum { ORC_PIXEL_RGBA8 = 0, ORC_PIXEL_BGRA8 = 1 };
/* Texture samplers (src/sampler/mod.rs, linear.rs, nearest.rs): filter x wrapping. */
enum {
    ORC_SAMPLER_DEFAULT = 0,         /* TEXTURED: Linear, TEX_SHADOW: Clamped<Linear>   */
    ORC_SAMPLER_LINEAR = 1, ORC_SAMPLER_LINEAR_CLAMPED = 2, ORC_SAMPLER_LINEAR_TILED = 3,
    ORC_SAMPLER_LINEAR_MIRRORED = 4,
    ORC_SAMPLER_NEAREST = 5, ORC_SAMPLER_NEAREST_CLAMPED = 6, ORC_SAMPLER_NEAREST_TILED = 7,
    ORC_SAMPLER_NEAREST_MIRRORED = 8
};
/* Fixed geometry stages (Pipeline::geometry, src/pipeline.rs:221-229) of a
 * triangle [a, b, c]: output it; output [a, c, b]; output both, in that order. */
enum { ORC_GEOMETRY_IDENTITY = 0, ORC_GEOMETRY_REVERSED = 1, ORC_GEOMETRY_BOTH_SIDES = 2 };
/* PrimitiveKind (src/primitives.rs): TriangleList, LineList, LineTriangleList. */
enum { ORC_PRIMS_TRIANGLE_LIST = 0, ORC_PRIMS_LINE_LIST = 1, ORC_PRIMS_LINE_TRIANGLE_LIST = 2 };

enum { ORC_DEPTH_NONE = 0, ORC_DEPTH_LESS = 1, ORC_DEPTH_EQUAL = 2, ORC_DEPTH_GREATER = 3 };
enum { ORC_CULL_NONE = 0, ORC_CULL_BACK = 1, ORC_CULL_FRONT = 2 };
enum { ORC_BANDS_PAR = 0, ORC_BANDS_SEQ = 1 };
enum { ORC_XFORM_CLIP = 0, ORC_XFORM_MVP = 1 };

enum {
    ORC_OK = 0,
    ORC_SIZE_MISMATCH = 1,  /* assert_eq!(pixel.size(), depth.size()) panics        */
    ORC_BAND_ZERO = 2,      /* render_par: tgt_size[1] / group_rows with G == 0     */
    ORC_INVALID = 3
};

typedef struct {
    int32_t shader;
    int32_t xform;          /* ORC_XFORM_*                                          */
    int32_t pixel_write;    /* PixelMode::write                                     */
    int32_t depth_test;     /* ORC_DEPTH_*  (DepthMode::test)                       */
    int32_t depth_write;    /* DepthMode::write                                     */
    int32_t y_up;           /* YAxisDirection::Up                                   */
    int32_t z_clip;         /* z_clip_range.is_some()                               */
    float z_min, z_max;     /* z_clip_range start/end (inclusive test)              */
    int32_t cull;           /* ORC_CULL_*                                           */
    int32_t aa_msaa;        /* AaMode::Msaa selected                                */
    int32_t aa_level;       /* AaMode::Msaa { level }                               */
    int32_t band_policy;    /* ORC_BANDS_*                                          */
    int32_t threads;        /* worker threads for PAR (available_parallelism)       */
    int32_t group_rows;     /* tests only: PAR band height override (0 = 20000*2^L/W) */
    int32_t pixel_format;   /* ORC_PIXEL_*: how blend packs Rgba<u8> into the u32 texel  */
    int32_t sampler;        /* ORC_SAMPLER_*: texture sampler (0 = the shader's own)  */
    int32_t geometry;       /* ORC_GEOMETRY_*: the fixed Pipeline::geometry stage     */
    int32_t primitives;     /* ORC_PRIMS_*: PrimitiveKind of the vertex stream        */
} orc_desc;

/* Vertex stream: element e -> (instance e / count, vertex idx[e % count]). */
typedef struct {
    const float *pos;       /* pos_comps floats per vertex                           */
    int32_t pos_comps;      /* 3 (object space, MVP) or 4 (clip space)               */
    const float *nrm;       /* 3 floats per vertex or NULL                           */
    const float *col;       /* 4 floats per vertex (Rgba<f32>) or NULL               */
    const float *uv;        /* 2 floats per vertex or NULL                           */
    const uint32_t *rgba8;  /* packed RGBA8 per vertex (FLAT) or NULL                */
    const uint32_t *idx;    /* index buffer or NULL (identity)                       */
    int64_t count;          /* stream elements per instance                          */
    int64_t n_vertices;     /* vertices addressable by idx                           */
    int32_t n_instances;
    const float *models;    /* n_instances * 16 floats, row-major                    */
} orc_stream;

typedef struct {
    const void *data;       /* RGBA8 (uint32) or f32 texels, row-major x + w*y      */
    int32_t w, h;
    int32_t is_f32;
} orc_texture;

/*
 * Render one pipeline call, as Pipeline::render (pipeline.rs:248-300) does with
 * the fixed shader `desc->shader`.  pixel/depth NULL = Empty (size [0,0]).
 * uniforms: 64 floats (layout documented in euc_oracle.c).
 */
int orc_render(const orc_desc *desc, const float *uniforms, const orc_stream *s,
               const orc_texture *tex, const orc_texture *shadow,
               uint32_t *pixel, int32_t pw, int32_t ph,
               float *depth, int32_t dw, int32_t dh);

/* Same, but also returns the number of emit_fragment calls (shaded fragments). */
int orc_render_count(const orc_desc *desc, const float *uniforms, const orc_stream *s,
                     const orc_texture *tex, const orc_texture *shadow,
                     uint32_t *pixel, int32_t pw, int32_t ph,
                     float *depth, int32_t dw, int32_t dh, int64_t *frags);

/* Per-triangle setup values of triangles.rs:27-199 for one band, for known-answer
 * tests.  out (37 floats): M[9], o[3], dx[3], dy[3], bmin[2], bmax[2],
 * verts_by_y[6], hom_z[3], allin, reversed, culled, winding.
 * Returns 0 when the triangle is culled (fields after `culled` still valid). */
int orc_setup(const float hom_in[12], int32_t w, int32_t h, int32_t by0, int32_t by1,
              int32_t y_up, int32_t z_clip, float z_min, float z_max, int32_t cull,
              float *out);

/* The pixels (x, y pairs into out_xy, at most max) of the clipped Bresenham line
 * the lines rasterizer draws (lines.rs:83-117); returns their number. */
int64_t orc_clip_line(int64_t x1, int64_t y1, int64_t x2, int64_t y2, int64_t cx0, int64_t cy0, int64_t cx1,
                      int64_t cy1, int64_t *out_xy, int64_t max);

/* a + d + d + ... (n times) in f32, the row walk of triangles.rs:301. */
float orc_walk(float a, float d, int64_t n);

/* Bilinear Clamped<Linear> sample of a texture (sampler/mod.rs:183-186,
 * sampler/linear.rs:30-66); out gets 4 (RGBA8 -> f32) or 1 (f32) values. */
void orc_sample(const orc_texture *t, float u, float v, int32_t clamped, float *out);

/* A sample through sampler `kind` (ORC_SAMPLER_LINEAR ... NEAREST_MIRRORED). */
void orc_sample_with(const orc_texture *t, int32_t kind, float u, float v, float *out);

/* f32::rem_euclid (the Tiled / Mirrored index maps, src/sampler/mod.rs:146-179). */
float orc_rem_euclid(float x, float m);

/* Pipeline::vertex of the fixed shader set for stream vertex `vi` of instance
 * `inst`: the clip position and the VertexData (returns its float count). */
int orc_vertex(const orc_desc *d, const float *uniforms, const orc_stream *s, int32_t inst, int64_t vi,
               float pos[4], float vd[12]);

/* Pipeline::fragment of the fixed shader set on VertexData vd (out: Rgba<f32>,
 * FLAT: colour bits in out[0]) and Pipeline::blend onto old_pixel (blended may
 * be NULL). */
int orc_fragment(const orc_desc *d, const float *uniforms, const orc_texture *tex, const orc_texture *shadow,
                 const float *vd, float out[4], uint32_t old_pixel, uint32_t *blended);

/* The fixed fragment-shader math, exposed for unit tests. */
void orc_lighting2(const float *uniforms, const float wpos[3], const float nrm[3], float out[4]);

#ifdef __cplusplus
}
#endif
#endif
