/*
 * euc_oracle.c -- CPU restatement of zesterer/euc 0.6.0's triangle path.
 *
 * TEST INFRASTRUCTURE ONLY (see euc_oracle.h).  Built with
 *   gcc -O2 -ffp-contract=off -fno-fast-math
 * so that every f32 operation below rounds exactly once, in the order written,
 * which is the order rustc evaluates the reference's expressions in.
 *
 * Reference files restated here (paths relative to the euc crate root):
 *   src/pipeline.rs:248-300   Pipeline::render (target size, MSAA level clamp)
 *   src/pipeline.rs:303-366   render_par (row bands, G = 20000*2^L/W, T = H/G)
 *   src/pipeline.rs:368-394   render_seq (one band)
 *   src/pipeline.rs:396-614   render_inner + BlitterImpl (depth test, emit, MSAA)
 *   src/rasterizer/triangles.rs:15-367   Triangles::rasterize and f32 helpers
 *   src/math.rs:23-41         WeightedSum for Mul<f32>+Add types
 *   src/sampler/linear.rs:30-66, src/sampler/mod.rs:183-186   Clamped<Linear>
 *   src/primitives.rs:23-44, src/index.rs:44-55   TriangleList / IndexedVertices
 *
 * The fixed shader set (vertex/fragment/blend) is this project's own
 * definition of the example shader models (examples/teapot.rs:9-141,
 * examples/triangle.rs:7-25, examples/texture_mapping.rs:12-35); the product's
 * device shaders implement the same definitions.
 *
 * Uniform block (64 floats):
 *   [0..16)  view_proj, row-major          [16..20) light 0: dir xyz, intensity
 *   [20..24) light 1: dir xyz, intensity   [24..28) camera position xyz
 *   [28..32) surface colour rgba           [32..36) ambient, k_diffuse, k_specular, shadow_bias
 *   [36..40) point light position xyz      [40..56) light view_proj, row-major
 */
#include "euc_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define U_VP 0
#define U_L0 16
#define U_L1 20
#define U_CAM 24
#define U_SURF 28
#define U_K 32
#define U_LPOS 36
#define U_LVP 40

#define MAX_VD 12

/* ------------------------------------------------------------------------- */
/* Rust scalar semantics                                                      */
/* ------------------------------------------------------------------------- */

/* `f as usize`: saturating, NaN -> 0 (Rust reference: numeric casts). */
static size_t sat_usize(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 18446744073709551616.0f) return SIZE_MAX;
    return (size_t)f;
}

/* `f as u8`: saturating truncation, NaN -> 0. */
static uint8_t sat_u8(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 255.0f) return 255;
    return (uint8_t)f;
}

/* f32::min / f32::max: the non-NaN operand wins (IEEE minNum/maxNum). */
static float rmin(float a, float b) { return fminf(a, b); }
static float rmax(float a, float b) { return fmaxf(a, b); }

/* f32::fract = self - self.trunc() */
static float rfract(float x) { return x - truncf(x); }

static size_t uclamp(size_t v, size_t lo, size_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ------------------------------------------------------------------------- */
/* triangles.rs:309-367 helpers                                               */
/* ------------------------------------------------------------------------- */

static void cross3(const float a[3], const float b[3], float r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static void sub3(const float a[3], const float b[3], float r[3]) {
    r[0] = a[0] - b[0];
    r[1] = a[1] - b[1];
    r[2] = a[2] - b[2];
}
static void add3(const float a[3], const float b[3], float r[3]) {
    r[0] = a[0] + b[0];
    r[1] = a[1] + b[1];
    r[2] = a[2] + b[2];
}
static float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static float magsq3(const float v[3]) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
static void matmul33(float a[3][3], float b[3][3], float r[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
}
static void mat3_mul_vec3(float m[3][3], const float v[3], float r[3]) {
    r[0] = m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2];
    r[1] = m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2];
    r[2] = m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2];
}
static float lerp1(float a, float b, float t) { return a + t * (b - a); }

/* ------------------------------------------------------------------------- */
/* Fixed shader set                                                           */
/* ------------------------------------------------------------------------- */

typedef struct {
    float pos[4];      /* homogeneous clip position returned by vertex()     */
    float vd[MAX_VD];  /* VertexData                                          */
} SVert;

typedef struct {
    float c[4];        /* Rgba<f32> fragment (or FLAT colour bits in c[0])    */
} Frag;

typedef struct {
    const orc_desc *d;
    const float *u;
    const orc_stream *s;
    const orc_texture *tex;
    const orc_texture *shadow;
    int nvd;           /* VertexData component count                         */
    int frag_kind;     /* 0 = Rgba<f32>, 1 = flat RGBA8 newtype, 2 = Unit     */
} Pipe;

/* row-major 4x4 times column vector, left-to-right sums */
static void mat4_vec4(const float *m, const float v[4], float r[4]) {
    for (int i = 0; i < 4; i++)
        r[i] = ((m[4 * i + 0] * v[0] + m[4 * i + 1] * v[1]) + m[4 * i + 2] * v[2]) + m[4 * i + 3] * v[3];
}

static void normalize3(const float v[3], float r[3]) {
    float l2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    float s = 1.0f / sqrtf(l2);
    r[0] = v[0] * s;
    r[1] = v[1] * s;
    r[2] = v[2] * s;
}

static float dot3s(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

/* x^30 by a fixed multiplication chain */
static float pow30(float x) {
    float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
    float x24 = x16 * x8, x28 = x24 * x4;
    return x28 * x2;
}

/* Two directional lights, ambient + diffuse + specular (examples/teapot.rs:107-131 model). */
void orc_lighting2(const float *u, const float wpos[3], const float nrm[3], float out[4]) {
    float n[3], vv[3], v[3];
    normalize3(nrm, n);
    vv[0] = u[U_CAM + 0] - wpos[0];
    vv[1] = u[U_CAM + 1] - wpos[1];
    vv[2] = u[U_CAM + 2] - wpos[2];
    normalize3(vv, v);
    float light = u[U_K + 0];
    for (int i = 0; i < 2; i++) {
        const float *L = u + (i == 0 ? U_L0 : U_L1);
        float ndl = dot3s(n, L);
        float diff = rmax(ndl, 0.0f) * u[U_K + 1];
        float t = ndl * 2.0f;
        float r[3] = {t * n[0] - L[0], t * n[1] - L[1], t * n[2] - L[2]};
        float rv = rmax(dot3s(r, v), 0.0f);
        float spec = pow30(rv) * u[U_K + 2];
        light = light + (diff + spec) * L[3];
    }
    for (int c = 0; c < 4; c++) out[c] = u[U_SURF + c] * light;
}

/* Texture::read of the u8->f32 mapped RGBA8 texture or an f32 depth map */
static void tex_read(const orc_texture *t, size_t x, size_t y, float *out) {
    size_t i = x + (size_t)t->w * y;
    if (t->is_f32) {
        out[0] = ((const float *)t->data)[i];
    } else {
        uint32_t p = ((const uint32_t *)t->data)[i];
        for (int c = 0; c < 4; c++) out[c] = (float)((p >> (8 * c)) & 0xffu);
    }
}

/* sampler/linear.rs:30-66, wrapped by Clamped (sampler/mod.rs:183-186) */
void orc_sample(const orc_texture *t, float x, float y, int32_t clamped, float *out) {
    if (clamped) {
        x = rmin(rmax(x, 0.0f), 1.0f);
        y = rmin(rmax(y, 0.0f), 1.0f);
    }
    size_t w = (size_t)t->w, h = (size_t)t->h;
    float itx = rfract(x) * (float)w;
    float ity = rfract(y) * (float)h;
    size_t px = sat_usize(truncf(itx));
    size_t py = sat_usize(truncf(ity));
    float fx = rfract(itx);
    float fy = rfract(ity);
    size_t p0x = (px + 0) < w - 1 ? px + 0 : w - 1;
    size_t p0y = (py + 0) < h - 1 ? py + 0 : h - 1;
    size_t p1x = (px + 1) < w - 1 ? px + 1 : w - 1;
    size_t p1y = (py + 1) < h - 1 ? py + 1 : h - 1;
    int nc = t->is_f32 ? 1 : 4;
    float t00[4], t10[4], t01[4], t11[4];
    tex_read(t, p0x, p0y, t00);
    tex_read(t, p1x, p0y, t10);
    tex_read(t, p0x, p1y, t01);
    tex_read(t, p1x, p1y, t11);
    for (int c = 0; c < nc; c++) {
        float t0 = t00[c] * (1.0f - fy) + t01[c] * fy;
        float t1 = t10[c] * (1.0f - fy) + t11[c] * fy;
        out[c] = t0 * (1.0f - fx) + t1 * fx;
    }
}

/* f32::rem_euclid: r = x % m (exact), then r + |m| when r < 0 */
float orc_rem_euclid(float x, float m) {
    float r = fmodf(x, m);
    return r < 0.0f ? r + fabsf(m) : r;
}

/* Denormalize f32 -> usize (math.rs:52-55): ((u * size as f32).max(0.0) as usize).min(size - 1) */
static size_t denormalize(float u, size_t size) {
    size_t v = sat_usize(rmax(u * (float)size, 0.0f));
    return v < size - 1 ? v : size - 1;
}

/* The index map of a sampler wrapper (sampler/mod.rs:98-179): none, Clamped,
 * Tiled, Mirrored. */
static float wrap_index(int wrap, float x) {
    switch (wrap) {
    case 1: return rmin(rmax(x, 0.0f), 1.0f);
    case 2: return orc_rem_euclid(x, 1.0f);
    case 3: return orc_rem_euclid(x, 2.0f) >= 1.0f ? 1.0f - orc_rem_euclid(x, 1.0f) : orc_rem_euclid(x, 1.0f);
    }
    return x;
}

/* Sampler `kind` (ORC_SAMPLER_LINEAR .. ORC_SAMPLER_NEAREST_MIRRORED): the
 * wrapper's index map, then Linear (sampler/linear.rs:30-66) or Nearest
 * (sampler/nearest.rs:20-38). */
void orc_sample_with(const orc_texture *t, int32_t kind, float u, float v, float *out) {
    const int nearest = kind >= ORC_SAMPLER_NEAREST;
    const int wrap = nearest ? kind - ORC_SAMPLER_NEAREST : kind - ORC_SAMPLER_LINEAR;
    u = wrap_index(wrap, u);
    v = wrap_index(wrap, v);
    if (nearest)
        tex_read(t, denormalize(u, (size_t)t->w), denormalize(v, (size_t)t->h), out);
    else
        orc_sample(t, u, v, 0, out);
}

static const float *vattr(const float *base, int comps, int64_t vi) { return base + (size_t)comps * (size_t)vi; }

/* Pipeline::vertex for the fixed shader set. */
static void vertex_shader(const Pipe *p, int32_t inst, int64_t vi, SVert *out) {
    const orc_stream *s = p->s;
    const float *u = p->u;
    memset(out, 0, sizeof(*out));
    const float *P = vattr(s->pos, s->pos_comps, vi);
    float wpos[4] = {0, 0, 0, 0}, wn[4] = {0, 0, 0, 0};
    const float *M = s->models ? s->models + 16 * (size_t)inst : NULL;
    if (p->d->xform == ORC_XFORM_CLIP) {
        for (int c = 0; c < 4; c++) out->pos[c] = s->pos_comps == 4 ? P[c] : (c < 3 ? P[c] : 1.0f);
        for (int c = 0; c < 3; c++) wpos[c] = out->pos[c];
    } else {
        float p4[4] = {P[0], P[1], P[2], 1.0f};
        mat4_vec4(M, p4, wpos);
        mat4_vec4(u + U_VP, wpos, out->pos);
    }
    if (s->nrm) {
        const float *N = vattr(s->nrm, 3, vi);
        float n4[4] = {N[0], N[1], N[2], 0.0f};
        if (p->d->shader == ORC_SHADER_PHONG_SHADOW) {
            n4[0] = -N[0];
            n4[1] = -N[1];
            n4[2] = -N[2];
        }
        if (p->d->xform == ORC_XFORM_CLIP)
            memcpy(wn, n4, sizeof(wn));
        else
            mat4_vec4(M, n4, wn);
    }
    switch (p->d->shader) {
    case ORC_SHADER_FLAT: {
        uint32_t c = s->rgba8[vi];
        memcpy(&out->vd[0], &c, 4);
        break;
    }
    case ORC_SHADER_COLOR:
        for (int c = 0; c < 4; c++) out->vd[c] = vattr(s->col, 4, vi)[c];
        break;
    case ORC_SHADER_GOURAUD:
        orc_lighting2(u, wpos, wn, out->vd);
        break;
    case ORC_SHADER_PHONG:
        for (int c = 0; c < 3; c++) out->vd[c] = wpos[c], out->vd[3 + c] = wn[c];
        break;
    case ORC_SHADER_DEPTH_ONLY:
        break;
    case ORC_SHADER_TEXTURED:
        out->vd[0] = vattr(s->uv, 2, vi)[0];
        out->vd[1] = vattr(s->uv, 2, vi)[1];
        break;
    case ORC_SHADER_PHONG_SHADOW:
    case ORC_SHADER_TEX_SHADOW: {
        float w1[4] = {wpos[0], wpos[1], wpos[2], 1.0f}, l4[4];
        mat4_vec4(u + U_LVP, w1, l4);
        for (int c = 0; c < 3; c++) out->vd[c] = wpos[c], out->vd[3 + c] = wn[c];
        int o = 6;
        if (p->d->shader == ORC_SHADER_TEX_SHADOW) {
            out->vd[6] = vattr(s->uv, 2, vi)[0];
            out->vd[7] = vattr(s->uv, 2, vi)[1];
            o = 8;
        }
        out->vd[o + 0] = l4[0] / l4[3];
        out->vd[o + 1] = l4[1] / l4[3];
        out->vd[o + 2] = l4[2] / l4[3];
        break;
    }
    }
}

/* Shadow-map lookup of examples/teapot.rs:123-128 */
static int in_light(const Pipe *p, const float lvp[3]) {
    float uvs[2] = {(lvp[0] * 1.0f) * 0.5f + 0.5f, (lvp[1] * -1.0f) * 0.5f + 0.5f};
    float ld;
    orc_sample(p->shadow, uvs[0], uvs[1], 1, &ld);
    ld = ld + p->u[U_K + 3];
    return lvp[2] < ld;
}

/* Pipeline::fragment for the fixed shader set. */
static void fragment_shader(const Pipe *p, const float *vd, Frag *f) {
    const float *u = p->u;
    memset(f, 0, sizeof(*f));
    switch (p->d->shader) {
    case ORC_SHADER_FLAT:
        f->c[0] = vd[0];
        break;
    case ORC_SHADER_COLOR:
    case ORC_SHADER_GOURAUD:
        for (int c = 0; c < 4; c++) f->c[c] = vd[c];
        break;
    case ORC_SHADER_PHONG:
        orc_lighting2(u, vd, vd + 3, f->c);
        break;
    case ORC_SHADER_DEPTH_ONLY:
        break;
    case ORC_SHADER_TEXTURED:
        orc_sample_with(p->tex, p->d->sampler ? p->d->sampler : ORC_SAMPLER_LINEAR, vd[0], vd[1], f->c);
        break;
    case ORC_SHADER_PHONG_SHADOW: {
        /* examples/teapot.rs:99-132 with an exact x^30 chain */
        float wn[3], cd[3], ld[3], t3[3];
        normalize3(vd + 3, wn);
        for (int c = 0; c < 3; c++) t3[c] = u[U_CAM + c] - vd[c];
        normalize3(t3, cd);
        for (int c = 0; c < 3; c++) t3[c] = vd[c] - u[U_LPOS + c];
        normalize3(t3, ld);
        float nl[3] = {-ld[0], -ld[1], -ld[2]};
        float ambient = u[U_K + 0];
        float diffuse = rmax(dot3s(wn, nl), 0.0f) * u[U_K + 1];
        /* (-light_dir).reflected(wnorm) = v - n * (2 * dot(v, n)) */
        float dn = dot3s(nl, wn) * 2.0f;
        float r[3] = {nl[0] - wn[0] * dn, nl[1] - wn[1] * dn, nl[2] - wn[2] * dn};
        float ncd[3] = {-cd[0], -cd[1], -cd[2]};
        float specular = pow30(rmax(dot3s(r, ncd), 0.0f)) * u[U_K + 2];
        float light = ambient + (in_light(p, vd + 6) ? diffuse + specular : 0.0f);
        for (int c = 0; c < 4; c++) f->c[c] = u[U_SURF + c] * light;
        break;
    }
    case ORC_SHADER_TEX_SHADOW: {
        float n[3], texc[4];
        normalize3(vd + 3, n);
        const float *L = u + U_L0;
        float ndl = rmax(dot3s(n, L), 0.0f);
        float light = u[U_K + 0] + (in_light(p, vd + 8) ? ndl * u[U_K + 1] : 0.0f);
        orc_sample_with(p->tex, p->d->sampler ? p->d->sampler : ORC_SAMPLER_LINEAR_CLAMPED, vd[6], vd[7], texc);
        float s = light * (1.0f / 255.0f);
        f->c[0] = texc[0] * s;
        f->c[1] = texc[1] * s;
        f->c[2] = texc[2] * s;
        f->c[3] = texc[3] * (1.0f / 255.0f);
        break;
    }
    }
}

/* Pipeline::blend (all fixed shaders replace the old pixel). */
static uint32_t blend_shader(const Pipe *p, uint32_t old, const Frag *f) {
    if (p->frag_kind == 2) return old; /* Unit fragment: the pixel is left as it was */
    if (p->frag_kind == 1) {
        uint32_t c;
        memcpy(&c, &f->c[0], 4);
        return c;
    }
    uint8_t b[4];
    if (p->d->shader == ORC_SHADER_TEXTURED) {
        for (int c = 0; c < 4; c++) b[c] = sat_u8(f->c[c]);
    } else {
        for (int c = 0; c < 4; c++) b[c] = sat_u8(rmin(rmax(f->c[c], 0.0f), 1.0f) * 255.0f);
    }
    if (p->d->pixel_format == ORC_PIXEL_BGRA8) { /* u32::from_le_bytes([b, g, r, a]) */
        const uint8_t r = b[0];
        b[0] = b[2];
        b[2] = r;
    }
    return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
}

/* math.rs:37-40 weighted_sum3 = v0*w0 + v1*w1 + v2*w2; the FLAT newtype keeps v0. */
static void vd_weighted_sum3(const Pipe *p, const float *v0, const float *v1, const float *v2, float w0, float w1,
                             float w2, float *r) {
    if (p->frag_kind == 1) {
        r[0] = v0[0];
        return;
    }
    for (int c = 0; c < p->nvd; c++) r[c] = v0[c] * w0 + v1[c] * w1 + v2[c] * w2;
}

/* math.rs:34-36 weighted_sum2 = v0*w0 + v1*w1 (FLAT keeps v0, Unit is Unit). */
static void frag_weighted_sum2(const Pipe *p, const Frag *a, const Frag *b, float w0, float w1, Frag *r) {
    if (p->frag_kind == 1) {
        *r = *a;
        return;
    }
    if (p->frag_kind == 2) {
        memset(r, 0, sizeof(*r));
        return;
    }
    for (int c = 0; c < 4; c++) r->c[c] = a->c[c] * w0 + b->c[c] * w1;
}

/* ------------------------------------------------------------------------- */
/* BlitterImpl (pipeline.rs:452-579)                                          */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint64_t prim;
    Frag frag;
} MsaaTexel;

typedef struct {
    const Pipe *pipe;
    int write_pixels;
    int depth_test, depth_write;
    size_t tgt_min[2], tgt_max[2], tgt_size[2];
    uint32_t *pixel;
    size_t pixel_w;  /* Buffer2d size[0] of the pixel target */
    float *depth;
    size_t depth_w;
    uint64_t primitive_count;
    size_t msaa_level;
    MsaaTexel *msaa_buf;
    size_t msaa_w;
    float msaa_div;
    int64_t frags;
} Blitter;

/* Interpolation state captured by the get_v_data closures of triangles.rs:274-294
 * and, with `line` set, lines.rs:99-112. */
typedef struct {
    float o[3], dx[3], dy[3];
    const float *vout[3];
    int line, use_x;
    float vs0x, vs0y, norm;
} VData;

static void get_v_data(const Blitter *b, const VData *g, float x, float y, float *vd) {
    if (g->line) {
        /* frac = (if use_x { x - vs0.x } else { y - vs0.y }) * norm;
         * V::weighted_sum2(v0, v1, 1.0 - frac, frac) = v0 * w0 + v1 * w1 (math.rs:34-36) */
        float frac = (g->use_x ? x - g->vs0x : y - g->vs0y) * g->norm;
        float w0 = 1.0f - frac;
        if (b->pipe->frag_kind == 1) {
            vd[0] = g->vout[0][0];
            return;
        }
        for (int c = 0; c < b->pipe->nvd; c++) vd[c] = g->vout[0][c] * w0 + g->vout[1][c] * frac;
        return;
    }
    float t1[3], wh[3];
    float dyy[3] = {g->dy[0] * y, g->dy[1] * y, g->dy[2] * y};
    float dxx[3] = {g->dx[0] * x, g->dx[1] * x, g->dx[2] * x};
    add3(g->o, dyy, t1);
    add3(t1, dxx, wh);
    float wu[3] = {wh[0], wh[1], wh[2] - wh[0] - wh[1]};
    float r = 1.0f / wh[2];
    float w[3] = {wu[0] * r, wu[1] * r, wu[2] * r};
    vd_weighted_sum3(b->pipe, g->vout[0], g->vout[1], g->vout[2], w[0], w[1], w[2], vd);
}

static int test_fragment(Blitter *b, size_t x, size_t y, float z) {
    if (b->depth_test == ORC_DEPTH_NONE) return 1;
    float old = b->depth ? b->depth[b->depth_w * y + x] : 0.0f; /* Empty reads T::default() */
    switch (b->depth_test) {
    case ORC_DEPTH_LESS: return z < old;
    case ORC_DEPTH_EQUAL: return z == old;
    case ORC_DEPTH_GREATER: return z > old;
    }
    return 0;
}

static void msaa_fragment(Blitter *b, const VData *g, size_t x, size_t y, Frag *out) {
    MsaaTexel *t = &b->msaa_buf[(x + 1) + b->msaa_w * (y + 1)];
    if (t->prim != b->primitive_count) {
        t->prim = b->primitive_count;
        float vd[MAX_VD];
        get_v_data(b, g, (float)(b->tgt_min[0] + (x << b->msaa_level)),
                   (float)(b->tgt_min[1] + (y << b->msaa_level)), vd);
        fragment_shader(b->pipe, vd, &t->frag);
    }
    *out = t->frag;
}

static void emit_fragment(Blitter *b, size_t x, size_t y, const VData *g, float z) {
    b->frags++;
    if (b->depth_write && b->depth) b->depth[b->depth_w * y + x] = z;
    if (b->write_pixels) {
        Frag frag;
        if (b->msaa_level == 0) {
            float vd[MAX_VD];
            get_v_data(b, g, (float)x, (float)y, vd);
            fragment_shader(b->pipe, vd, &frag);
        } else {
            float fractx = rfract((float)(x - b->tgt_min[0]) * b->msaa_div);
            float fracty = rfract((float)(y - b->tgt_min[1]) * b->msaa_div);
            size_t posix = (x - b->tgt_min[0]) >> b->msaa_level;
            size_t posiy = (y - b->tgt_min[1]) >> b->msaa_level;
            Frag t00, t10, t01, t11, t0, t1;
            msaa_fragment(b, g, posix + 0, posiy + 0, &t00);
            msaa_fragment(b, g, posix + 1, posiy + 0, &t10);
            msaa_fragment(b, g, posix + 0, posiy + 1, &t01);
            msaa_fragment(b, g, posix + 1, posiy + 1, &t11);
            frag_weighted_sum2(b->pipe, &t00, &t01, 1.0f - fracty, fracty, &t0);
            frag_weighted_sum2(b->pipe, &t10, &t11, 1.0f - fracty, fracty, &t1);
            frag_weighted_sum2(b->pipe, &t0, &t1, 1.0f - fractx, fractx, &frag);
        }
        if (b->pixel) {
            size_t i = b->pixel_w * y + x;
            b->pixel[i] = blend_shader(b->pipe, b->pixel[i], &frag);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Triangles::rasterize (triangles.rs:15-306)                                 */
/* ------------------------------------------------------------------------- */

typedef struct {
    int y_up, z_clip;
    float z_min, z_max;
} Coords;

static int passes_z_clip(const Coords *c, float z) { return !c->z_clip || (c->z_min <= z && z <= c->z_max); }

typedef struct {
    float M[3][3];
    float o[3], dx[3], dy[3];
    size_t bmin[2], bmax[2];
    float vby[3][2];
    float hom[3][4];
    int allin, reversed, culled;
    float winding;
} Setup;

/* triangles.rs:55-171 for one triangle; returns 0 if culled. */
static int setup_triangle(const float hom_in[3][4], size_t size_w, size_t size_h, const size_t tgt_min[2],
                          const size_t tgt_max[2], const Coords *coords, int cull_mode, Setup *S, int order[3]) {
    float size_x = (float)size_w, size_y = (float)size_h;
    float to_ndc[3][3] = {{2.0f / size_x, 0.0f, -1.0f}, {0.0f, -2.0f / size_y, 1.0f}, {0.0f, 0.0f, 1.0f}};
    float flip0 = 1.0f, flip1 = coords->y_up ? -1.0f : 1.0f;
    float hom[3][4], euc[3][3];
    for (int i = 0; i < 3; i++) {
        hom[i][0] = hom_in[i][0] * flip0;
        hom[i][1] = hom_in[i][1] * flip1;
        hom[i][2] = hom_in[i][2];
        hom[i][3] = hom_in[i][3];
    }
    for (int i = 0; i < 3; i++) {
        euc[i][0] = hom[i][0] / hom[i][3];
        euc[i][1] = hom[i][1] / hom[i][3];
        euc[i][2] = hom[i][2] / hom[i][3];
    }
    float e10[3], e20[3], cr[3];
    sub3(euc[1], euc[0], e10);
    sub3(euc[2], euc[0], e20);
    cross3(e10, e20, cr);
    float winding = cr[2];
    S->winding = winding;
    S->culled = 0;
    S->reversed = 0;
    order[0] = 0, order[1] = 1, order[2] = 2;
    float cull_dir = cull_mode == ORC_CULL_BACK ? 1.0f : -1.0f;
    if (cull_mode != ORC_CULL_NONE && winding * cull_dir < 0.0f) {
        S->culled = 1;
        return 0;
    } else if (winding >= 0.0f) {
        float th[3][4], te[3][3];
        memcpy(th, hom, sizeof(th));
        memcpy(te, euc, sizeof(te));
        for (int i = 0; i < 3; i++) {
            memcpy(hom[i], th[2 - i], sizeof(hom[i]));
            memcpy(euc[i], te[2 - i], sizeof(euc[i]));
        }
        order[0] = 2, order[1] = 1, order[2] = 0;
        S->reversed = 1;
    }
    /* coords_to_weights (triangles.rs:86-102) */
    {
        float c[3] = {hom[2][0], hom[2][1], hom[2][3]};
        float a3[3] = {hom[0][0], hom[0][1], hom[0][3]};
        float b3[3] = {hom[1][0], hom[1][1], hom[1][3]};
        float ca[3], cb[3], n[3];
        sub3(a3, c, ca);
        sub3(b3, c, cb);
        cross3(ca, cb, n);
        float rec_det = magsq3(n) > 0.0f ? 1.0f / rmin(dot3(n, c), -FLT_EPSILON) : 1.0f;
        float rows[3][3];
        cross3(cb, c, rows[0]);
        cross3(c, ca, rows[1]);
        memcpy(rows[2], n, sizeof(n));
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) rows[i][j] = rows[i][j] * rec_det;
        matmul33(rows, to_ndc, S->M);
    }
    /* verts_screen (triangles.rs:110-111) */
    float vs[3][2];
    for (int i = 0; i < 3; i++) {
        vs[i][0] = size_x * (euc[i][0] * 0.5f + 0.5f);
        vs[i][1] = size_y * (euc[i][1] * -0.5f + 0.5f);
    }
    /* bounds (triangles.rs:114-139) */
    S->bmin[0] = uclamp(sat_usize(rmin(rmin(vs[0][0], vs[1][0]), vs[2][0]) + 0.0f), tgt_min[0], tgt_max[0]);
    S->bmin[1] = uclamp(sat_usize(rmin(rmin(vs[0][1], vs[1][1]), vs[2][1]) + 0.0f), tgt_min[1], tgt_max[1]);
    S->bmax[0] = uclamp(sat_usize(rmax(rmax(vs[0][0], vs[1][0]), vs[2][0]) + 1.0f), tgt_min[0], tgt_max[0]);
    S->bmax[1] = uclamp(sat_usize(rmax(rmax(vs[0][1], vs[1][1]), vs[2][1]) + 1.0f), tgt_min[1], tgt_max[1]);
    /* weights (triangles.rs:142-145) */
    {
        float p0[3] = {0.0f, 0.0f, 1.0f}, px[3] = {1000.0f, 0.0f, 1.0f}, py[3] = {0.0f, 1000.0f, 1.0f};
        float wx[3], wy[3], t[3];
        mat3_mul_vec3(S->M, p0, S->o);
        mat3_mul_vec3(S->M, px, wx);
        mat3_mul_vec3(S->M, py, wy);
        sub3(wx, S->o, t);
        for (int i = 0; i < 3; i++) S->dx[i] = t[i] * (1.0f / 1000.0f);
        sub3(wy, S->o, t);
        for (int i = 0; i < 3; i++) S->dy[i] = t[i] * (1.0f / 1000.0f);
    }
    /* verts_by_y (triangles.rs:148-171) */
    {
        float min_y = rmin(rmin(vs[0][1], vs[1][1]), vs[2][1]);
        int a, b, c;
        if (vs[0][1] == min_y) {
            if (vs[1][1] < vs[2][1]) a = 0, b = 1, c = 2;
            else a = 0, b = 2, c = 1;
        } else if (vs[1][1] == min_y) {
            if (vs[0][1] < vs[2][1]) a = 1, b = 0, c = 2;
            else a = 1, b = 2, c = 0;
        } else {
            if (vs[0][1] < vs[1][1]) a = 2, b = 0, c = 1;
            else a = 2, b = 1, c = 0;
        }
        memcpy(S->vby[0], vs[a], sizeof(vs[a]));
        memcpy(S->vby[1], vs[b], sizeof(vs[b]));
        memcpy(S->vby[2], vs[c], sizeof(vs[c]));
    }
    S->allin = passes_z_clip(coords, euc[0][2]) && passes_z_clip(coords, euc[1][2]) &&
               passes_z_clip(coords, euc[2][2]);
    memcpy(S->hom, hom, sizeof(hom));
    return 1;
}

/* triangles.rs:203-304 */
static void rasterize_rows(Blitter *b, const Coords *coords, const Setup *S, const VData *g) {
    for (size_t y = S->bmin[1]; y < S->bmax[1]; y++) {
        size_t extent0 = S->bmax[0] - S->bmin[0];
        size_t extent1 = S->bmax[1] - S->bmin[1];
        size_t x0, x1;
        if (extent0 * extent1 < 128) {
            x0 = S->bmin[0];
            x1 = S->bmax[0];
        } else {
            const float *a = S->vby[0], *bb = S->vby[1], *c = S->vby[2];
            float Y = (float)y;
            float ac = lerp1(a[0], c[0], (Y - a[1]) / (c[1] - a[1]));
            float lo, hi;
            if (Y < bb[1]) {
                float ab = lerp1(a[0], bb[0], (Y - a[1]) / (bb[1] - a[1]));
                lo = rmin(ab, ac);
                hi = rmax(ab, ac);
            } else {
                float bc = lerp1(bb[0], c[0], (Y - bb[1]) / (c[1] - bb[1]));
                lo = rmin(bc, ac);
                hi = rmax(bc, ac);
            }
            float fl = floorf(lo), ce = ceilf(hi);
            x0 = (fl >= (float)S->bmin[0] && fl < (float)S->bmax[0]) ? sat_usize(fl) : S->bmin[0];
            x1 = (ce >= (float)S->bmin[0] && ce < (float)S->bmax[0]) ? sat_usize(ce) : S->bmax[0];
        }
        float w[3], t1[3];
        float Y = (float)y, X0 = (float)x0;
        float dyy[3] = {S->dy[0] * Y, S->dy[1] * Y, S->dy[2] * Y};
        float dxx[3] = {S->dx[0] * X0, S->dx[1] * X0, S->dx[2] * X0};
        add3(S->o, dyy, t1);
        add3(t1, dxx, w);
        for (size_t x = x0; x < x1; x++) {
            float wu[3] = {w[0], w[1], w[2] - w[0] - w[1]};
            if (wu[0] >= 0.0f && wu[1] >= 0.0f && wu[2] >= 0.0f) {
                float hz[3] = {S->hom[0][2], S->hom[1][2], S->hom[2][2]};
                float z = dot3(hz, wu);
                if ((S->allin || passes_z_clip(coords, z)) && test_fragment(b, x, y, z)) emit_fragment(b, x, y, g, z);
            }
            float nw[3];
            add3(w, S->dx, nw);
            memcpy(w, nw, sizeof(w));
        }
    }
}

/* `f as isize` (saturating, NaN -> 0). */
static int64_t sat_isize(float f) {
    if (f != f) return 0;
    if (f >= 9223372036854775808.0f) return INT64_MAX;
    if (f <= -9223372036854775808.0f) return INT64_MIN;
    return (int64_t)f;
}

/* f32::clamp (NaN stays NaN). */
static float fclamp(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

/* Line endpoints further than this from the origin (in pixels) are not drawn:
 * the clipline crate's isize arithmetic is not defined there. */
#define LINE_COORD_LIMIT ((int64_t)1 << 29)

/* The Bresenham line (x1, y1) -> (x2, y2), both ends included, restricted to the
 * rectangle [cx0, cx1] x [cy0, cy1] -- the contract of `clipline::clipline`
 * (clipline 0.2, not vendored: parity unpinned at this boundary): clipping never
 * moves a pixel.  Along the major axis (the longer |delta|) step k's minor
 * coordinate is m1 + sign * floor((2k|dm| + |dM| - 1) / (2|dM|)), the line's
 * value rounded to nearest with halves rounded back towards the start (the
 * decision `e > 0` of an error term e = 2|dm| - |dM|).  Calls plot(x, y). */
typedef void (*PlotFn)(void *ctx, int64_t x, int64_t y);

static void clip_line(int64_t x1, int64_t y1, int64_t x2, int64_t y2, int64_t cx0, int64_t cy0, int64_t cx1,
                      int64_t cy1, PlotFn plot, void *ctx) {
    if (cx1 < cx0 || cy1 < cy0) return;
    const int64_t dx = x2 - x1, dy = y2 - y1;
    const int64_t adx = dx < 0 ? -dx : dx, ady = dy < 0 ? -dy : dy;
    const int xmajor = adx >= ady;
    const int64_t M1 = xmajor ? x1 : y1, m1 = xmajor ? y1 : x1;
    const int64_t dM = xmajor ? adx : ady, dm = xmajor ? ady : adx;
    const int64_t sM = (xmajor ? dx : dy) < 0 ? -1 : 1, sm = (xmajor ? dy : dx) < 0 ? -1 : 1;
    const int64_t c0 = xmajor ? cx0 : cy0, c1 = xmajor ? cx1 : cy1;
    const int64_t n0 = xmajor ? cy0 : cx0, n1 = xmajor ? cy1 : cx1;
    /* steps k in [0, dM] whose major coordinate M1 + sM k lies in [c0, c1] */
    int64_t klo = sM > 0 ? c0 - M1 : M1 - c1, khi = sM > 0 ? c1 - M1 : M1 - c0;
    if (klo < 0) klo = 0;
    if (khi > dM) khi = dM;
    for (int64_t k = klo; k <= khi; k++) {
        const int64_t j = dM ? (2 * k * dm + dM - 1) / (2 * dM) : 0;
        const int64_t M = M1 + sM * k, m = m1 + sm * j;
        if (m < n0 || m > n1) continue;
        if (xmajor) plot(ctx, M, m);
        else plot(ctx, m, M);
    }
}

typedef struct {
    Blitter *b;
    const Coords *coords;
    VData *g;
    float e0z, e1z;
} LinePlot;

static void line_pixel(void *ctx, int64_t xi, int64_t yi) {
    LinePlot *L = (LinePlot *)ctx;
    const size_t x = (size_t)xi, y = (size_t)yi;
    /* frac of lines.rs:88-92, z of :95 */
    float frac = (L->g->use_x ? (float)x - L->g->vs0x : (float)y - L->g->vs0y) * L->g->norm;
    float z = L->e0z + frac * (L->e1z - L->e0z);
    if (passes_z_clip(L->coords, z) && test_fragment(L->b, x, y, z)) emit_fragment(L->b, x, y, L->g, z);
}

/* Lines::rasterize (lines.rs:15-121) of the pairs of the stream `vidx`. */
static void rasterize_lines(Blitter *b, const Coords *coords, const SVert *verts, const int64_t *vidx,
                            int64_t nverts) {
    const float size0 = (float)b->tgt_size[0], size1 = (float)b->tgt_size[1];
    const float flip1 = coords->y_up ? -1.0f : 1.0f;
    const float smin0 = (float)b->tgt_min[0], smin1 = (float)b->tgt_min[1];
    const float smax0 = (float)b->tgt_max[0], smax1 = (float)b->tgt_max[1];
    for (int64_t l = 0; l + 1 < nverts; l += 2) {
        b->primitive_count = b->primitive_count + 1;
        const SVert *v[2] = {&verts[vidx[l]], &verts[vidx[l + 1]]};
        float euc[2][3], vs[2][2];
        for (int i = 0; i < 2; i++) {
            const float hom[4] = {v[i]->pos[0] * 1.0f, v[i]->pos[1] * flip1, v[i]->pos[2], v[i]->pos[3]};
            const float w = rmax(hom[3], 0.0001f);
            euc[i][0] = hom[0] / w, euc[i][1] = hom[1] / w, euc[i][2] = hom[2] / w;
            vs[i][0] = size0 * (euc[i][0] * 0.5f + 0.5f);
            vs[i][1] = size1 * (euc[i][1] * -0.5f + 0.5f);
        }
        const int64_t x1 = sat_isize(vs[0][0]), y1 = sat_isize(vs[0][1]);
        const int64_t x2 = sat_isize(vs[1][0]), y2 = sat_isize(vs[1][1]);
        const int64_t wx1 = sat_isize(fclamp(rmin(vs[0][0], vs[1][0]) + 0.0f, smin0, smax0));
        const int64_t wy1 = sat_isize(fclamp(rmin(vs[0][1], vs[1][1]) + 0.0f, smin1, smax1));
        const int64_t wx2 = sat_isize(fclamp(rmax(vs[0][0], vs[1][0]) + 1.0f, smin0, smax0));
        const int64_t wy2 = sat_isize(fclamp(rmax(vs[0][1], vs[1][1]) + 1.0f, smin1, smax1));
        if (x1 <= -LINE_COORD_LIMIT || x1 >= LINE_COORD_LIMIT || y1 <= -LINE_COORD_LIMIT || y1 >= LINE_COORD_LIMIT ||
            x2 <= -LINE_COORD_LIMIT || x2 >= LINE_COORD_LIMIT || y2 <= -LINE_COORD_LIMIT || y2 >= LINE_COORD_LIMIT)
            continue;
        const int64_t adx = x1 - x2 < 0 ? x2 - x1 : x1 - x2, ady = y1 - y2 < 0 ? y2 - y1 : y1 - y2;
        VData g;
        memset(&g, 0, sizeof(g));
        g.line = 1;
        g.use_x = adx > ady;
        g.vs0x = vs[0][0], g.vs0y = vs[0][1];
        g.norm = 1.0f / (g.use_x ? vs[1][0] - vs[0][0] : vs[1][1] - vs[0][1]);
        g.vout[0] = v[0]->vd, g.vout[1] = v[1]->vd;
        LinePlot L = {b, coords, &g, euc[0][2], euc[1][2]};
        clip_line(x1, y1, x2, y2, wx1, wy1, wx2 - 1, wy2 - 1, line_pixel, &L);
    }
}

/* verts: the shaded vertex stream; vidx: the stream after the geometry stage,
 * as indices into verts (pipeline.rs:277-289). */
static void rasterize_all(Blitter *b, const Coords *coords, int cull_mode, const SVert *verts, const int64_t *vidx,
                          int64_t nverts) {
    for (int64_t t = 0; t + 2 < nverts; t += 3) {
        b->primitive_count = b->primitive_count + 1; /* begin_primitive (wrapping_add) */
        float hom_in[3][4];
        for (int i = 0; i < 3; i++) memcpy(hom_in[i], verts[vidx[t + i]].pos, sizeof(hom_in[i]));
        Setup S;
        int order[3];
        if (!setup_triangle(hom_in, b->tgt_size[0], b->tgt_size[1], b->tgt_min, b->tgt_max, coords, cull_mode, &S,
                            order))
            continue;
        VData g;
        memset(&g, 0, sizeof(g));
        memcpy(g.o, S.o, sizeof(g.o));
        memcpy(g.dx, S.dx, sizeof(g.dx));
        memcpy(g.dy, S.dy, sizeof(g.dy));
        for (int i = 0; i < 3; i++) g.vout[i] = verts[vidx[t + order[i]]].vd;
        rasterize_rows(b, coords, &S, &g);
    }
}

/* ------------------------------------------------------------------------- */
/* render_inner (pipeline.rs:396-614) and the schedulers                    */
/* ------------------------------------------------------------------------- */

typedef struct {
    const Pipe *pipe;
    const SVert *verts;
    const int64_t *vidx;   /* the stream after the geometry stage (indices into verts) */
    int64_t nverts;        /* its length */
    size_t tgt_size[2];
    uint32_t *pixel;
    size_t pw;
    float *depth;
    size_t dw;
    size_t msaa_level;
    Coords coords;
    /* render_par */
    size_t group_rows;
    _Atomic size_t row;
    _Atomic int64_t frags;
} RenderJob;

static void render_inner(RenderJob *j, size_t by0, size_t by1) {
    const orc_desc *d = j->pipe->d;
    Blitter b;
    memset(&b, 0, sizeof(b));
    b.pipe = j->pipe;
    b.write_pixels = d->pixel_write;
    b.depth_test = d->depth_test;
    b.depth_write = d->depth_write;
    b.tgt_size[0] = j->tgt_size[0];
    b.tgt_size[1] = j->tgt_size[1];
    b.tgt_min[0] = 0;
    b.tgt_min[1] = by0;
    b.tgt_max[0] = j->tgt_size[0];
    b.tgt_max[1] = by1;
    b.pixel = j->pixel;
    b.pixel_w = j->pw;
    b.depth = j->depth;
    b.depth_w = j->dw;
    b.primitive_count = 0;
    b.msaa_level = j->msaa_level;
    b.msaa_div = 1.0f / (float)(1u << j->msaa_level);
    if (j->msaa_level > 0) {
        b.msaa_w = ((b.tgt_max[0] - b.tgt_min[0]) >> j->msaa_level) + 3;
        size_t mh = ((b.tgt_max[1] - b.tgt_min[1]) >> j->msaa_level) + 3;
        b.msaa_buf = (MsaaTexel *)malloc(sizeof(MsaaTexel) * b.msaa_w * mh);
        for (size_t i = 0; i < b.msaa_w * mh; i++) b.msaa_buf[i].prim = UINT64_MAX;
    }
    if (d->primitives == ORC_PRIMS_TRIANGLE_LIST)
        rasterize_all(&b, &j->coords, d->cull, j->verts, j->vidx, j->nverts);
    else
        rasterize_lines(&b, &j->coords, j->verts, j->vidx, j->nverts);
    free(b.msaa_buf);
    atomic_fetch_add(&j->frags, b.frags);
}

static void *par_worker(void *arg) {
    RenderJob *j = (RenderJob *)arg;
    for (;;) {
        size_t row_start = atomic_fetch_add_explicit(&j->row, j->group_rows, memory_order_relaxed);
        if (row_start >= j->tgt_size[1]) break;
        size_t row_end = row_start + j->group_rows < j->tgt_size[1] ? row_start + j->group_rows : j->tgt_size[1];
        render_inner(j, row_start, row_end);
    }
    return NULL;
}

static int render_impl(const orc_desc *d, const float *u, const orc_stream *s, const orc_texture *tex,
                       const orc_texture *shadow, uint32_t *pixel, int32_t pw, int32_t ph, float *depth, int32_t dw,
                       int32_t dh, int64_t *frags_out) {
    if (frags_out) *frags_out = 0;
    int uses_depth = d->depth_test != ORC_DEPTH_NONE || d->depth_write;
    size_t psz[2] = {pixel ? (size_t)pw : 0, pixel ? (size_t)ph : 0};
    size_t dsz[2] = {depth ? (size_t)dw : 0, depth ? (size_t)dh : 0};
    size_t tgt[2];
    if (!d->pixel_write && !uses_depth) return ORC_OK;
    if (d->pixel_write && !uses_depth) {
        tgt[0] = psz[0], tgt[1] = psz[1];
    } else if (!d->pixel_write && uses_depth) {
        tgt[0] = dsz[0], tgt[1] = dsz[1];
    } else {
        if (psz[0] != dsz[0] || psz[1] != dsz[1]) return ORC_SIZE_MISMATCH;
        tgt[0] = psz[0], tgt[1] = psz[1];
    }
    Pipe p;
    p.d = d, p.u = u, p.s = s, p.tex = tex, p.shadow = shadow;
    p.frag_kind = d->shader == ORC_SHADER_FLAT ? 1 : (d->shader == ORC_SHADER_DEPTH_ONLY ? 2 : 0);
    switch (d->shader) {
    case ORC_SHADER_FLAT: p.nvd = 1; break;
    case ORC_SHADER_COLOR:
    case ORC_SHADER_GOURAUD: p.nvd = 4; break;
    case ORC_SHADER_PHONG: p.nvd = 6; break;
    case ORC_SHADER_DEPTH_ONLY: p.nvd = 0; break;
    case ORC_SHADER_TEXTURED: p.nvd = 2; break;
    case ORC_SHADER_PHONG_SHADOW: p.nvd = 9; break;
    case ORC_SHADER_TEX_SHADOW: p.nvd = 11; break;
    default: return ORC_INVALID;
    }
    if (d->primitives < ORC_PRIMS_TRIANGLE_LIST || d->primitives > ORC_PRIMS_LINE_TRIANGLE_LIST) return ORC_INVALID;
    if (d->primitives == ORC_PRIMS_LINE_LIST && d->geometry != ORC_GEOMETRY_IDENTITY) return ORC_INVALID;
    /* vertex shading of the whole stream (render_par collects first, pipeline.rs:322);
     * a trailing partial primitive is dropped (primitives.rs:28-32, 96-100) */
    const int64_t per = d->primitives == ORC_PRIMS_LINE_LIST ? 2 : 3;
    int64_t total = s->count * (int64_t)s->n_instances;
    int64_t nverts = total - total % per;
    SVert *verts = (SVert *)malloc(sizeof(SVert) * (size_t)(nverts > 0 ? nverts : 1));
    for (int64_t e = 0; e < nverts; e++) {
        int32_t inst = (int32_t)(e / s->count);
        int64_t k = e % s->count;
        int64_t vi = s->idx ? (int64_t)s->idx[k] : k;
        vertex_shader(&p, inst, vi, &verts[e]);
    }
    /* the geometry stage (each triangle [a, b, c] becomes its outputs, in order),
     * then PrimitiveKind::primitive_vertices: a LineTriangleList triangle is the
     * lines a-b, b-c, c-a (primitives.rs:71-84) */
    const int64_t k = d->geometry == ORC_GEOMETRY_BOTH_SIDES ? 2 : 1;
    const int64_t emit = d->primitives == ORC_PRIMS_LINE_TRIANGLE_LIST ? 6 : 3;
    int64_t *vidx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nverts * k * 2 > 0 ? nverts * k * 2 : 1));
    int64_t nv = 0;
    if (d->primitives == ORC_PRIMS_LINE_LIST) {
        for (int64_t e = 0; e < nverts; e++) vidx[nv++] = e;
    } else {
        for (int64_t t = 0; t + 2 < nverts; t += 3) {
            for (int out = 0; out < 2; out++) {
                if (out == 0 && d->geometry == ORC_GEOMETRY_REVERSED) continue;
                if (out == 1 && d->geometry == ORC_GEOMETRY_IDENTITY) continue;
                const int64_t a = t, b = out ? t + 2 : t + 1, c = out ? t + 1 : t + 2;
                if (emit == 3) {
                    vidx[nv++] = a, vidx[nv++] = b, vidx[nv++] = c;
                } else {
                    vidx[nv++] = a, vidx[nv++] = b, vidx[nv++] = b, vidx[nv++] = c, vidx[nv++] = c, vidx[nv++] = a;
                }
            }
        }
    }
    size_t msaa_level = d->aa_msaa ? (size_t)(d->aa_level < 0 ? 0 : (d->aa_level > 6 ? 6 : d->aa_level)) : 0;
    RenderJob j;
    memset(&j, 0, sizeof(j));
    j.pipe = &p;
    j.verts = verts;
    j.vidx = vidx;
    j.nverts = nv;
    j.tgt_size[0] = tgt[0], j.tgt_size[1] = tgt[1];
    j.pixel = pixel, j.pw = psz[0];
    j.depth = depth, j.dw = dsz[0];
    j.msaa_level = msaa_level;
    j.coords.y_up = d->y_up, j.coords.z_clip = d->z_clip, j.coords.z_min = d->z_min, j.coords.z_max = d->z_max;
    int rc = ORC_OK;
    if (d->band_policy == ORC_BANDS_SEQ) {
        render_inner(&j, 0, tgt[1]);
    } else {
        size_t group_rows = (size_t)20000 * ((size_t)1 << msaa_level) / (tgt[0] > 1 ? tgt[0] : 1);
        if (d->group_rows > 0) group_rows = (size_t)d->group_rows;
        if (group_rows == 0) {
            rc = ORC_BAND_ZERO;
        } else {
            size_t threads = d->threads > 0 ? (size_t)d->threads : 1;
            size_t needed = tgt[1] / group_rows;
            if (needed > threads) needed = threads;
            j.group_rows = group_rows;
            atomic_store(&j.row, 0);
            if (needed == 1) {
                par_worker(&j);
            } else if (needed > 1) {
                pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * needed);
                for (size_t i = 0; i < needed; i++) pthread_create(&th[i], NULL, par_worker, &j);
                for (size_t i = 0; i < needed; i++) pthread_join(th[i], NULL);
                free(th);
            }
        }
    }
    if (frags_out) *frags_out = atomic_load(&j.frags);
    free(vidx);
    free(verts);
    return rc;
}

int orc_render(const orc_desc *d, const float *u, const orc_stream *s, const orc_texture *tex,
               const orc_texture *shadow, uint32_t *pixel, int32_t pw, int32_t ph, float *depth, int32_t dw,
               int32_t dh) {
    return render_impl(d, u, s, tex, shadow, pixel, pw, ph, depth, dw, dh, NULL);
}

int orc_render_count(const orc_desc *d, const float *u, const orc_stream *s, const orc_texture *tex,
                     const orc_texture *shadow, uint32_t *pixel, int32_t pw, int32_t ph, float *depth, int32_t dw,
                     int32_t dh, int64_t *frags) {
    return render_impl(d, u, s, tex, shadow, pixel, pw, ph, depth, dw, dh, frags);
}

int orc_setup(const float hom_in[12], int32_t w, int32_t h, int32_t by0, int32_t by1, int32_t y_up, int32_t z_clip,
              float z_min, float z_max, int32_t cull, float *out) {
    float hi[3][4];
    memcpy(hi, hom_in, sizeof(hi));
    size_t tmin[2] = {0, (size_t)by0}, tmax[2] = {(size_t)w, (size_t)by1};
    Coords c = {y_up, z_clip, z_min, z_max};
    Setup S;
    int order[3];
    memset(&S, 0, sizeof(S));
    int ok = setup_triangle(hi, (size_t)w, (size_t)h, tmin, tmax, &c, cull, &S, order);
    float *o = out;
    for (int i = 0; i < 3; i++)
        for (int j2 = 0; j2 < 3; j2++) *o++ = S.M[i][j2];
    for (int i = 0; i < 3; i++) *o++ = S.o[i];
    for (int i = 0; i < 3; i++) *o++ = S.dx[i];
    for (int i = 0; i < 3; i++) *o++ = S.dy[i];
    *o++ = (float)S.bmin[0], *o++ = (float)S.bmin[1];
    *o++ = (float)S.bmax[0], *o++ = (float)S.bmax[1];
    for (int i = 0; i < 3; i++) *o++ = S.vby[i][0], *o++ = S.vby[i][1];
    for (int i = 0; i < 3; i++) *o++ = S.hom[i][2];
    *o++ = (float)S.allin, *o++ = (float)S.reversed, *o++ = (float)S.culled, *o++ = S.winding;
    return ok;
}

float orc_walk(float a, float d, int64_t n) {
    for (int64_t i = 0; i < n; i++) a = a + d;
    return a;
}

/* The fixed vertex and fragment shaders of one pipeline, exposed for the
 * closed-form tests of the shader models (tests/test_oracle.py). */
static int pipe_init(Pipe *p, const orc_desc *d, const float *u, const orc_stream *s, const orc_texture *tex,
                     const orc_texture *shadow) {
    p->d = d, p->u = u, p->s = s, p->tex = tex, p->shadow = shadow;
    p->frag_kind = d->shader == ORC_SHADER_FLAT ? 1 : (d->shader == ORC_SHADER_DEPTH_ONLY ? 2 : 0);
    switch (d->shader) {
    case ORC_SHADER_FLAT: p->nvd = 1; break;
    case ORC_SHADER_COLOR:
    case ORC_SHADER_GOURAUD: p->nvd = 4; break;
    case ORC_SHADER_PHONG: p->nvd = 6; break;
    case ORC_SHADER_DEPTH_ONLY: p->nvd = 0; break;
    case ORC_SHADER_TEXTURED: p->nvd = 2; break;
    case ORC_SHADER_PHONG_SHADOW: p->nvd = 9; break;
    case ORC_SHADER_TEX_SHADOW: p->nvd = 11; break;
    default: return ORC_INVALID;
    }
    return ORC_OK;
}

int orc_vertex(const orc_desc *d, const float *u, const orc_stream *s, int32_t inst, int64_t vi, float pos[4],
               float vd[12]) {
    Pipe p;
    if (pipe_init(&p, d, u, s, NULL, NULL)) return ORC_INVALID;
    SVert v;
    vertex_shader(&p, inst, vi, &v);
    memcpy(pos, v.pos, sizeof(v.pos));
    memcpy(vd, v.vd, sizeof(v.vd));
    return p.nvd;
}

int orc_fragment(const orc_desc *d, const float *u, const orc_texture *tex, const orc_texture *shadow,
                 const float *vd, float out[4], uint32_t old_pixel, uint32_t *blended) {
    Pipe p;
    if (pipe_init(&p, d, u, NULL, tex, shadow)) return ORC_INVALID;
    Frag f;
    fragment_shader(&p, vd, &f);
    memcpy(out, f.c, sizeof(f.c));
    if (blended) *blended = blend_shader(&p, old_pixel, &f);
    return ORC_OK;
}

typedef struct {
    int64_t *out;
    int64_t n, max;
} PlotCollect;

static void collect_pixel(void *ctx, int64_t x, int64_t y) {
    PlotCollect *c = (PlotCollect *)ctx;
    if (c->n < c->max) c->out[2 * c->n] = x, c->out[2 * c->n + 1] = y;
    c->n++;
}

int64_t orc_clip_line(int64_t x1, int64_t y1, int64_t x2, int64_t y2, int64_t cx0, int64_t cy0, int64_t cx1,
                      int64_t cy1, int64_t *out_xy, int64_t max) {
    PlotCollect c = {out_xy, 0, max};
    clip_line(x1, y1, x2, y2, cx0, cy0, cx1, cy1, collect_pixel, &c);
    return c.n;
}
