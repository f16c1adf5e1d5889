// shade.hpp -- device-side scalar semantics and the fixed shader set.
//
// Every function here is evaluated with -ffp-contract=off and IEEE f32
// division/sqrt, so each operation rounds once in the order written.  The
// order follows the reference's expressions (triangles.rs, math.rs,
// sampler/linear.rs) and this project's fixed-shader definitions.  Shaders are
// template parameters so that VertexData arrays have compile-time sizes and
// live in registers.
#pragma once
#include "internal.hpp"

namespace euchip {

#define EUCHIP_DEV __device__ __forceinline__

// VertexData array length of a shader (at least 1 so arrays stay legal).
template <uint32_t SH>
struct VDN {
    static constexpr uint32_t N = vd_count(SH) ? vd_count(SH) : 1;
};

// f32::min / f32::max (the non-NaN operand wins).
EUCHIP_DEV float rmin(float a, float b) { return fminf(a, b); }
EUCHIP_DEV float rmax(float a, float b) { return fmaxf(a, b); }
// f32::fract (self - self.trunc()).
EUCHIP_DEV float rfract(float x) { return x - truncf(x); }

// `(f as usize).clamp(lo, hi)` for hi < 2^24: saturating, NaN -> 0.
EUCHIP_DEV uint32_t sat_clamp(float f, uint32_t lo, uint32_t hi) {
    uint32_t v;
    if (!(f > 0.0f)) v = 0;
    else if (f >= (float)hi) v = hi;
    else v = (uint32_t)f;
    return v < lo ? lo : v;
}

// `f as u8`: saturating truncation, NaN -> 0.
EUCHIP_DEV uint32_t sat_u8(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 255.0f) return 255;
    return (uint32_t)f;
}

EUCHIP_DEV void mat4_vec4(const float *m, const float v[4], float r[4]) {
#pragma unroll
    for (int i = 0; i < 4; i++)
        r[i] = ((m[4 * i + 0] * v[0] + m[4 * i + 1] * v[1]) + m[4 * i + 2] * v[2]) + m[4 * i + 3] * v[3];
}

EUCHIP_DEV void normalize3(const float v[3], float r[3]) {
    float l2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    float s = 1.0f / sqrtf(l2);
    r[0] = v[0] * s;
    r[1] = v[1] * s;
    r[2] = v[2] * s;
}

EUCHIP_DEV float dot3s(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

EUCHIP_DEV float pow30(float x) {
    float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
    float x24 = x16 * x8, x28 = x24 * x4;
    return x28 * x2;
}

// Two directional lights: ambient + sum_i I_i * (kd * max(n.l, 0) + ks * max(r.v, 0)^30).
EUCHIP_DEV void lighting2(const float *u, const float wpos[3], const float nrm[3], float out[4]) {
    float n[3], vv[3], v[3];
    normalize3(nrm, n);
    vv[0] = u[U_CAM + 0] - wpos[0];
    vv[1] = u[U_CAM + 1] - wpos[1];
    vv[2] = u[U_CAM + 2] - wpos[2];
    normalize3(vv, v);
    float light = u[U_K + 0];
#pragma unroll
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
#pragma unroll
    for (int c = 0; c < 4; c++) out[c] = u[U_SURF + c] * light;
}

// Bilinear sample, Clamped<Linear> (sampler/mod.rs:183-186, sampler/linear.rs:30-66)
// of an RGBA8 texture mapped to f32 per channel (NC = 4), or of an f32 map (NC = 1).
template <int NC>
EUCHIP_DEV void sample_bilinear(const void *data, uint32_t w, uint32_t h, float x, float y, bool clamped,
                                float *out) {
    if (clamped) {
        x = rmin(rmax(x, 0.0f), 1.0f);
        y = rmin(rmax(y, 0.0f), 1.0f);
    }
    float itx = rfract(x) * (float)w;
    float ity = rfract(y) * (float)h;
    float tx = truncf(itx), ty = truncf(ity);
    uint32_t px = !(tx > 0.0f) ? 0u : (tx >= (float)w ? w : (uint32_t)tx);
    uint32_t py = !(ty > 0.0f) ? 0u : (ty >= (float)h ? h : (uint32_t)ty);
    float fx = rfract(itx);
    float fy = rfract(ity);
    uint32_t p0x = px < w - 1 ? px : w - 1;
    uint32_t p0y = py < h - 1 ? py : h - 1;
    uint32_t p1x = px + 1 < w - 1 ? px + 1 : w - 1;
    uint32_t p1y = py + 1 < h - 1 ? py + 1 : h - 1;
    if (NC == 1) {
        const float *t = (const float *)data;
        float t00 = t[p0x + w * p0y], t10 = t[p1x + w * p0y], t01 = t[p0x + w * p1y], t11 = t[p1x + w * p1y];
        float t0 = t00 * (1.0f - fy) + t01 * fy;
        float t1 = t10 * (1.0f - fy) + t11 * fy;
        out[0] = t0 * (1.0f - fx) + t1 * fx;
    } else {
        const uint32_t *t = (const uint32_t *)data;
        uint32_t q00 = t[p0x + w * p0y], q10 = t[p1x + w * p0y], q01 = t[p0x + w * p1y], q11 = t[p1x + w * p1y];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            float t00 = (float)((q00 >> (8 * c)) & 0xffu), t10 = (float)((q10 >> (8 * c)) & 0xffu);
            float t01 = (float)((q01 >> (8 * c)) & 0xffu), t11 = (float)((q11 >> (8 * c)) & 0xffu);
            float t0 = t00 * (1.0f - fy) + t01 * fy;
            float t1 = t10 * (1.0f - fy) + t11 * fy;
            out[c] = t0 * (1.0f - fx) + t1 * fx;
        }
    }
}

// f32::rem_euclid for m = 1 or 2: x % m (exact), plus m when negative.  x % m =
// x - m * trunc(x / m) with the sign of x: x / m and m * trunc(...) are exact
// for these m, and the difference is the low bits of x, so exact too (NaN and
// +-Inf give NaN as fmod does).
EUCHIP_DEV float rem_euclid(float x, float m) {
    const float r = copysignf(x - m * truncf(x / m), x);
    return r < 0.0f ? r + m : r;
}

// Texture sample through sampler `kind` (EUCHIP_SAMPLER_LINEAR ..
// NEAREST_MIRRORED): the wrapper's index map (sampler/mod.rs:98-179: none,
// Clamped, Tiled, Mirrored), then Linear or Nearest (sampler/nearest.rs:20-38
// with Denormalize f32 -> usize, math.rs:52-55) of an RGBA8 texture.
EUCHIP_DEV void sample_rgba8(const uint32_t *tex, uint32_t w, uint32_t h, uint32_t kind, float x, float y, float out[4]) {
    const bool nearest = kind >= SAMPLER_NEAREST;
    const uint32_t wrap = nearest ? kind - SAMPLER_NEAREST : kind - SAMPLER_LINEAR;
    if (wrap == 1u) {
        x = rmin(rmax(x, 0.0f), 1.0f);
        y = rmin(rmax(y, 0.0f), 1.0f);
    } else if (wrap == 2u) {
        x = rem_euclid(x, 1.0f);
        y = rem_euclid(y, 1.0f);
    } else if (wrap == 3u) {
        x = rem_euclid(x, 2.0f) >= 1.0f ? 1.0f - rem_euclid(x, 1.0f) : rem_euclid(x, 1.0f);
        y = rem_euclid(y, 2.0f) >= 1.0f ? 1.0f - rem_euclid(y, 1.0f) : rem_euclid(y, 1.0f);
    }
    if (!nearest) {
        sample_bilinear<4>(tex, w, h, x, y, false, out);
        return;
    }
    // ((u * size as f32).max(0.0) as usize).min(size - 1)
    const float fx = rmax(x * (float)w, 0.0f), fy = rmax(y * (float)h, 0.0f);
    const uint32_t px = fx >= (float)w ? w - 1u : ((uint32_t)fx < w - 1u ? (uint32_t)fx : w - 1u);
    const uint32_t py = fy >= (float)h ? h - 1u : ((uint32_t)fy < h - 1u ? (uint32_t)fy : h - 1u);
    const uint32_t q = tex[px + w * py];
#pragma unroll
    for (int c = 0; c < 4; c++) out[c] = (float)((q >> (8 * c)) & 0xffu);
}

// ---------------------------------------------------------------------------
// Pipeline::vertex of the fixed shader set
// ---------------------------------------------------------------------------

// World position (model * (p, 1)) or, for CLIP input, the position itself.
EUCHIP_DEV void vertex_world(const FrameParams &p, const View &v, uint32_t inst, uint32_t vi, float wpos[4]) {
    const float *P = v.pos + (size_t)v.pos_comps * vi;
    if (p.xform == 0) {
        wpos[0] = P[0], wpos[1] = P[1], wpos[2] = P[2];
        wpos[3] = v.pos_comps == 4 ? P[3] : 1.0f;
    } else {
        const float p4[4] = {P[0], P[1], P[2], 1.0f};
        mat4_vec4(v.models + 16u * inst, p4, wpos);
    }
}

// VertexData returned by vertex().
template <uint32_t SH>
EUCHIP_DEV void vertex_data(const FrameParams &p, const View &v, uint32_t inst, uint32_t vi, float *vd) {
    if constexpr (SH == FLAT) {
        vd[0] = __uint_as_float(v.rgba8[vi]);
    } else if constexpr (SH == COLOR) {
        const float *c = v.col + 4u * vi;
        vd[0] = c[0], vd[1] = c[1], vd[2] = c[2], vd[3] = c[3];
    } else if constexpr (SH == TEXTURED) {
        vd[0] = v.uv[2u * vi], vd[1] = v.uv[2u * vi + 1];
    } else if constexpr (SH == DEPTH_ONLY) {
        (void)vd;
    } else {
        float wpos[4], wn[4];
        vertex_world(p, v, inst, vi, wpos);
        const float *N = v.nrm + 3u * vi;
        float n4[4] = {N[0], N[1], N[2], 0.0f};
        if constexpr (SH == PHONG_SHADOW) n4[0] = -N[0], n4[1] = -N[1], n4[2] = -N[2];
        if (p.xform == 0) wn[0] = n4[0], wn[1] = n4[1], wn[2] = n4[2], wn[3] = n4[3];
        else mat4_vec4(v.models + 16u * inst, n4, wn);
        if constexpr (SH == GOURAUD) {
            lighting2(v.u, wpos, wn, vd);
        } else {
            vd[0] = wpos[0], vd[1] = wpos[1], vd[2] = wpos[2];
            vd[3] = wn[0], vd[4] = wn[1], vd[5] = wn[2];
            if constexpr (SH == PHONG_SHADOW || SH == TEX_SHADOW) {
                const float w1[4] = {wpos[0], wpos[1], wpos[2], 1.0f};
                float l4[4];
                mat4_vec4(v.u + U_LVP, w1, l4);
                int o = 6;
                if constexpr (SH == TEX_SHADOW) {
                    vd[6] = v.uv[2u * vi], vd[7] = v.uv[2u * vi + 1];
                    o = 8;
                }
                vd[o + 0] = l4[0] / l4[3];
                vd[o + 1] = l4[1] / l4[3];
                vd[o + 2] = l4[2] / l4[3];
            }
        }
    }
}

// Shadow-map lookup of examples/teapot.rs:123-128.
EUCHIP_DEV bool in_light(const View &v, const float lvp[3]) {
    float sx = (lvp[0] * 1.0f) * 0.5f + 0.5f;
    float sy = (lvp[1] * -1.0f) * 0.5f + 0.5f;
    float ld;
    sample_bilinear<1>(v.shadow, v.sh_w, v.sh_h, sx, sy, true, &ld);
    ld = ld + v.u[U_K + 3];
    return lvp[2] < ld;
}

// Pipeline::fragment of the fixed shader set.  out: Rgba<f32> (FLAT: colour bits in out[0]).
// GS: the texture is read through `sampler` (any EUCHIP_SAMPLER_* kind);
// otherwise through the shader's own one (TEXTURED: Linear, TEX_SHADOW:
// Clamped<Linear>), whose kernels then carry no code of the others.
template <uint32_t SH, bool GS = false>
EUCHIP_DEV void fragment_shader(const View &v, uint32_t sampler, const float *vd, float out[4]) {
    const float *u = v.u;
    out[0] = out[1] = out[2] = out[3] = 0.0f;
    if constexpr (SH == FLAT) {
        out[0] = vd[0];
    } else if constexpr (SH == COLOR || SH == GOURAUD) {
        out[0] = vd[0], out[1] = vd[1], out[2] = vd[2], out[3] = vd[3];
    } else if constexpr (SH == PHONG) {
        lighting2(u, vd, vd + 3, out);
    } else if constexpr (SH == TEXTURED) {
        if constexpr (GS) sample_rgba8(v.tex, v.tex_w, v.tex_h, sampler, vd[0], vd[1], out);
        else sample_bilinear<4>(v.tex, v.tex_w, v.tex_h, vd[0], vd[1], false, out);
    } else if constexpr (SH == PHONG_SHADOW) {
        float wn[3], cd[3], ld[3], t3[3];
        normalize3(vd + 3, wn);
        for (int c = 0; c < 3; c++) t3[c] = u[U_CAM + c] - vd[c];
        normalize3(t3, cd);
        for (int c = 0; c < 3; c++) t3[c] = vd[c] - u[U_LPOS + c];
        normalize3(t3, ld);
        float nl[3] = {-ld[0], -ld[1], -ld[2]};
        float diffuse = rmax(dot3s(wn, nl), 0.0f) * u[U_K + 1];
        float dn = dot3s(nl, wn) * 2.0f;
        float r[3] = {nl[0] - wn[0] * dn, nl[1] - wn[1] * dn, nl[2] - wn[2] * dn};
        float ncd[3] = {-cd[0], -cd[1], -cd[2]};
        float specular = pow30(rmax(dot3s(r, ncd), 0.0f)) * u[U_K + 2];
        float light = u[U_K + 0] + (in_light(v, vd + 6) ? diffuse + specular : 0.0f);
        for (int c = 0; c < 4; c++) out[c] = u[U_SURF + c] * light;
    } else if constexpr (SH == TEX_SHADOW) {
        float n[3], texc[4];
        normalize3(vd + 3, n);
        float ndl = rmax(dot3s(n, u + U_L0), 0.0f);
        float light = u[U_K + 0] + (in_light(v, vd + 8) ? ndl * u[U_K + 1] : 0.0f);
        if constexpr (GS) sample_rgba8(v.tex, v.tex_w, v.tex_h, sampler, vd[6], vd[7], texc);
        else sample_bilinear<4>(v.tex, v.tex_w, v.tex_h, vd[6], vd[7], true, texc);
        float s = light * (1.0f / 255.0f);
        out[0] = texc[0] * s;
        out[1] = texc[1] * s;
        out[2] = texc[2] * s;
        out[3] = texc[3] * (1.0f / 255.0f);
    }
}

// Pipeline::blend: every fixed shader replaces the old pixel (DEPTH_ONLY's Unit keeps it).
template <uint32_t SH>
EUCHIP_DEV uint32_t blend_shader(uint32_t old, const float f[4], uint32_t bgra) {
    if constexpr (SH == FLAT) return __float_as_uint(f[0]);
    if constexpr (SH == DEPTH_ONLY) return old;
    uint32_t b[4];
    if constexpr (SH == TEXTURED) {
#pragma unroll
        for (int c = 0; c < 4; c++) b[c] = sat_u8(f[c]);
    } else {
#pragma unroll
        for (int c = 0; c < 4; c++) b[c] = sat_u8(rmin(rmax(f[c], 0.0f), 1.0f) * 255.0f);
    }
    if (bgra) return b[2] | (b[1] << 8) | (b[0] << 16) | (b[3] << 24);
    return b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24);
}

}  // namespace euchip
