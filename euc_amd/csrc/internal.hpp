// internal.hpp -- structures shared by the host runtime and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace euchip {

constexpr int kUniforms = 64;
constexpr int kTileW = 32;  // raster tile: 32 x 8 pixels, one pixel per thread of a 256-thread workgroup
constexpr int kTileH = 8;
constexpr int kTileShiftX = 5;
constexpr int kTileShiftY = 3;
constexpr int kMaxVD = 12;  // VertexData floats of the widest fixed shader

// Uniform block offsets (same layout as include/euchip.h documents).
constexpr int U_VP = 0, U_L0 = 16, U_L1 = 20, U_CAM = 24, U_SURF = 28, U_K = 32, U_LPOS = 36, U_LVP = 40;

enum Shader : uint32_t { FLAT = 0, COLOR = 1, GOURAUD = 2, PHONG = 3, DEPTH_ONLY = 4, TEXTURED = 5, PHONG_SHADOW = 6, TEX_SHADOW = 7 };
enum DepthTest : uint32_t { DT_NONE = 0, DT_LESS = 1, DT_EQUAL = 2, DT_GREATER = 3 };

constexpr __host__ __device__ uint32_t vd_count(uint32_t shader) {
    return shader == FLAT ? 1
         : (shader == COLOR || shader == GOURAUD) ? 4
         : shader == PHONG ? 6
         : shader == DEPTH_ONLY ? 0
         : shader == TEXTURED ? 2
         : shader == PHONG_SHADOW ? 9
         : shader == TEX_SHADOW ? 11 : 0;
}

// One triangle after setup (triangles.rs:55-171).  Lives in registers: the
// raster pass recomputes it from the vertices' clip positions instead of
// storing it.
struct TriRec {
    float o[3], dx[3], dy[3];   // w_hom_origin, w_hom_dx, w_hom_dy
    float hz[3];                // verts_hom[i][2] after flip/reversal
    float vby[6];               // verts_by_y (screen x, y) x 3
    uint32_t xmin, xmax;        // bounds_clamped x, clamped to [0, W]
    uint32_t ymin, ymax;        // bounds_clamped y, clamped to [0, H]
    uint32_t flags;             // bit 0: all three vertices pass the z clip
    uint32_t sid[3];            // shaded-vertex id (instance * n_vertices + vertex), post-reversal
};

enum FrameFlags : uint32_t { FLAG_BIN_OVERFLOW = 1u };

// Device counters of one context.  ovf[f & 1] counts the overflow entries of
// frame f; the raster pass of frame f zeroes ovf[(f + 1) & 1] for the next one.
// Each counter sits on its own 128-byte line.
struct Counters {
    uint32_t ovf[2][32];
    uint32_t next[2][8][32];  // per frame parity and XCD: next tile of the dynamic tile loop
    uint32_t flags;      // FrameFlags, sticky until the host clears them
    uint32_t max_count;  // largest per-tile triangle count seen (host resizes bins)
    uint32_t pad[30];
};

struct FrameParams {
    // target (tgt_size) and the two targets; null = Empty
    uint32_t W, H;
    uint32_t *pixel;
    float *depth;
    // pipeline
    uint32_t shader, xform, pixel_write, depth_test, depth_write, y_up, z_clip;
    float z_min, z_max;
    uint32_t cull, msaa_level, band_par, G;
    uint32_t clear_color_en, clear_color, clear_depth_en;
    float clear_depth;
    float ndc_x, ndc_y;   // 2/size_x and -2/size_y of to_ndc (triangles.rs:44-48)
    // mesh (device)
    const float *pos;
    const float *nrm;
    const float *col;
    const float *uv;
    const uint32_t *rgba8;
    const uint32_t *idx;
    const float *models;
    uint32_t pos_comps, count, n_vertices, n_instances, n_tris;
    // textures (device)
    const uint32_t *tex;
    uint32_t tex_w, tex_h;
    const float *shadow;
    uint32_t sh_w, sh_h;
    // per-frame scratch (device)
    float4 *clip;          // clip position per shaded vertex (instance * n_vertices + vertex)
    float *vary;           // VertexData per shaded vertex (instance * n_vertices + vertex)
    Counters *ctr;
    uint32_t *tile_count;  // per tile, zero between frames
    uint32_t *bins;        // tile * tile_cap + rank -> primitive
    uint32_t tile_cap;     // bin slots per tile
    uint2 *ovf;            // (tile, primitive) of ranks >= tile_cap
    uint32_t ovf_cap;
    uint32_t frame;        // frame parity selects the overflow counter
    uint32_t ntx, nty;
    float u[kUniforms];
};

// Kernel launchers (kernels.hip).  All enqueue on `s`.
hipError_t launch_setup(const FrameParams &p, hipStream_t s);
hipError_t launch_raster(const FrameParams &p, hipStream_t s, uint32_t max_groups);
hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t n, hipStream_t s);

}  // namespace euchip
