// internal.hpp -- structures shared by the host runtime and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace euchip {

constexpr int kUniforms = 64;
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

// One triangle after setup (triangles.rs:55-171).
struct TriRec {
    float o[3], dx[3], dy[3];   // w_hom_origin, w_hom_dx, w_hom_dy
    float hz[3];                // verts_hom[i][2] after flip/reversal
    float vby[6];               // verts_by_y (screen x, y) x 3
    uint32_t xmin, xmax;        // bounds_clamped x, clamped to [0, W]
    uint32_t ymin, ymax;        // bounds_clamped y, clamped to [0, H]
    uint32_t flags;             // TRI_ALL_IN | TRI_SWAPPED
    uint32_t sid[3];            // shaded-vertex id in vary (request base + instance * n_vertices + vertex), post-reversal
};

enum TriFlags : uint32_t {
    TRI_ALL_IN = 1u,   // all three vertices pass the z clip
    TRI_SWAPPED = 2u,  // vertices 0 and 2 were exchanged (winding >= 0)
};

// Setup records written by k_setup for every kept triangle, read by k_raster.
// rec0 (one 64-byte line): o[3], dx[3], dy[3], hz[3], xmin, xmax, ymin,
//                          ymax | flags << 30
// rec1 (32 bytes):         verts_by_y x, y x 3 (read only for the spans of
//                          large triangles), 2 words unused
constexpr uint32_t kRec0Quads = 4, kRec1Quads = 2;

enum FrameFlags : uint32_t { FLAG_BIN_OVERFLOW = 1u };

// Device counters of one launch stream, per frame parity f & 1: setup of frame f
// counts into the [f & 1] entries, its raster pass reads them and zeroes the
// [(f + 1) & 1] entries for the next frame.  Each counter sits on its own
// 128-byte line.
struct Counters {
    uint32_t pool[2][32];     // words of the overflow pool taken
    uint32_t next[2][8][32];  // per XCD: next tile of the dynamic tile loop
    uint32_t empty[2][8][32]; // per XCD: next group of tiles the empty-tile clears look at
    uint32_t heavy[2][32];    // entries of the heavy-tile list
    uint32_t big[2][32];      // entries of the big-triangle list
    uint32_t busy[2][8][32];  // per XCD: entries of the busy-tile list
    uint32_t flags[2][32];    // FrameFlags of the frame (the pool ran out: it was dropped)
    uint32_t max_count[2][32];// largest bin count of a tile past its capacity (one-request frames)
};

// What a one-request frame's raster pass reports to host-visible memory when
// it starts (the counts of its own setup): the host grows bins and the pool
// from it without waiting for the stream.  `frame` is written last.
struct Report {
    uint32_t max_count;  // largest tile count past its bin capacity
    uint32_t flags;      // FrameFlags
    uint32_t pool_used;  // pool words the frame asked for (more than pool_cap when it ran out)
    uint32_t dropped;    // frames dropped so far on this context (running count)
    uint32_t frame;      // the frame these counts belong to
};

// Per-tile bin state, 8 bytes per tile so the raster pass loads both at once:
// `count` bin entries (ranks below the tile's capacity sit in its bin), and the
// head of the tile's overflow segments (pool word + 1, 0 = none).  Every setup
// workgroup whose ranks for the tile pass the capacity takes one segment of
// the pool, [next head, n, n primitives], and pushes it on the tile's list.
struct TileBin {
    uint32_t count, head;
};

// Bin entries the raster pass stages per step.
constexpr uint32_t kBatch = 128;
// Tiles whose bin reaches this many entries are listed by setup and rasterised
// first (the longest tiles start early instead of ending the launch late).
constexpr uint32_t kHeavyTile = 128u;
// Which of the 8 busy-tile lists (one per XCD) a tile goes to: its XCD's
// contiguous tile range (batches: each XCD gets whole requests), or its block
// of 4 x 4 tiles dealt over the XCDs (one-request frames: each XCD gets a share
// of every screen region, so dense regions do not load one XCD).
enum BusyMap : uint32_t { BUSY_RANGES = 0, BUSY_BLOCKS = 1 };

// One-request frames of meshes with no more triangles than the target has tiles
// ("big-triangle mode").  With row checkpoints (below) no triangle is binned:
// the span blocks walk them all and every tile reads its checkpoint list.
// Without them a triangle whose box covers more tiles than kBigTiles is not
// binned (binning it costs one global atomic per tile when the workgroup's LDS
// grouping overflows) but goes to the big-triangle list (while it has room),
// and every tile walks those from the row starts.
constexpr uint32_t kBigTiles = 64;
constexpr uint32_t kBigCap = 4096;

// Row checkpoints of big triangles.  The exact walk of a row (w += dx from the
// row start x0, triangles.rs:257-301) reaches a tile only after every pixel
// left of it; the span blocks of the setup launch walk each row of each big
// triangle once and keep, per tile column the row's span touches, the weights
// at the first pixel of the span in that column.  They set up their triangle
// themselves (the setup blocks leave big triangles out of the bins), so the
// two kinds of block run side by side, and every slot has a fixed place: the
// checkpoint of triangle t, tile column c and row y is slot (t * ntx + c) * H
// + y of the pool (one 16-byte slot per triangle, column and row of the
// target, so the 8 rows of one tile are consecutive): (w0, w1, w2) at xs, then
// (xs - X0) | (xe - X0) << 16, the row's span inside the column that starts at
// X0.  Each tile some row of which has such a span gets the 8-byte entry
// (t << 8 | mask of those rows, klo) in its list (cktab, tile * mstride + k, k
// from the tile's count, ck_count[tile].x).  The raster pass reads its tile's
// list and walks those rows from their checkpoints.  Tiles are then at most 8
// rows high.
// Occlusion (write modes, Less or Greater with depth write, where the winner is
// the fragment of least key): klo is a lower bound of the key order (zord of z,
// or ~zord) of every fragment the triangle has in the tile.  A triangle whose
// spans cover every pixel of the tile, with margins that hold under the walk's
// roundings, raises ck_count[tile].y to ~khi, khi an upper bound of its keys'
// order there.  An entry with klo above the tile's least khi can win no pixel:
// the raster skips it.  (Bounds: the triangle's weight and depth planes at the
// tile's corners, widened by what the row walk's roundings can add; the
// derivation is at tile_bounds in kernels.hip.)
constexpr uint32_t kSpanRows = 64;             // rows per item of the checkpoint walk (one wave)
// Bytes of one span wave's LDS tables: two row-mask bytes per (tile row of
// the item, tile column).
constexpr uint32_t span_table_bytes(uint32_t sy, uint32_t ntx) {
    return 2u * ((((kSpanRows >> sy) * ntx) + 3u) & ~3u);
}
// Per triangle of a checkpoint frame (big_rec): (primitive, flags << 30, dx0,
// dx1), (dx2, hz0, hz1, hz2), what the raster's row tasks read besides the slot.
constexpr uint32_t kBigRecQuads = 2;
constexpr size_t kCkMaxSlots = (size_t)1 << 26;  // 1 GiB of checkpoints at most

// One render request: a target pair, a mesh, a uniform block, and where the
// request's vertices, tiles and bins sit in the launch's scratch memory.
struct View {
    // target (tgt_size) and the two targets; null = Empty
    uint32_t W, H;
    uint32_t G;            // render_par band height (rows), H for render_seq
    uint32_t ntx, nty;     // tiles of the target
    uint32_t tile_base;    // first tile of this request in the launch's tile space
    uint32_t sv_base;      // first shaded vertex of this request in vary
    uint32_t tri_base;     // first triangle of this request in rec0 / rec1
    uint32_t tile_cap;     // bin slots per tile of this request
    uint64_t bin_base;     // first bin slot of this request
    uint32_t *pixel;
    float *depth;
    uint32_t clear_color_en, clear_color, clear_depth_en;
    float clear_depth;
    float ndc_x, ndc_y;    // 2/size_x and -2/size_y of to_ndc (triangles.rs:44-48)
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
    float u[kUniforms];
};

// Sampler kinds (EUCHIP_SAMPLER_*): LINEAR + wrap, NEAREST + wrap; wrap 0 none,
// 1 Clamped, 2 Tiled, 3 Mirrored.
enum SamplerKind : uint32_t { SAMPLER_LINEAR = 1, SAMPLER_LINEAR_CLAMPED = 2, SAMPLER_NEAREST = 5 };
enum Geometry : uint32_t { GEOM_IDENTITY = 0, GEOM_REVERSED = 1, GEOM_BOTH_SIDES = 2 };

struct FrameParams {
    // pipeline state shared by every request of the launch
    uint32_t shader, xform, pixel_write, depth_test, depth_write, y_up, z_clip;
    float z_min, z_max;
    uint32_t cull, msaa_level, band_par;
    uint32_t bgra;         // blend packs [b, g, r, a]
    uint32_t sampler;      // texture sampler kind (resolved: never EUCHIP_SAMPLER_DEFAULT)
    uint32_t geometry;     // Geometry: primitive t of the stream is stage output t % k of input triangle t / k
    // lines (one-request frames): LINE_LIST or LINE_TRIANGLE_LIST, n_lines lines;
    // a 64-byte record per line and a 64-bit winner key per target pixel
    uint32_t primitives, n_lines;
    uint32_t *lrec;
    uint64_t *lkey;
    // per-launch scratch (device)
    uint4 *rec0;           // kRec0Quads per triangle of the launch
    uint4 *rec1;           // kRec1Quads per triangle of the launch
    float *vary;           // VertexData per shaded vertex
    Counters *ctr;
    TileBin *tile_count;   // per tile of the launch, zero before setup, read-only in raster
    uint32_t *bins;        // bin_base + local tile * tile_cap + rank -> primitive
    uint32_t *req_max;     // batches, per request: largest tile count above its tile_cap (host grows the bins)
    uint32_t *pool;        // overflow segments (ranks >= tile_cap), pool_cap words
    uint32_t pool_cap;
    Report *report;        // host-visible report of one-request frames, or null
    uint32_t *heavy;       // tiles with at least kHeavyTile entries (heavy_cap slots)
    uint32_t heavy_cap;    // 0: no heavy-first pass
    // Setup lists the tiles it bins into (busy) in 8 lists of up to ntiles
    // entries (list x at busy + x * ntiles, chosen by busy_map); raster clears
    // the empty tiles itself, takes only busy tiles from the lists, leaves
    // `tile_count` as it is and zeroes the first `zero_n` counts of
    // `count_next` (one-request frames: the other half of a two-frame count
    // buffer; batches zero their counts with a fill before setup).
    uint32_t *busy;
    uint32_t busy_map;
    uint32_t *big;         // big triangles (primitive ids), big_cap slots; big_cap 0: off (batches)
    uint32_t big_cap;
    // Row checkpoints (ck_on): per triangle its big record (kBigRecQuads
    // quads), the pool of n_tris x ntx x H slots and the tiles' lists
    // (ntiles x mstride entries, mstride >= n_tris) with their counts (two
    // halves by frame parity like tile_count: ck_count this frame's,
    // ck_count_next the other, whose first ck_zero_n counts a big-triangle
    // frame's raster zeroes).
    // A one-request setup launch has tri_blocks triangle blocks, then vertex
    // blocks up to setup_blocks; a checkpoint frame's has setup_blocks blocks
    // of triangles and vertices, then the span blocks (span_parts items of
    // kSpanRows rows per triangle, 4 per block).
    uint4 *big_rec;
    uint4 *ck;
    uint32_t ck_on;
    uint2 *cktab;
    uint32_t mstride;
    uint2 *ck_count, *ck_count_next;  // per tile (list length, ~least occluder bound)
    uint32_t ck_zero_n;
    uint32_t tri_blocks, setup_blocks, span_parts;
    TileBin *count_next;
    uint32_t zero_n;
    uint32_t frame;        // frame parity selects the overflow counter and tile queue
    uint32_t ntiles;       // tiles of all requests
    uint32_t tile_sx;      // tiles are (1 << tile_sx) x (256 >> tile_sx) pixels, tile_sx in 3..5
    // requests: one in `v0` (kernel argument), or `n_views` in device memory
    uint32_t n_views;
    const View *views;
    const uint2 *blk;          // batched setup: workgroup -> (request, workgroup within the request)
    const uint32_t *tile_req;  // batched raster: tile -> request
    View v0;
};

// euchip_render_noalloc: every per-frame buffer is carved from one static device
// array (noalloc.hip) of kNoAllocBytes, laid out below; requests beyond the
// limits are rejected.
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
constexpr uint32_t kNaMaxTris = 1u << 18;         // triangles (instances x count / 3)
constexpr uint32_t kNaMaxVaryFloats = 3u << 20;   // shaded vertices x VertexData floats
constexpr uint32_t kNaMaxTiles = 1u << 14;        // 32 x 8 tiles, e.g. 2048 x 2048 pixels
constexpr uint32_t kNaBinSlots = 3u << 20;        // bin slots over all tiles
constexpr size_t kNaRec0 = 0;
constexpr size_t kNaRec1 = kNaRec0 + (size_t)kNaMaxTris * kRec0Quads * 16;
constexpr size_t kNaVary = kNaRec1 + (size_t)kNaMaxTris * kRec1Quads * 16;
constexpr size_t kNaBins = kNaVary + (size_t)kNaMaxVaryFloats * 4;
constexpr size_t kNaTiles = kNaBins + (size_t)kNaBinSlots * 4;
constexpr size_t kNaBusy = kNaTiles + (size_t)kNaMaxTiles * sizeof(TileBin);
constexpr size_t kNaHeavy = kNaBusy + (size_t)8 * kNaMaxTiles * 4;
constexpr size_t kNaBig = kNaHeavy + (size_t)kNaMaxTiles * 4;
constexpr size_t kNaCtr = kNaBig + (size_t)kBigCap * 4;
constexpr size_t kNaPool = kNaCtr + ((sizeof(Counters) + 255) / 256) * 256;
constexpr size_t kNaPoolWords = (kNoAllocBytes - kNaPool) / 4;
static_assert(kNaPool < kNoAllocBytes && kNaPoolWords >= (1u << 20), "no-allocation scratch layout");

// The device address of the no-allocation scratch array on `device` (the
// current device), looked up once per device (noalloc.hip).
hipError_t noalloc_scratch(int device, void **out);

// Kernel launchers (kernels.hip).  All enqueue on `s`.  `batched` selects the
// kernels that read their requests from p.views (setup then runs `blocks`
// workgroups as p.blk maps them); one-request launches read p.v0.
hipError_t launch_setup(const FrameParams &p, hipStream_t s, bool batched, uint32_t blocks);
hipError_t launch_raster(const FrameParams &p, hipStream_t s, bool batched, uint32_t max_groups);
// Row checkpoints of the big triangles (one-request big-triangle frames with a
// checkpoint pool), between setup and raster.
hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t n, hipStream_t s);
// The lines path of one request (p.v0): vertex stage + line setup, winner keys
// initialised, lines walked, then every pixel resolved (colour, depth, clears).
hipError_t launch_lines(const FrameParams &p, hipStream_t s);

// PrimitiveKind (EUCHIP_PRIMS_*).
enum Prims : uint32_t { PRIMS_TRIANGLES = 0, PRIMS_LINE_LIST = 1, PRIMS_LINE_TRIANGLES = 2 };
// Line endpoints further than this from the origin (pixels) are not drawn.
constexpr int32_t kLineCoordLimit = 1 << 29;
constexpr uint32_t kLineRecWords = 16;

}  // namespace euchip
