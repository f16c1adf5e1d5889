// kernels.hip -- gfx950 kernels of the triangle path.
//
// Frame = setup -> raster, for one request (the view in the kernel
// argument) or for a batch of requests (views in device memory, each with its
// own target, mesh, uniforms, tiles and bins):
//   k_setup   one lane per shaded vertex (clip position and VertexData of the
//             fixed vertex shader) and one lane per triangle: euc's triangle
//             setup (triangles.rs:55-171) in IEEE f32, culling, and binning of
//             the triangle into every tile its bounding box covers.  Pairs are
//             grouped by tile in an LDS hash so each distinct tile costs one
//             global atomic per workgroup; ranks past a tile's bin capacity go
//             to overflow segments of that tile taken from a pool.  With few
//             triangles per tile (big-triangle mode) and row checkpoints no
//             triangle is binned: span blocks of the same launch set each one
//             up again, walk its rows once and store the row's weights where
//             it enters each tile column, plus per (tile, triangle) the rows
//             that reach into the tile and an occlusion bound (without
//             checkpoints, triangles covering many tiles go on one list).
//   k_raster  256-thread workgroups that take tiles (256 pixels, one per
//             thread; 32x8, 16x16 or 8x32) from per-XCD queues.  Each pixel
//             keeps a 64-bit key in LDS that encodes the winning fragment; the
//             depth rule becomes an order-independent reduction (atomic min/max
//             of the key) because every fixed shader's blend replaces the old
//             pixel: the winner is a function of the set of fragments, with
//             primitive order as the tie-break (see init_key).  The (triangle,
//             row) pairs of a tile are spread over the lanes, each walking its
//             row exactly as triangles.rs:219-303 does (big triangles from
//             their checkpoints).  The winner is then shaded once and colour
//             and depth are written once per pixel.
//   k_line_*  the lines rasteriser (lines.rs:15-121) for one request.
#include "internal.hpp"
#include "shade.hpp"

#include <algorithm>

namespace euchip {

EUCHIP_DEV uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

EUCHIP_DEV bool passes_z_clip(const FrameParams &p, float z) {
    return !p.z_clip || (p.z_min <= z && z <= p.z_max);
}

// Inclusive wave64 prefix sum on DPP lane moves: shifts of 1, 2, 4 and 8 lanes
// inside each 16-lane row (lanes without a source add the `old` operand, 0),
// then row_bcast:15 adds row 0's (row 2's) total to row 1 (row 3) and
// row_bcast:31 adds lane 31's running total to rows 2 and 3.
EUCHIP_DEV uint32_t wave_inclusive_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15 -> rows 1, 3
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31 -> rows 2, 3
    return v;
}

// ---------------------------------------------------------------------------
// Visit every tile of each active lane's tile rectangle.  Small rectangles are
// walked by their own lane; rectangles of more than 8 tiles are spread over the
// whole wave so one huge triangle does not serialise a lane.  f(tile, j, a, b)
// gets the tile, its index j in the owner's rectangle and the owner's (a, b).
// ---------------------------------------------------------------------------
template <class F>
EUCHIP_DEV void for_each_tile(bool act, uint32_t tx0, uint32_t ty0, uint32_t nx, uint32_t ny, uint32_t ntx,
                              uint32_t a, uint32_t b, F f) {
    const uint32_t n = act ? nx * ny : 0u;
    const bool big = n > 8u;
    if (!big) {
        for (uint32_t j = 0; j < n; j++) {
            uint32_t jy = j / nx, jx = j - jy * nx;
            f((ty0 + jy) * ntx + tx0 + jx, j, a, b);
        }
    }
    uint64_t m = __ballot(big);
    const uint32_t lane = lane_id();
    while (m) {
        const int l = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        const uint32_t btx0 = __builtin_amdgcn_readlane(tx0, l);
        const uint32_t bty0 = __builtin_amdgcn_readlane(ty0, l);
        const uint32_t bnx = __builtin_amdgcn_readlane(nx, l);
        const uint32_t bn = __builtin_amdgcn_readlane(n, l);
        const uint32_t ba = __builtin_amdgcn_readlane(a, l);
        const uint32_t bb = __builtin_amdgcn_readlane(b, l);
        for (uint32_t j = lane; j < bn; j += 64u) {
            uint32_t jy = j / bnx, jx = j - jy * bnx;
            f((bty0 + jy) * ntx + btx0 + jx, j, ba, bb);
        }
    }
}

// Stream element of corner i of primitive t after the fixed geometry stage
// (Pipeline::geometry, pipeline.rs:221-229): output t % k of input triangle
// t / k, [a, b, c] or [a, c, b].
EUCHIP_DEV uint32_t corner_element(uint32_t geometry, uint32_t t, int i) {
    uint32_t tri = t, rev = 0u;
    if (geometry == GEOM_BOTH_SIDES) tri = t >> 1, rev = t & 1u;
    else if (geometry == GEOM_REVERSED) rev = 1u;
    const uint32_t c = (i == 0 || !rev) ? (uint32_t)i : 3u - (uint32_t)i;
    return 3u * tri + c;
}

// ---------------------------------------------------------------------------
// triangle setup (triangles.rs:55-171) of primitive t of request v, whole-target
// band clamp, from the vertex shader's clip positions.  Returns false when the
// triangle is culled or its bounding box is empty.
// ---------------------------------------------------------------------------
EUCHIP_DEV bool tri_setup(const FrameParams &p, const View &v, uint32_t t, TriRec &R) {
    float hom[3][4], euc[3][3];
    uint32_t sid[3], inst[3], vi[3];
    const float flip1 = p.y_up ? -1.0f : 1.0f;
    // the corners' clip positions (vertex(): VP * model * (p, 1), or the CLIP
    // input itself), the uniform branches taken outside the corner loop so the
    // loads of all three corners are in flight together
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const uint32_t e = corner_element(p.geometry, t, i);
        inst[i] = e / v.count;
        vi[i] = e - inst[i] * v.count;
    }
    if (v.idx) {
#pragma unroll
        for (int i = 0; i < 3; i++) vi[i] = v.idx[vi[i]];
    }
    float c[3][4];
    if (p.xform == 0u) {
        if (v.pos_comps == 4u) {
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const float *P = v.pos + 4u * vi[i];
                c[i][0] = P[0], c[i][1] = P[1], c[i][2] = P[2], c[i][3] = P[3];
            }
        } else {
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const float *P = v.pos + (size_t)v.pos_comps * vi[i];
                c[i][0] = P[0], c[i][1] = P[1], c[i][2] = P[2], c[i][3] = 1.0f;
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float *P = v.pos + (size_t)v.pos_comps * vi[i];
            const float p4[4] = {P[0], P[1], P[2], 1.0f};
            float w[4];
            mat4_vec4(v.models + 16u * inst[i], p4, w);
            mat4_vec4(v.u + U_VP, w, c[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        sid[i] = v.sv_base + inst[i] * v.n_vertices + vi[i];
        hom[i][0] = c[i][0] * 1.0f;
        hom[i][1] = c[i][1] * flip1;
        hom[i][2] = c[i][2];
        hom[i][3] = c[i][3];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        euc[i][0] = hom[i][0] / hom[i][3];
        euc[i][1] = hom[i][1] / hom[i][3];
        euc[i][2] = hom[i][2] / hom[i][3];
    }
    // winding = cross(e1 - e0, e2 - e0).z
    const float a0 = euc[1][0] - euc[0][0], a1 = euc[1][1] - euc[0][1];
    const float b0 = euc[2][0] - euc[0][0], b1 = euc[2][1] - euc[0][1];
    const float winding = a0 * b1 - a1 * b0;
    if (p.cull == 1u && winding * 1.0f < 0.0f) return false;
    if (p.cull == 2u && winding * -1.0f < 0.0f) return false;
    const bool swapped = winding >= 0.0f;
    if (swapped) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const float tmp = hom[0][c];
            hom[0][c] = hom[2][c];
            hom[2][c] = tmp;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float tmp = euc[0][c];
            euc[0][c] = euc[2][c];
            euc[2][c] = tmp;
        }
        const uint32_t ts = sid[0];
        sid[0] = sid[2];
        sid[2] = ts;
    }
    const float sx = (float)v.W, sy = (float)v.H;
    const float T[3][3] = {{v.ndc_x, 0.0f, -1.0f}, {0.0f, v.ndc_y, 1.0f}, {0.0f, 0.0f, 1.0f}};
    // coords_to_weights (triangles.rs:86-102)
    const float c3[3] = {hom[2][0], hom[2][1], hom[2][3]};
    const float ca[3] = {hom[0][0] - c3[0], hom[0][1] - c3[1], hom[0][3] - c3[2]};
    const float cb[3] = {hom[1][0] - c3[0], hom[1][1] - c3[1], hom[1][3] - c3[2]};
    const float n[3] = {ca[1] * cb[2] - ca[2] * cb[1], ca[2] * cb[0] - ca[0] * cb[2], ca[0] * cb[1] - ca[1] * cb[0]};
    const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    const float nc = n[0] * c3[0] + n[1] * c3[1] + n[2] * c3[2];
    const float rd = nn > 0.0f ? 1.0f / rmin(nc, -__FLT_EPSILON__) : 1.0f;
    float rows[3][3] = {{cb[1] * c3[2] - cb[2] * c3[1], cb[2] * c3[0] - cb[0] * c3[2], cb[0] * c3[1] - cb[1] * c3[0]},
                        {c3[1] * ca[2] - c3[2] * ca[1], c3[2] * ca[0] - c3[0] * ca[2], c3[0] * ca[1] - c3[1] * ca[0]},
                        {n[0], n[1], n[2]}};
    float M[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int j = 0; j < 3; j++) rows[i][j] = rows[i][j] * rd;
#pragma unroll
        for (int j = 0; j < 3; j++) M[i][j] = rows[i][0] * T[0][j] + rows[i][1] * T[1][j] + rows[i][2] * T[2][j];
    }
    // screen coordinates and bounds (triangles.rs:110-139); the band clamp of y
    // is applied per row by the raster pass, here the clamp is the whole target
    float vs[3][2];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        vs[i][0] = sx * (euc[i][0] * 0.5f + 0.5f);
        vs[i][1] = sy * (euc[i][1] * -0.5f + 0.5f);
    }
    R.xmin = sat_clamp(rmin(rmin(vs[0][0], vs[1][0]), vs[2][0]) + 0.0f, 0u, v.W);
    R.ymin = sat_clamp(rmin(rmin(vs[0][1], vs[1][1]), vs[2][1]) + 0.0f, 0u, v.H);
    R.xmax = sat_clamp(rmax(rmax(vs[0][0], vs[1][0]), vs[2][0]) + 1.0f, 0u, v.W);
    R.ymax = sat_clamp(rmax(rmax(vs[0][1], vs[1][1]), vs[2][1]) + 1.0f, 0u, v.H);
    // w_hom at (0,0) and the per-pixel steps (triangles.rs:142-145)
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float o = M[i][0] * 0.0f + M[i][1] * 0.0f + M[i][2] * 1.0f;
        const float wx = M[i][0] * 1000.0f + M[i][1] * 0.0f + M[i][2] * 1.0f;
        const float wy = M[i][0] * 0.0f + M[i][1] * 1000.0f + M[i][2] * 1.0f;
        R.o[i] = o;
        R.dx[i] = (wx - o) * (1.0f / 1000.0f);
        R.dy[i] = (wy - o) * (1.0f / 1000.0f);
        R.hz[i] = hom[i][2];
        R.sid[i] = sid[i];
    }
    // verts_by_y (triangles.rs:148-171), ties resolved as the reference does; kept in
    // scalars (a runtime index into a local array would live in scratch memory)
    const float x0 = vs[0][0], y0 = vs[0][1], x1 = vs[1][0], y1 = vs[1][1], x2 = vs[2][0], y2 = vs[2][1];
    const float min_y = rmin(rmin(y0, y1), y2);
    float ax, ay, bx, by, cx, cy;
    if (y0 == min_y) {
        ax = x0, ay = y0;
        if (y1 < y2) bx = x1, by = y1, cx = x2, cy = y2;
        else bx = x2, by = y2, cx = x1, cy = y1;
    } else if (y1 == min_y) {
        ax = x1, ay = y1;
        if (y0 < y2) bx = x0, by = y0, cx = x2, cy = y2;
        else bx = x2, by = y2, cx = x0, cy = y0;
    } else {
        ax = x2, ay = y2;
        if (y0 < y1) bx = x0, by = y0, cx = x1, cy = y1;
        else bx = x1, by = y1, cx = x0, cy = y0;
    }
    R.vby[0] = ax, R.vby[1] = ay, R.vby[2] = bx, R.vby[3] = by, R.vby[4] = cx, R.vby[5] = cy;
    const bool allin = passes_z_clip(p, euc[0][2]) && passes_z_clip(p, euc[1][2]) && passes_z_clip(p, euc[2][2]);
    R.flags = (allin ? TRI_ALL_IN : 0u) | (swapped ? TRI_SWAPPED : 0u);
    return R.xmin < R.xmax && R.ymin < R.ymax;
}

// A segment of n overflow entries of `tile`, taken from the pool and pushed on
// the tile's list: returns the pool index of its first entry, or ~0u when the
// pool ran out (the frame is then dropped and reported; the host re-runs a
// blocking render with a larger pool).
EUCHIP_DEV uint32_t take_segment(const FrameParams &p, uint32_t tile, uint32_t n) {
    const uint32_t words = n + 2u;
    const uint32_t base = atomicAdd(&p.ctr->pool[p.frame & 1u][0], words);
    if (base >= p.pool_cap || p.pool_cap - base < words) {
        atomicOr(&p.ctr->flags[p.frame & 1u][0], FLAG_BIN_OVERFLOW);
        return ~0u;
    }
    p.pool[base] = atomicExch(&p.tile_count[tile].head, base + 1u);
    p.pool[base + 1u] = n;
    return base + 2u;
}

// Bin slot of rank `rank` (below the capacity) of tile `tile` (launch tile space)
// of request v.
EUCHIP_DEV uint32_t *bin_slot(const FrameParams &p, const View &v, uint32_t tile, uint32_t rank) {
    return p.bins + v.bin_base + (size_t)(tile - v.tile_base) * v.tile_cap + rank;
}

// Range [beg, end) of this XCD's tiles: the hardware deals workgroups
// round-robin to the 8 XCDs, so workgroup b runs on XCD b % 8; each XCD gets a
// contiguous run of tiles so triangles shared by neighbouring tiles stay in
// that XCD's L2.
EUCHIP_DEV void xcd_range(uint32_t ntiles, uint32_t xcd, uint32_t &beg, uint32_t &end) {
    const uint32_t q = ntiles >> 3, r = ntiles & 7u;
    beg = xcd * q + (xcd < r ? xcd : r);
    end = beg + q + (xcd < r ? 1u : 0u);
}

// Busy list x: BUSY_BLOCKS lists hold up to ntiles entries at p.busy + x *
// ntiles; BUSY_RANGES list x holds the tiles of range x in that range's own slots.
EUCHIP_DEV void busy_list_span(const FrameParams &p, uint32_t x, uint32_t &base, uint32_t &cap) {
    if (p.busy_map == BUSY_BLOCKS) {
        base = x * p.ntiles, cap = p.ntiles;
    } else {
        uint32_t e;
        xcd_range(p.ntiles, x, base, e);
        cap = e - base;
    }
}

// Busy list of a tile.
EUCHIP_DEV uint32_t busy_list_of(const FrameParams &p, uint32_t tile) {
    if (p.busy_map == BUSY_BLOCKS) {  // one request: blocks of 4 x 4 tiles dealt over the XCDs
        const uint32_t ty = tile / p.v0.ntx, tx = tile - ty * p.v0.ntx;
        return ((tx >> 2) + 3u * (ty >> 2)) & 7u;
    }
    // the XCD whose contiguous tile range (xcd_range) holds the tile
    const uint32_t q = p.ntiles >> 3, r = p.ntiles & 7u;
    const uint32_t big = r * (q + 1u);  // tiles of the ranges one longer
    return tile < big ? tile / (q + 1u) : r + (tile - big) / q;
}

// A tile whose count went from `old` to `old + n` may have become heavy.
EUCHIP_DEV void note_heavy(const FrameParams &p, uint32_t tile, uint32_t old, uint32_t n) {
    if (p.heavy_cap && old < kHeavyTile && old + n >= kHeavyTile) {
        const uint32_t k = atomicAdd(&p.ctr->heavy[p.frame & 1u][0], 1u);
        if (k < p.heavy_cap) p.heavy[k] = tile;
    }
}

// Bin append of one pair by a global atomic (the fallback of the LDS-grouped
// path).  maxp: where counts past the capacity are recorded for the host.
EUCHIP_DEV void bin_direct(const FrameParams &p, const View &v, uint32_t tile, uint32_t prim, uint32_t *maxp) {
    const uint32_t rank = atomicAdd(&p.tile_count[tile].count, 1u);
    note_heavy(p, tile, rank, 1u);
    if (p.busy && rank == 0u) {
        const uint32_t x = busy_list_of(p, tile);
        uint32_t lb, lc;
        busy_list_span(p, x, lb, lc);
        const uint32_t k = atomicAdd(&p.ctr->busy[p.frame & 1u][x][0], 1u);
        if (k < lc) p.busy[(size_t)lb + k] = tile;
    }
    if (rank < v.tile_cap) {
        *bin_slot(p, v, tile, rank) = prim;
    } else {
        atomicMax(maxp, rank + 1u);
        const uint32_t e = take_segment(p, tile, 1u);
        if (e != ~0u) p.pool[e] = prim;
    }
}

// The request a workgroup or tile works on: the kernel argument's, or one of
// the batch's in device memory.
template <bool BATCH>
EUCHIP_DEV const View &view_of(const FrameParams &p, uint32_t r) {
    if constexpr (BATCH) return p.views[r];
    else return p.v0;
}

// Stores that bypass the caches on their way to memory: setup records (read
// back by another launch, usually on another XCD) and the targets (not read
// again by this frame).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

EUCHIP_DEV void store_stream(uint4 *dst, const uint4 &v) {
    u32x4 x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<u32x4 *>(dst));
}
EUCHIP_DEV void store_stream(uint32_t *dst, uint32_t v) { __builtin_nontemporal_store(v, dst); }
EUCHIP_DEV void store_stream(float *dst, float v) { __builtin_nontemporal_store(v, dst); }

// Setup records (layout in internal.hpp).
EUCHIP_DEV void pack_rec(const TriRec &R, uint4 q0[kRec0Quads], uint4 q1[kRec1Quads]) {
    q0[0] = make_uint4(__float_as_uint(R.o[0]), __float_as_uint(R.o[1]), __float_as_uint(R.o[2]), __float_as_uint(R.dx[0]));
    q0[1] = make_uint4(__float_as_uint(R.dx[1]), __float_as_uint(R.dx[2]), __float_as_uint(R.dy[0]), __float_as_uint(R.dy[1]));
    q0[2] = make_uint4(__float_as_uint(R.dy[2]), __float_as_uint(R.hz[0]), __float_as_uint(R.hz[1]), __float_as_uint(R.hz[2]));
    q0[3] = make_uint4(R.xmin, R.xmax, R.ymin, R.ymax | (R.flags << 30));
    q1[0] = make_uint4(__float_as_uint(R.vby[0]), __float_as_uint(R.vby[1]), __float_as_uint(R.vby[2]), __float_as_uint(R.vby[3]));
    q1[1] = make_uint4(__float_as_uint(R.vby[4]), __float_as_uint(R.vby[5]), 0u, 0u);
}

// o, dx, dy, hz, bounds and flags from a record's first kRec0Quads quads.
EUCHIP_DEV void load_rec0_at(const uint4 *q, TriRec &R) {
    const uint4 a = q[0], b = q[1], c = q[2], d = q[3];
    R.o[0] = __uint_as_float(a.x), R.o[1] = __uint_as_float(a.y), R.o[2] = __uint_as_float(a.z);
    R.dx[0] = __uint_as_float(a.w), R.dx[1] = __uint_as_float(b.x), R.dx[2] = __uint_as_float(b.y);
    R.dy[0] = __uint_as_float(b.z), R.dy[1] = __uint_as_float(b.w), R.dy[2] = __uint_as_float(c.x);
    R.hz[0] = __uint_as_float(c.y), R.hz[1] = __uint_as_float(c.z), R.hz[2] = __uint_as_float(c.w);
    R.xmin = d.x, R.xmax = d.y, R.ymin = d.z, R.ymax = d.w & 0x3fffffffu, R.flags = d.w >> 30;
}

// o, dx, dy, hz, bounds and flags of record `rid`.
EUCHIP_DEV void load_rec0(const FrameParams &p, uint32_t rid, TriRec &R) {
    load_rec0_at(p.rec0 + (size_t)rid * kRec0Quads, R);
}

// verts_by_y from a record's kRec1Quads secondary quads.
EUCHIP_DEV void load_rec1_at(const uint4 *q, TriRec &R) {
    const uint4 a = q[0], b = q[1];
    R.vby[0] = __uint_as_float(a.x), R.vby[1] = __uint_as_float(a.y), R.vby[2] = __uint_as_float(a.z);
    R.vby[3] = __uint_as_float(a.w), R.vby[4] = __uint_as_float(b.x), R.vby[5] = __uint_as_float(b.y);
}

// verts_by_y of record `rid`.
EUCHIP_DEV void load_rec1(const FrameParams &p, uint32_t rid, TriRec &R) {
    load_rec1_at(p.rec1 + (size_t)rid * kRec1Quads, R);
}

// Shaded-vertex ids of primitive t's corners after the winding swap.
EUCHIP_DEV void tri_sids(const View &v, uint32_t geometry, uint32_t t, uint32_t flags, uint32_t sid[3]) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const uint32_t e = corner_element(geometry, t, i);
        const uint32_t inst = e / v.count;
        const uint32_t k = e - inst * v.count;
        sid[i] = v.sv_base + inst * v.n_vertices + (v.idx ? v.idx[k] : k);
    }
    if (flags & TRI_SWAPPED) {
        const uint32_t s0 = sid[0];
        sid[0] = sid[2];
        sid[2] = s0;
    }
}

// A triangle's rows all take the bounding-box span ("stupid version",
// triangles.rs:229-233) when its whole-target box is small; band clamping can
// only shrink the box, so then no row needs verts_by_y.
EUCHIP_DEV bool all_rows_boxed(const TriRec &R) {
    const uint32_t bw = R.xmax - R.xmin, bh = R.ymax - R.ymin;
    return bw < 128u && bh < 128u && bw * bh < 128u;
}

// Fragment test of one pixel (triangles.rs:262-301): weights, coverage, z, z-clip.
EUCHIP_DEV bool frag_z(const FrameParams &p, const TriRec &R, float w0, float w1, float w2, float &z) {
    const float u2 = (w2 - w0) - w1;
    if (!(w0 >= 0.0f && w1 >= 0.0f && u2 >= 0.0f)) return false;
    z = (R.hz[0] * w0 + R.hz[1] * w1) + R.hz[2] * u2;
    return (R.flags & TRI_ALL_IN) || passes_z_clip(p, z);
}

// Triangles whose box has at most this many pixels are tested for coverage in
// setup; one that covers no pixel corner emits no fragment and is not binned.
constexpr uint32_t kCoverTestArea = 8;

// Does a boxed triangle (all_rows_boxed) cover any pixel of its box, with the
// raster pass's arithmetic (rows walked from xmin, triangles.rs:257-301)?
EUCHIP_DEV bool covers_any(const FrameParams &p, const TriRec &R) {
    for (uint32_t y = R.ymin; y < R.ymax; y++) {
        const float Y = (float)y, X0 = (float)R.xmin;
        float w0 = (R.o[0] + R.dy[0] * Y) + R.dx[0] * X0;
        float w1 = (R.o[1] + R.dy[1] * Y) + R.dx[1] * X0;
        float w2 = (R.o[2] + R.dy[2] * Y) + R.dx[2] * X0;
        for (uint32_t x = R.xmin; x < R.xmax; x++) {
            float z;
            if (frag_z(p, R, w0, w1, w2, z)) return true;
            w0 = w0 + R.dx[0];
            w1 = w1 + R.dx[1];
            w2 = w2 + R.dx[2];
        }
    }
    return false;
}

constexpr uint32_t kHash = 512;    // LDS hash slots (distinct tiles) per setup workgroup
constexpr uint32_t kPairs = 1024;  // (tile, primitive) pairs grouped in LDS per workgroup
static_assert(kHash == 512u && kPairs <= (1u << 23), "pairs pack a 9-bit hash slot and their rank in the slot");

// ---------------------------------------------------------------------------
// k_setup: per-vertex clip position + VertexData (thread t = shaded vertex t)
// and per-triangle culling + binning (thread t = primitive t).
// Binning groups the workgroup's pairs by tile in an LDS hash table, so each
// distinct tile costs one global atomic per workgroup; ranks inside the
// group come from LDS.
// ---------------------------------------------------------------------------
// A span block of a checkpoint frame's setup launch (defined with the raster's
// row helpers below).
EUCHIP_DEV void span_item(const FrameParams &p, uint32_t b, void *static_lds);

// LDS of k_setup: the workgroup's records on their way out, then the binning tables.
union SetupShared {
    struct {
        uint4 q0[256 * kRec0Quads];
        uint32_t kept[256];
    } rec;
    struct {
        uint32_t hkey[kHash], hcnt[kHash], hbase[kHash];
        uint32_t hseg[kHash];                  // pool index of the slot's overflow entries
        uint32_t pl_prim[kPairs], pl_sl[kPairs];  // pair: primitive; hash slot | rank in the slot << 9
    } bin;
};

// SPAN: the launch of a checkpoint frame, whose blocks past setup_blocks are
// span blocks (compiled apart so other frames' setup carries none of their code).
template <uint32_t SH, bool BATCH, bool SPAN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_setup(FrameParams p) {
    __shared__ SetupShared sm;
    __shared__ uint32_t npl;
    __shared__ uint32_t lcnt[9], lbase[9];  // new busy tiles per list (8) and new heavy tiles (8)
    const uint32_t tid = threadIdx.x;
    if constexpr (!BATCH) {
        if (blockIdx.x >= p.tri_blocks) {  // (checkpoint frames: tri_blocks = setup_blocks)
            if constexpr (SPAN) {
                span_item(p, blockIdx.x - p.setup_blocks, &sm);
            } else if constexpr (vd_count(SH) > 0) {  // a vertex block
                const View &v = p.v0;
                const uint32_t t = (blockIdx.x - p.tri_blocks) * 256u + tid;
                if (t < v.n_instances * v.n_vertices) {
                    const uint32_t inst = t / v.n_vertices;
                    float vd[VDN<SH>::N];
                    vertex_data<SH>(p, v, inst, t - inst * v.n_vertices, vd);
#pragma unroll
                    for (uint32_t q = 0; q < vd_count(SH); q++) p.vary[((size_t)v.sv_base + t) * vd_count(SH) + q] = vd[q];
                }
            }
            return;
        }
    }
    uint32_t r = 0, j = blockIdx.x;
    if constexpr (BATCH) {
        const uint2 bj = p.blk[blockIdx.x];
        r = bj.x, j = bj.y;
    }
    const View &v = view_of<BATCH>(p, r);
    const uint32_t t = j * 256u + tid;
    const uint32_t sx = p.tile_sx, sy = 8u - p.tile_sx;
    bool keep = false;
    uint32_t tx0 = 0, ty0 = 0, nx = 0, ny = 0;
    if (t < v.n_tris) {
        TriRec R;
        keep = tri_setup(p, v, t, R);
        if (keep && all_rows_boxed(R) && (R.xmax - R.xmin) * (R.ymax - R.ymin) <= kCoverTestArea)
            keep = covers_any(p, R);
        if (keep) {
            tx0 = R.xmin >> sx;
            ty0 = R.ymin >> sy;
            nx = ((R.xmax - 1u) >> sx) - tx0 + 1u;
            ny = ((R.ymax - 1u) >> sy) - ty0 + 1u;
            if (p.ck_on) {
                // checkpoint frames: the span blocks walk every triangle (a lane
                // binning a triangle of many tiles here would take them one by
                // one, each behind dependent LDS atomics)
                nx = ny = 0u;
            } else if (p.big_cap && nx * ny > kBigTiles) {  // on the big-triangle list (while it has room)
                const uint32_t k = atomicAdd(&p.ctr->big[p.frame & 1u][0], 1u);
                if (k < p.big_cap) {
                    p.big[k] = t;
                    nx = ny = 0u;
                }
            }
            uint4 q0[kRec0Quads], q1[kRec1Quads];
            pack_rec(R, q0, q1);
#pragma unroll
            for (uint32_t k = 0; k < kRec0Quads; k++) sm.rec.q0[tid * kRec0Quads + k] = q0[k];
            // rec1: this lane's 32 bytes, contiguous with its neighbours' (no
            // staging); only triangles with a span row read it (all_rows_boxed)
            if (!all_rows_boxed(R)) {
#pragma unroll
                for (uint32_t k = 0; k < kRec1Quads; k++)
                    store_stream(&p.rec1[((size_t)v.tri_base + t) * kRec1Quads + k], q1[k]);
            }
        }
    }
    sm.rec.kept[tid] = keep ? 1u : 0u;
    // Batches and checkpoint frames: the vertex stage after the triangles' loads
    // (the VertexData stores would otherwise order those loads behind them);
    // other one-request frames have vertex blocks of their own.
    if constexpr ((BATCH || SPAN) && vd_count(SH) > 0) {
        if (t < v.n_instances * v.n_vertices) {
            const uint32_t inst = t / v.n_vertices;
            const uint32_t vi = t - inst * v.n_vertices;
            const size_t sv = (size_t)v.sv_base + t;
            float vd[VDN<SH>::N];
            vertex_data<SH>(p, v, inst, vi, vd);
#pragma unroll
            for (uint32_t q = 0; q < vd_count(SH); q++) p.vary[sv * vd_count(SH) + q] = vd[q];
        }
    }
    __syncthreads();
    // the workgroup's records are contiguous: write them out 16 bytes per lane,
    // skipping the culled triangles' slots
    {
        const size_t base = (size_t)v.tri_base + (size_t)j * 256u;
#pragma unroll
        for (uint32_t k = 0; k < kRec0Quads; k++) {
            const uint32_t q = k * 256u + tid;
            if (sm.rec.kept[q / kRec0Quads]) store_stream(&p.rec0[base * kRec0Quads + q], sm.rec.q0[q]);
        }
    }
    __syncthreads();  // the binning tables reuse the record staging
    for (uint32_t i = tid; i < kHash; i += 256u) sm.bin.hkey[i] = 0u, sm.bin.hcnt[i] = 0u;
    if (tid == 0u) npl = 0u;
    __syncthreads();
    const uint32_t tbase = v.tile_base;
    uint32_t *maxp = BATCH ? p.req_max + r : &p.ctr->max_count[p.frame & 1u][0];
    for_each_tile(keep, tx0, ty0, nx, ny, v.ntx, t, 0u, [&](uint32_t lt, uint32_t, uint32_t prim, uint32_t) {
        const uint32_t tile = tbase + lt;
        uint32_t h = (tile * 2654435761u) >> 23;  // 9-bit multiplicative hash
        for (uint32_t probe = 0; probe < 16u; probe++) {
            const uint32_t old = atomicCAS(&sm.bin.hkey[h], 0u, tile + 1u);
            if (old == 0u || old == tile + 1u) {
                const uint32_t k = atomicAdd(&npl, 1u);
                if (k < kPairs) {
                    sm.bin.pl_prim[k] = prim;
                    sm.bin.pl_sl[k] = h | (atomicAdd(&sm.bin.hcnt[h], 1u) << 9);
                    return;
                }
                break;
            }
            h = (h + 1u) & (kHash - 1u);
        }
        bin_direct(p, v, tile, prim, maxp);  // table or pair list full
    });
    __syncthreads();
    // One global atomic per distinct tile; tiles that became busy or heavy get a
    // rank among this workgroup's new ones (kept in hcnt, no longer needed):
    // bits 0-10 busy rank, 11-13 busy list, 14 busy, 15-25 heavy rank, 26 heavy.
    if (tid < 9u) lcnt[tid] = 0u;
    __syncthreads();
    for (uint32_t i = tid; i < kHash; i += 256u) {
        const uint32_t n = sm.bin.hcnt[i];
        if (n) {
            const uint32_t tile = sm.bin.hkey[i] - 1u;
            const uint32_t old = atomicAdd(&p.tile_count[tile].count, n);
            sm.bin.hbase[i] = old;
            // ranks past the capacity: one overflow segment for this workgroup's share
            if (old + n > v.tile_cap) {
                atomicMax(maxp, old + n);
                const uint32_t first = old > v.tile_cap ? old : v.tile_cap;
                sm.bin.hseg[i] = take_segment(p, tile, old + n - first);
            }
            uint32_t code = 0u;
            if (p.busy && old == 0u) {
                const uint32_t x = busy_list_of(p, tile);
                code |= atomicAdd(&lcnt[x], 1u) | (x << 11) | (1u << 14);
            }
            if (p.heavy_cap && old < kHeavyTile && old + n >= kHeavyTile) code |= (atomicAdd(&lcnt[8], 1u) << 15) | (1u << 26);
            sm.bin.hcnt[i] = code;
        }
    }
    __syncthreads();
    // list bases (one global atomic per list) in flight while the bins are written
    uint32_t any_new = 0u;
#pragma unroll
    for (uint32_t q = 0; q < 9u; q++) any_new |= lcnt[q];
    uint32_t my_base = 0u;
    if (tid < 9u && lcnt[tid]) {
        my_base = tid < 8u ? atomicAdd(&p.ctr->busy[p.frame & 1u][tid][0], lcnt[tid])
                           : atomicAdd(&p.ctr->heavy[p.frame & 1u][0], lcnt[tid]);
    }
    const uint32_t np = npl < kPairs ? npl : kPairs;
    for (uint32_t k = tid; k < np; k += 256u) {
        const uint32_t h = sm.bin.pl_sl[k] & (kHash - 1u), rank = sm.bin.hbase[h] + (sm.bin.pl_sl[k] >> 9);
        if (rank < v.tile_cap) {
            *bin_slot(p, v, sm.bin.hkey[h] - 1u, rank) = sm.bin.pl_prim[k];
        } else {
            const uint32_t base = sm.bin.hbase[h], first = base > v.tile_cap ? base : v.tile_cap;
            const uint32_t e = sm.bin.hseg[h];
            if (e != ~0u) p.pool[e + (rank - first)] = sm.bin.pl_prim[k];
        }
    }
    if (!any_new) return;  // no tile became busy or heavy here (uniform)
    if (tid < 9u) lbase[tid] = my_base;
    __syncthreads();
    for (uint32_t i = tid; i < kHash; i += 256u) {
        const uint32_t code = sm.bin.hcnt[i];
        if (sm.bin.hkey[i] == 0u || code == 0u) continue;
        const uint32_t tile = sm.bin.hkey[i] - 1u;
        if (code & (1u << 14)) {
            const uint32_t x = (code >> 11) & 7u, k = lbase[x] + (code & 0x7ffu);
            uint32_t lb, lc;
            busy_list_span(p, x, lb, lc);
            if (k < lc) p.busy[(size_t)lb + k] = tile;
        }
        if (code & (1u << 26)) {
            const uint32_t k = lbase[8] + ((code >> 15) & 0x7ffu);
            if (k < p.heavy_cap) p.heavy[k] = tile;
        }
    }
}

// ---------------------------------------------------------------------------
// raster
// ---------------------------------------------------------------------------

// Total order of f32 for the winner keys; -0 and +0 map to one value.
EUCHIP_DEV uint32_t zord(float z) {
    const uint32_t u = __float_as_uint(z);
    if ((u & 0x7fffffffu) == 0u) return 0x80000000u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
EUCHIP_DEV float zunord(uint32_t o, uint32_t negzero) {
    if (o == 0x80000000u) return negzero ? -0.0f : 0.0f;
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// Winner keys.  mode 1 (test Less, write): min of (zord(z), prim+1, sign of zero)
// = the first fragment of the smallest z below D0.  mode 2 (Greater, write):
// the same with ~zord.  mode 0 (every other DepthMode): max of (prim+1, z bits)
// over fragments passing the test against D0 = the last passing fragment.
// In the reference's sequential order (pipeline.rs:519-538) these are exactly
// the fragments whose emit_fragment determines the final pixel and depth.
EUCHIP_DEV uint64_t init_key(uint32_t mode, float D0) {
    if (mode == 0u) return 0ull;
    if (D0 != D0) return 0ull;  // nothing compares with NaN
    const uint32_t o = mode == 1u ? zord(D0) : ~zord(D0);
    return (uint64_t)o << 32;
}

EUCHIP_DEV void update_key(uint64_t *key, uint32_t mode, uint32_t test, float z, float D0, uint32_t prim) {
    if (mode == 0u) {
        bool pass;
        switch (test) {
        case DT_LESS: pass = z < D0; break;
        case DT_EQUAL: pass = z == D0; break;
        case DT_GREATER: pass = z > D0; break;
        default: pass = true; break;
        }
        if (pass) atomicMax((unsigned long long *)key, ((uint64_t)(prim + 1u) << 32) | __float_as_uint(z));
    } else {
        if (z != z) return;
        const uint32_t o = mode == 1u ? zord(z) : ~zord(z);
        const uint32_t nz = __float_as_uint(z) == 0x80000000u ? 1u : 0u;
        atomicMin((unsigned long long *)key, ((uint64_t)o << 32) | ((prim + 1u) << 1) | nz);
    }
}

// Row span [x0, x1) of row y for a triangle (triangles.rs:219-254); the
// "stupid version" switch uses the band-clamped bounding box.
EUCHIP_DEV void row_span(const TriRec &R, uint32_t y, uint32_t by0, uint32_t by1, uint32_t &x0, uint32_t &x1) {
    const uint32_t bw = R.xmax - R.xmin;
    const uint32_t yl = R.ymin < by0 ? by0 : (R.ymin > by1 ? by1 : R.ymin);
    const uint32_t yh = R.ymax < by0 ? by0 : (R.ymax > by1 ? by1 : R.ymax);
    const uint32_t bh = yh - yl;
    if (bw < 128u && bh < 128u && bw * bh < 128u) {
        x0 = R.xmin;
        x1 = R.xmax;
        return;
    }
    const float Y = (float)y;
    const float ax = R.vby[0], ay = R.vby[1], bx = R.vby[2], byy = R.vby[3], cx = R.vby[4], cy = R.vby[5];
    const float ac = ax + ((Y - ay) / (cy - ay)) * (cx - ax);
    float lo, hi;
    if (Y < byy) {
        const float ab = ax + ((Y - ay) / (byy - ay)) * (bx - ax);
        lo = rmin(ab, ac);
        hi = rmax(ab, ac);
    } else {
        const float bc = bx + ((Y - byy) / (cy - byy)) * (cx - bx);
        lo = rmin(bc, ac);
        hi = rmax(bc, ac);
    }
    const float fl = floorf(lo), ce = ceilf(hi);
    const float fxmin = (float)R.xmin, fxmax = (float)R.xmax;
    x0 = (fl >= fxmin && fl < fxmax) ? (uint32_t)fl : R.xmin;
    x1 = (ce >= fxmin && ce < fxmax) ? (uint32_t)ce : R.xmax;
}

EUCHIP_DEV void band_of(const FrameParams &p, const View &v, uint32_t y, uint32_t &by0, uint32_t &by1) {
    if (p.band_par) {
        by0 = (y / v.G) * v.G;
        by1 = by0 + v.G < v.H ? by0 + v.G : v.H;
    } else {
        by0 = 0;
        by1 = v.H;
    }
}

template <uint32_t SH>
EUCHIP_DEV void v_data(const TriRec &R, const float (*V)[VDN<SH>::N], float X, float Y, float *vd) {
    float w[3];
#pragma unroll
    for (int i = 0; i < 3; i++) w[i] = (R.o[i] + R.dy[i] * Y) + R.dx[i] * X;
    const float u2 = (w[2] - w[0]) - w[1];
    const float r = 1.0f / w[2];
    const float n0 = w[0] * r, n1 = w[1] * r, n2 = u2 * r;
    if constexpr (SH == FLAT) {
        vd[0] = V[0][0];
    } else {
#pragma unroll
        for (uint32_t c = 0; c < vd_count(SH); c++) vd[c] = (V[0][c] * n0 + V[1][c] * n1) + V[2][c] * n2;
    }
}

// Deferred shading of one pixel's winning fragment (pipeline.rs:539-576).
// MS: the frame's MSAA level is above 0 (coarse shading); the one-sample and
// the coarse kernels are compiled apart so the first keep the registers the
// second needs for its four samples.
template <uint32_t SH, bool GS, bool MS>
EUCHIP_DEV uint32_t shade_pixel(const FrameParams &p, const View &v, const TriRec &R, uint32_t px, uint32_t py,
                                uint32_t by0, uint32_t old) {
    constexpr uint32_t NV = VDN<SH>::N;
    float V[3][NV];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float *src = p.vary + (size_t)R.sid[i] * vd_count(SH);
#pragma unroll
        for (uint32_t c = 0; c < vd_count(SH); c++) V[i][c] = src[c];
    }
    float frag[4], vd[NV];
    const uint32_t L = p.msaa_level;
    if constexpr (!MS) {
        v_data<SH>(R, V, (float)px, (float)py, vd);
        fragment_shader<SH, GS>(v, p.sampler, vd, frag);
    } else {
        // coarse shading on the band-relative 2^L grid with bilinear reconstruction
        // (pipeline.rs:541-570); tgt_min = (0, by0)
        const float div = 1.0f / (float)(1u << L);
        const uint32_t rx = px, ry = py - by0;
        const float fx = rfract((float)rx * div);
        const float fy = rfract((float)ry * div);
        const uint32_t cx = rx >> L, cy = ry >> L;
        float t[4][4];  // t00, t10, t01, t11
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t ix = cx + (uint32_t)(q & 1), iy = cy + (uint32_t)(q >> 1);
            v_data<SH>(R, V, (float)(ix << L), (float)(by0 + (iy << L)), vd);
            fragment_shader<SH, GS>(v, p.sampler, vd, t[q]);
        }
        if constexpr (SH == FLAT) {
            frag[0] = t[0][0];
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const float t0 = t[0][c] * (1.0f - fy) + t[2][c] * fy;
                const float t1 = t[1][c] * (1.0f - fy) + t[3][c] * fy;
                frag[c] = t0 * (1.0f - fx) + t1 * fx;
            }
        }
    }
    return blend_shader<SH>(old, frag, p.bgra);
}

// Tile geometry: TW x TH pixels at (X0, Y0); TW = 1 << sx, TH = 256 >> sx.
struct TileGeom {
    uint32_t X0, Y0, sx, TW, TH;
};

// The tile's pixels [xs, xe) of one row (LDS index pi0 of the first), w_hom at
// xs given: each pixel's fragment updates its winner key.  Rows of small
// triangles in key mode 0 (the test reads the depth before the pass); the
// write modes take walk_span's loop.
template <uint32_t MODE>
EUCHIP_DEV void walk_row(const FrameParams &p, const TriRec &R, uint32_t prim, uint32_t xs, uint32_t xe, uint32_t pi0,
                         float w0, float w1, float w2, uint64_t *key, const float *d0s) {
    for (uint32_t x = xs; x < xe; x++) {
        float z;
        if (frag_z(p, R, w0, w1, w2, z)) {
            const uint32_t pi = pi0 + (x - xs);
            update_key(key + pi, MODE, p.depth_test, z, MODE == 0u ? d0s[pi] : 0.0f, prim);
        }
        w0 = w0 + R.dx[0];
        w1 = w1 + R.dx[1];
        w2 = w2 + R.dx[2];
    }
}

// The write modes' row loop: coverage, z, z clip and the key of every pixel,
// then one LDS atomic min per pixel, with no branch: an uncovered pixel's
// atomic takes the all-ones key, which changes nothing.  The loop holds no LDS
// read, so nothing waits on the atomics of earlier pixels.
//   * A NaN weight makes z NaN (a NaN factor or term stays NaN), so z == z and
//     the minimum of the three weights >= 0 decide exactly what the three
//     tests w >= 0 and z == z decide.
//   * -0 + 0 = +0 gives both zeros one order value; z + 0 differs from z only
//     for z = -0 (NaN z never wins), in the sign bit alone, so that bit of
//     z ^ (z + 0) shifted in beside (prim + 1) keeps the sign in the key.
//   * An uncovered pixel's key has an all-ones high word: no stored key has
//     one (order values of non-NaN depths stay below it), so it never wins.
//   * PAIRS: two pixels per turn (the second atomic at the next address)
//     halve the loop control; a row of odd length ends with one more pixel.
//     It pays for long rows only.
template <uint32_t MODE, bool CLIP, bool PAIRS>
EUCHIP_DEV void walk_keys(const TriRec &R, bool clip, float zlo, float zhi, uint32_t pid, unsigned long long *kp,
                          unsigned long long *const ke, float w0, float w1, float w2) {
    auto pixel = [&](unsigned long long *at) {
        const float u2 = (w2 - w0) - w1;
        const float z = (R.hz[0] * w0 + R.hz[1] * w1) + R.hz[2] * u2;
        bool ok = (__builtin_fminf(__builtin_fminf(w0, w1), u2) >= 0.0f) & (z == z);
        if constexpr (CLIP) ok = ok & (!clip | ((zlo <= z) & (z <= zhi)));
        const uint32_t bz = __float_as_uint(z), u = __float_as_uint(z + 0.0f);
        const uint32_t o = u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
        const uint32_t lo = __builtin_amdgcn_alignbit(pid, bz ^ u, 31u);  // (pid << 1) | (z is -0)
        const uint32_t hi = ok ? (MODE == 1u ? o : ~o) : ~0u;
        atomicMin(at, ((uint64_t)hi << 32) | lo);
        w0 = w0 + R.dx[0];
        w1 = w1 + R.dx[1];
        w2 = w2 + R.dx[2];
    };
    if constexpr (PAIRS) {
        const uint32_t n = (uint32_t)(ke - kp);
        unsigned long long *const pairs_end = kp + (n & ~1u);
        for (; kp < pairs_end; kp += 2) {
            pixel(kp);
            pixel(kp + 1);
        }
        if (n & 1u) pixel(kp);
    } else {
        for (; kp < ke; kp++) pixel(kp);
    }
}

// walk_row without branches on the coverage; the write modes take walk_keys,
// with the z clip only where a lane of the wave needs it.
template <uint32_t MODE, bool PAIRS = false>
EUCHIP_DEV void walk_span(const FrameParams &p, const TriRec &R, uint32_t prim, uint32_t xs, uint32_t xe, uint32_t pi0,
                          float w0, float w1, float w2, uint64_t *key, const float *d0s) {
    const bool clip = !(R.flags & TRI_ALL_IN) && p.z_clip;
    if constexpr (MODE != 0u) {
        unsigned long long *kp = (unsigned long long *)(key + pi0);
        unsigned long long *const ke = kp + (xe - xs);
        if (__builtin_amdgcn_ballot_w64(clip) != 0ull)
            walk_keys<MODE, true, PAIRS>(R, clip, p.z_min, p.z_max, prim + 1u, kp, ke, w0, w1, w2);
        else
            walk_keys<MODE, false, PAIRS>(R, clip, p.z_min, p.z_max, prim + 1u, kp, ke, w0, w1, w2);
        return;
    }
    const float zlo = p.z_min, zhi = p.z_max;
    const uint32_t test = p.depth_test;
    const uint32_t pid = prim + 1u;
    for (uint32_t x = xs; x < xe; x++) {
        const float u2 = (w2 - w0) - w1;
        const float z = (R.hz[0] * w0 + R.hz[1] * w1) + R.hz[2] * u2;
        bool ok = (w0 >= 0.0f) & (w1 >= 0.0f) & (u2 >= 0.0f);
        ok = ok & (!clip | ((zlo <= z) & (z <= zhi)));
        const uint32_t pi = pi0 + (x - xs);
        const float D0 = d0s[pi];
        const bool pass = test == DT_LESS ? z < D0 : test == DT_EQUAL ? z == D0 : test == DT_GREATER ? z > D0 : true;
        if (ok & pass) atomicMax((unsigned long long *)(key + pi), ((uint64_t)pid << 32) | __float_as_uint(z));
        w0 = w0 + R.dx[0];
        w1 = w1 + R.dx[1];
        w2 = w2 + R.dx[2];
    }
}

// One row y of a triangle inside the tile: the row span (triangles.rs:219-254),
// then the walk from the row start exactly as the reference does
// (triangles.rs:257-301), testing the tile's pixels of the span.
template <uint32_t SKIP_UNROLL>
EUCHIP_DEV void raster_row(const FrameParams &p, const TriRec &R, uint32_t prim, uint32_t y, const TileGeom &g,
                           uint32_t mode, uint64_t *key, const float *d0s, const uint32_t *band) {
    uint32_t x0, x1;
    const uint32_t r = y - g.Y0;
    row_span(R, y, band[2 * r], band[2 * r + 1], x0, x1);
    const uint32_t xs = x0 > g.X0 ? x0 : g.X0;
    const uint32_t xe = x1 < g.X0 + g.TW ? x1 : g.X0 + g.TW;
    if (xs >= xe) return;
    const float Y = (float)y, X0 = (float)x0;
    float w0 = (R.o[0] + R.dy[0] * Y) + R.dx[0] * X0;
    float w1 = (R.o[1] + R.dy[1] * Y) + R.dx[1] * X0;
    float w2 = (R.o[2] + R.dy[2] * Y) + R.dx[2] * X0;
    // to the tile's first pixel, one addition per pixel as the reference walks
    // (SKIP_UNROLL pixels per turn: the same additions, less loop control)
    uint32_t x = x0;
    for (; x + SKIP_UNROLL <= xs; x += SKIP_UNROLL) {
#pragma unroll
        for (uint32_t q = 0; q < SKIP_UNROLL; q++) {
            w0 = w0 + R.dx[0];
            w1 = w1 + R.dx[1];
            w2 = w2 + R.dx[2];
        }
    }
    for (; x < xs; x++) {
        w0 = w0 + R.dx[0];
        w1 = w1 + R.dx[1];
        w2 = w2 + R.dx[2];
    }
    const uint32_t pi0 = r * (g.TW + 1u) + (xs - g.X0);
    // the long-row variant (big-triangle kernels) walks two pixels per turn
    if (mode == 1u) walk_span<1u, (SKIP_UNROLL > 4u)>(p, R, prim, xs, xe, pi0, w0, w1, w2, key, d0s);
    else if (mode == 2u) walk_span<2u, (SKIP_UNROLL > 4u)>(p, R, prim, xs, xe, pi0, w0, w1, w2, key, d0s);
    else walk_row<0u>(p, R, prim, xs, xe, pi0, w0, w1, w2, key, d0s);
}

constexpr uint32_t kMaxTileH = 32;
static_assert(kBatch <= 256u, "a batch is staged one entry per lane");

// Setup of one bin entry of the current batch, kept in LDS (24 words).
struct TriLds {
    float o[3], dx[3], dy[3], hz[3], vby[6];
    uint32_t xmin, xmax, ymin, ymax, flags, prim;
};

struct TileShared {
    // Per pixel of the tile, rows padded by one entry (pixel i at i + i / TW):
    // lanes walking the same column of different rows hit different banks.
    uint64_t key[256 + kMaxTileH];  // winner key
    float d0s[256 + kMaxTileH];     // depth before the pass (the test's reference)
    TriLds tri[kBatch];             // the batch's triangles
    uint32_t rows_incl[256];        // inclusive prefix sum of their row counts in the tile
    uint8_t owner[256];             // entry of each (entry, row) item of the current 256
    uint32_t wave_rows[4];          // row counts per wave (prefix-sum carry)
    uint32_t band[2 * kMaxTileH];   // [by0, by1) of the render_par band of each row
    uint32_t next;                  // tile taken from the dynamic queue
    uint32_t next_tile, next_count, next_head;  // batches: that tile and its bin state
};

// Rows of bin entry `prim` inside the tile, its setup kept in sh.tri[slot].
EUCHIP_DEV uint32_t stage_entry(const FrameParams &p, const View &v, TileShared &sh, uint32_t slot, uint32_t prim,
                                const TileGeom &g) {
    TriRec R;
    load_rec0(p, v.tri_base + prim, R);
    const uint32_t ry0 = R.ymin > g.Y0 ? R.ymin : g.Y0;
    const uint32_t ry1 = R.ymax < g.Y0 + g.TH ? R.ymax : g.Y0 + g.TH;
    const uint32_t xs = R.xmin > g.X0 ? R.xmin : g.X0;
    const uint32_t xe = R.xmax < g.X0 + g.TW ? R.xmax : g.X0 + g.TW;
    if (ry0 >= ry1 || xs >= xe) return 0u;
    if (!all_rows_boxed(R)) {
        load_rec1(p, v.tri_base + prim, R);
    } else {
#pragma unroll
        for (int i = 0; i < 6; i++) R.vby[i] = 0.0f;  // boxed rows never read verts_by_y
    }
    TriLds &T = sh.tri[slot];
#pragma unroll
    for (int i = 0; i < 3; i++) T.o[i] = R.o[i], T.dx[i] = R.dx[i], T.dy[i] = R.dy[i], T.hz[i] = R.hz[i];
#pragma unroll
    for (int i = 0; i < 6; i++) T.vby[i] = R.vby[i];
    T.xmin = R.xmin, T.xmax = R.xmax, T.ymin = R.ymin, T.ymax = R.ymax, T.flags = R.flags, T.prim = prim;
    return ry1 - ry0;
}

// Inclusive prefix sum over the workgroup (4 waves); every thread must call it.
EUCHIP_DEV uint32_t block_inclusive_sum(TileShared &sh, uint32_t tid, uint32_t v) {
    const uint32_t incl = wave_inclusive_sum(v);
    if ((tid & 63u) == 63u) sh.wave_rows[tid >> 6] = incl;
    __syncthreads();
    uint32_t carry = 0;
    for (uint32_t w = 0; w < (tid >> 6); w++) carry += sh.wave_rows[w];
    return incl + carry;
}

// One batch of up to 256 bin entries: each lane stages its entry's setup in LDS,
// then the (entry, row) pairs of the batch are spread over all lanes so no
// lane walks much more than one row.
template <bool OWNER, bool LONG_ROWS>
EUCHIP_DEV void raster_batch(const FrameParams &p, const View &v, TileShared &sh, uint32_t tid, bool valid,
                             uint32_t prim, const TileGeom &g, uint32_t mode) {
    const uint32_t rows = valid ? stage_entry(p, v, sh, tid, prim, g) : 0u;
    const uint32_t incl = block_inclusive_sum(sh, tid, rows);
    const uint32_t excl = incl - rows;
    sh.rows_incl[tid] = incl;
    // items [excl, incl) are this entry's rows; the owner table maps the first
    // 256 items back to their entries with one LDS read, later ones (batches
    // past 256 rows) search the prefix sums
    if constexpr (OWNER)
        for (uint32_t k = excl; k < incl && k < 256u; k++) sh.owner[k] = (uint8_t)tid;
    __syncthreads();
    const uint32_t total = sh.rows_incl[255];  // lanes past the batch add nothing
    for (uint32_t k = tid; k < total; k += 256u) {
        uint32_t e;
        if (OWNER && k < 256u) {
            e = sh.owner[k];
        } else {
            uint32_t lo = 0, hi = kBatch - 1u;  // first entry whose inclusive sum exceeds k
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sh.rows_incl[mid] > k) hi = mid;
                else lo = mid + 1u;
            }
            e = lo;
        }
        const uint32_t base = e ? sh.rows_incl[e - 1u] : 0u;
        const TriLds &T = sh.tri[e];
        TriRec R;
#pragma unroll
        for (int i = 0; i < 3; i++) R.o[i] = T.o[i], R.dx[i] = T.dx[i], R.dy[i] = T.dy[i], R.hz[i] = T.hz[i];
#pragma unroll
        for (int i = 0; i < 6; i++) R.vby[i] = T.vby[i];
        R.xmin = T.xmin, R.xmax = T.xmax, R.ymin = T.ymin, R.ymax = T.ymax, R.flags = T.flags;
        const uint32_t ry0 = R.ymin > g.Y0 ? R.ymin : g.Y0;
        raster_row<LONG_ROWS ? 8u : 4u>(p, R, T.prim, ry0 + (k - base), g, mode, sh.key, sh.d0s, sh.band);
    }
    __syncthreads();  // the next batch reuses sh.tri and sh.rows_incl
}

// The big triangles of `tile` from their row checkpoints: the tile's list of
// (triangle, rows whose span reaches into the tile), 256 entries at a time
// (one per lane, coalesced; the first 256 issued when the tile starts).  An
// entry whose key-order bound lies above `occl`, the least bound of the tiles'
// full-cover triangles, has no rows here (it loses every pixel).  A
// prefix sum over the row counts places each entry's rows in a task table
// (entry, row), and each lane takes table entries and walks one row from its
// checkpoint (weights and the span inside the tile's column, at the slot of
// triangle, column and row), with the triangle's dx, hz, flags and primitive
// from its big record (two loads beside the checkpoint's).
struct BigTaskLds {
    uint32_t tri[256];        // triangle of each entry of the round
    uint16_t task[256 * 8];   // entry << 3 | row of the tile
};
static_assert(sizeof(BigTaskLds) <= sizeof(TriLds) * kBatch, "big-triangle tasks fit the staging space");

EUCHIP_DEV void big_tasks(const FrameParams &p, TileShared &sh, uint32_t tid, uint32_t tile, uint32_t nbig,
                          uint32_t occl, const TileGeom &g, uint32_t mode, uint2 e_first) {
    BigTaskLds &bt = *reinterpret_cast<BigTaskLds *>(sh.tri);
    const uint2 *list = p.cktab + (size_t)tile * p.mstride;
    const uint4 *cktile = p.ck + (size_t)(g.X0 >> g.sx) * p.v0.H + g.Y0;  // triangle t's at + t * ntx * H
    const size_t ckstride = (size_t)p.v0.ntx * p.v0.H;
    for (uint32_t s0 = 0; s0 < nbig; s0 += 256u) {
        const uint2 e2 = s0 ? (s0 + tid < nbig ? list[s0 + tid] : make_uint2(0u, 0u)) : e_first;
        const uint32_t e = e2.x;
        // behind the tile's occluder at every pixel: no rows
        const uint32_t m = s0 + tid < nbig && e2.y <= occl ? e & 0xffu : 0u;
        const uint32_t mine = __builtin_popcount(m);
        uint32_t at = block_inclusive_sum(sh, tid, mine) - mine;
        const uint32_t total = sh.wave_rows[0] + sh.wave_rows[1] + sh.wave_rows[2] + sh.wave_rows[3];
        bt.tri[tid] = e >> 8;
        for (uint32_t mm = m; mm; mm &= mm - 1u) bt.task[at++] = (uint16_t)((tid << 3) | (uint32_t)__builtin_ctz(mm));
        __syncthreads();
        // two rows per lane per turn, both rows' loads issued before either walk
        for (uint32_t k0 = 0; k0 < total; k0 += 512u) {
            uint4 c[2], x6[2], x7[2];
            uint32_t r[2];
#pragma unroll
            for (uint32_t h = 0; h < 2u; h++) {
                const uint32_t k = k0 + h * 256u + tid;
                r[h] = 8u;  // no row
                if (k >= total) continue;
                const uint32_t tk = bt.task[k], tri = bt.tri[tk >> 3];
                r[h] = tk & 7u;
                const uint4 *q = p.big_rec + (size_t)tri * kBigRecQuads;
                c[h] = cktile[tri * ckstride + r[h]];
                x6[h] = q[0], x7[h] = q[1];
            }
#pragma unroll
            for (uint32_t h = 0; h < 2u; h++) {
                if (r[h] >= 8u) continue;
                const uint32_t xs = g.X0 + (c[h].w & 0xffffu), xe = g.X0 + (c[h].w >> 16);
                TriRec R;
                R.dx[0] = __uint_as_float(x6[h].z), R.dx[1] = __uint_as_float(x6[h].w);
                R.dx[2] = __uint_as_float(x7[h].x);
                R.hz[0] = __uint_as_float(x7[h].y), R.hz[1] = __uint_as_float(x7[h].z);
                R.hz[2] = __uint_as_float(x7[h].w);
                R.flags = x6[h].y >> 30;
                const uint32_t pi0 = r[h] * (g.TW + 1u) + (xs - g.X0);
                const float w0 = __uint_as_float(c[h].x), w1 = __uint_as_float(c[h].y);
                const float w2 = __uint_as_float(c[h].z);
                if (mode == 1u) walk_span<1u, true>(p, R, x6[h].x, xs, xe, pi0, w0, w1, w2, sh.key, sh.d0s);
                else if (mode == 2u) walk_span<2u, true>(p, R, x6[h].x, xs, xe, pi0, w0, w1, w2, sh.key, sh.d0s);
                else walk_span<0u>(p, R, x6[h].x, xs, xe, pi0, w0, w1, w2, sh.key, sh.d0s);
            }
        }
        __syncthreads();  // the next round reuses the table and the wave totals
    }
}

// Waves per SIMD the raster pass is compiled for (measured per shader): latency
// hiding pays for a few spilled registers, up to a point.  Without coarse
// shading every one-request and batched kernel fits these without spills.
// The big-triangle kernels hold two row tasks per lane and take 6 (their
// frames have few tiles, so fewer waves per SIMD than that are resident anyway).
constexpr uint32_t raster_waves(uint32_t sh, bool ms, bool big) {
    return ms ? ((sh == PHONG_SHADOW || sh == TEX_SHADOW) ? 5u : sh == TEXTURED ? 7u : (sh == DEPTH_ONLY || sh == FLAT) ? 8u : 7u)
              : big ? 6u : ((sh == PHONG_SHADOW || sh == TEX_SHADOW) ? 7u : 8u);
}

template <uint32_t SH, bool BATCH, bool BIG, bool GS, bool MS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(raster_waves(SH, MS, BIG)))) void k_raster(FrameParams p) {
    __shared__ TileShared sh;
    const uint32_t tid = threadIdx.x;
    const uint32_t ntiles = p.ntiles;
    const uint32_t par = p.frame & 1u, npar = (p.frame + 1u) & 1u;
    const uint32_t xcd = blockIdx.x & 7u;
    const uint32_t per_xcd = gridDim.x >> 3, w = blockIdx.x >> 3;
    // The loads the first tile waits for, issued together: frame flags, pool use,
    // list lengths, this workgroup's first busy tile and its bin state (read at a
    // clamped index before the tile is known to be in the list; used only if it is).
    uint32_t lbase, lcap;
    busy_list_span(p, xcd, lbase, lcap);
    const uint32_t *list = p.busy + lbase;  // this XCD's busy list
    const uint32_t flags = p.ctr->flags[par][0];
    const uint32_t pool_used0 = p.ctr->pool[par][0];
    uint32_t nh = __builtin_amdgcn_readfirstlane(p.ctr->heavy[par][0]);
    // big triangles (with checkpoints any triangle may be on any tile's list)
    uint32_t nbig = !BIG ? 0u : p.ck_on ? p.v0.n_tris : __builtin_amdgcn_readfirstlane(p.ctr->big[par][0]);
    uint32_t nbusy = __builtin_amdgcn_readfirstlane(p.ctr->busy[par][xcd][0]);
    const uint32_t first = w < lcap ? __builtin_amdgcn_readfirstlane(list[w]) : 0u;
    const TileBin first_bin = p.tile_count[first < ntiles ? first : 0u];
    if (p.report && blockIdx.x == 0u && tid == 0u) {
        // this frame's setup counts, for the host (frame number last)
        Report *rp = p.report;
        rp->max_count = p.ctr->max_count[par][0];
        rp->flags = flags;
        rp->pool_used = pool_used0;
        if (flags & FLAG_BIN_OVERFLOW) rp->dropped = rp->dropped + 1u;
        __threadfence_system();
        rp->frame = p.frame;
    }
    // The next frame's counters start at zero (its setup counts into them).
    if (blockIdx.x == 0u && tid == 0u) {
        p.ctr->pool[npar][0] = 0u;
        p.ctr->flags[npar][0] = 0u;
        p.ctr->max_count[npar][0] = 0u;
        p.ctr->heavy[npar][0] = 0u;
        p.ctr->big[npar][0] = 0u;
    }
    if (blockIdx.x < 8u && tid == 0u) {
        p.ctr->next[npar][xcd][0] = 0u;
        p.ctr->empty[npar][xcd][0] = 0u;
        p.ctr->busy[npar][xcd][0] = 0u;
    }
    for (uint32_t i = blockIdx.x * 256u + tid; i < p.zero_n; i += gridDim.x * 256u) p.count_next[i] = TileBin{0u, 0u};
    if constexpr (BIG)  // after a checkpoint frame: its list counts
        for (uint32_t i = blockIdx.x * 256u + tid; i < p.ck_zero_n; i += gridDim.x * 256u) p.ck_count_next[i] = make_uint2(0u, 0u);
    // The pool ran out in setup: entries were lost, the frame is dropped (the
    // host sees the flag and re-runs it with a larger pool).
    if (flags & FLAG_BIN_OVERFLOW) return;
    const uint32_t pool_used = pool_used0 < p.pool_cap ? pool_used0 : p.pool_cap;
    const bool uses_depth = p.depth_test != DT_NONE || p.depth_write;
    const uint32_t mode = (p.depth_write && p.depth_test == DT_LESS) ? 1u
                        : (p.depth_write && p.depth_test == DT_GREATER) ? 2u : 0u;
    const uint32_t sx = p.tile_sx, sy = 8u - sx;
    uint32_t *next = &p.ctr->next[par][xcd][0];
    // big triangles from their row checkpoints (else walked from the row starts)
    const bool ckp = BIG && p.ck_on;

    // One tile with `cnt` bin entries and overflow list `head`: winner keys in
    // LDS, then the resolve.
    // With `claim`, thread 0 takes the next queue index when the keys are done
    // (the atomic's latency hides behind the resolve) and leaves it in sh.next.
    // Ends with a barrier (LDS is reused by the next tile).
    auto run_tile = [&](uint32_t tile, uint32_t cnt, uint32_t head, bool claim) {
        // Batches (hundreds of tiles per workgroup) take the next queue index
        // when the tile starts and load that tile's list entry and bin state
        // before the resolve, so the next tile starts with no load to wait for.
        uint32_t claimed = 0u;
        if (BATCH && claim && tid == 0u) claimed = per_xcd + atomicAdd(next, 1u);
        uint32_t r = 0;
        if constexpr (BATCH) r = __builtin_amdgcn_readfirstlane(p.tile_req[tile]);
        const View &v = view_of<BATCH>(p, r);
        const uint32_t lt = tile - v.tile_base;
        const uint32_t ty = lt / v.ntx, tx = lt - ty * v.ntx;
        TileGeom g;
        g.sx = sx, g.TW = 1u << sx, g.TH = 256u >> sx;
        g.X0 = tx << sx, g.Y0 = ty << sy;
        const uint32_t px = g.X0 + (tid & (g.TW - 1u)), py = g.Y0 + (tid >> sx);
        const bool inside = px < v.W && py < v.H;
        const size_t pidx = (size_t)py * v.W + px;
        float D0 = 0.0f;
        if (inside && uses_depth) D0 = v.clear_depth_en ? v.clear_depth : v.depth[pidx];
        const uint32_t lp = tid + (tid >> sx);  // padded LDS index of this lane's pixel
        sh.key[lp] = init_key(mode, D0);
        sh.d0s[lp] = D0;
        if (tid < g.TH) {
            uint32_t by0, by1;
            band_of(p, v, g.Y0 + tid, by0, by1);
            sh.band[2 * tid] = by0;
            sh.band[2 * tid + 1] = by1;
        }
        const uint32_t cap = v.tile_cap;
        const uint32_t *bin = p.bins + v.bin_base + (size_t)lt * cap;
        // the tile's checkpoint list: its length and first 256 entries, loaded
        // now (used after the bin; entries past the length are ignored)
        uint32_t nck = 0u, occl = 0u;
        uint2 e_first = make_uint2(0u, 0u);
        if (BIG && ckp) {
            const uint2 cc = p.ck_count[tile];
            nck = cc.x, occl = ~cc.y;
            if (tid < p.mstride) e_first = p.cktab[(size_t)tile * p.mstride + tid];
        }
        __syncthreads();
        // Runs of entries, kBatch at a time: this tile's bin and the big
        // triangles (staging drops those that miss it), then (only when it
        // overflowed) its overflow segments.  Segments were taken from the pool
        // in this frame, each pointing at the one taken before it; a walk of at
        // most pool_used / 3 segments whose every read stays below pool_used
        // ends even on a damaged list.  One loop, so one copy of raster_batch.
        const uint32_t nb = cnt < cap ? cnt : cap;
        uint32_t run = nb + (ckp ? 0u : nbig);  // entries of the current run
        uint32_t base = 0u;                 // pool index of the current segment's entries (0: bin + big)
        uint32_t seg = cnt > cap ? head : 0u, guard = 0u;
        for (uint32_t b = 0;; b += kBatch) {
            while (b >= run) {              // next segment (uniform)
                if (!seg || guard++ > pool_used / 3u || seg + 1u > pool_used) {
                    seg = 0u;
                    break;
                }
                const uint32_t s0 = seg - 1u;
                seg = __builtin_amdgcn_readfirstlane(p.pool[s0]);
                const uint32_t len = __builtin_amdgcn_readfirstlane(p.pool[s0 + 1u]);
                base = s0 + 2u;
                run = len < pool_used - base ? len : pool_used - base;
                b = 0u;
            }
            if (b >= run) break;
            const uint32_t i = b + tid;
            bool valid = false;
            uint32_t prim = 0;
            if (tid >= kBatch || i >= run) {
            } else if (base) {
                valid = true, prim = p.pool[base + i];
            } else if (i < nb) {
                valid = true, prim = bin[i];
            } else if constexpr (BIG) {
                valid = true, prim = p.big[i - nb];
            }
            raster_batch<!BATCH, BIG>(p, v, sh, tid, valid, prim, g, mode);
        }
        if (BIG && ckp) {
            nck = __builtin_amdgcn_readfirstlane(nck);
            if (nck) big_tasks(p, sh, tid, tile, nck, __builtin_amdgcn_readfirstlane(occl), g, mode, e_first);
        }
        uint32_t ntile = 0u;
        TileBin ntb{0u, 0u};
        if (BATCH && claim && tid == 0u && claimed < nbusy) {
            ntile = list[claimed];
            ntb = p.tile_count[ntile];
        }
        if (!BATCH && claim && tid == 0u) claimed = per_xcd + atomicAdd(next, 1u);
        bool shade = false;  // this lane's pixel has a winner to shade
        if (inside) {
            const uint64_t kk = sh.key[lp];
            bool has;  // a winner (its primitive is read again by the shading lanes)
            float z;
            if (mode == 0u) {
                has = kk != 0ull;
                z = __uint_as_float((uint32_t)kk);
            } else {
                has = ((uint32_t)kk) != 0u;
                const uint32_t o = (uint32_t)(kk >> 32);
                z = zunord(mode == 1u ? o : ~o, (uint32_t)kk & 1u);
            }
            shade = v.pixel && p.pixel_write && has && SH != DEPTH_ONLY;
            if (v.pixel && !shade && v.clear_color_en) store_stream(&v.pixel[pidx], v.clear_color);
            if (v.depth) {
                if (has && p.depth_write) store_stream(&v.depth[pidx], z);
                else if (v.clear_depth_en) store_stream(&v.depth[pidx], v.clear_depth);
            }
        }
        // The pixels to shade, packed onto the first lanes in tile order: a wave
        // with none of them skips the shader.
        if constexpr (SH != DEPTH_ONLY) {
            const uint64_t bal = __ballot(shade);
            if ((tid & 63u) == 0u) sh.wave_rows[tid >> 6] = (uint32_t)__popcll(bal);
            __syncthreads();
            uint32_t at = 0, total = 0;
#pragma unroll
            for (uint32_t w4 = 0; w4 < 4u; w4++) {
                const uint32_t c = sh.wave_rows[w4];
                at += w4 < (tid >> 6) ? c : 0u;
                total += c;
            }
            if (shade) sh.owner[at + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u))] = (uint8_t)tid;
            __syncthreads();
            if (tid < total) {
                const uint32_t t = sh.owner[tid];
                const uint32_t tx2 = g.X0 + (t & (g.TW - 1u)), ty2 = g.Y0 + (t >> sx);
                const uint64_t kk = sh.key[t + (t >> sx)];
                const uint32_t prim = mode == 0u ? (uint32_t)(kk >> 32) - 1u : ((uint32_t)kk >> 1) - 1u;
                TriRec R;
                load_rec0(p, v.tri_base + prim, R);
                tri_sids(v, p.geometry, prim, R.flags, R.sid);
                store_stream(&v.pixel[(size_t)ty2 * v.W + tx2],
                             shade_pixel<SH, GS, MS>(p, v, R, tx2, ty2, sh.band[2 * (t >> sx)],
                                                     v.clear_color_en ? v.clear_color : 0u));
            }
        }
        if (claim && tid == 0u) {
            sh.next = claimed;
            if constexpr (BATCH) sh.next_tile = ntile, sh.next_count = ntb.count, sh.next_head = ntb.head;
        }
        __syncthreads();
    };

    if (BIG && !p.ck_on) nbig = nbig < p.big_cap ? nbig : p.big_cap;
    // With big triangles every tile may hold fragments: the loop below then takes
    // every tile of this XCD's contiguous range instead of its busy list.
    if (BIG && nbig) {
        uint32_t b0, b1;
        xcd_range(ntiles, xcd, b0, b1);
        nbusy = b1 - b0;
    }
    nh = nh < p.heavy_cap ? nh : p.heavy_cap;
    nbusy = nbusy < lcap ? nbusy : lcap;
    // 1. Heavy tiles first, workgroup k taking heavy[k] (the longest tiles
    //    start early instead of ending the launch late).
    for (uint32_t k = blockIdx.x; k < nh; k += gridDim.x) {
        const uint32_t tile = __builtin_amdgcn_readfirstlane(p.heavy[k]);
        const TileBin tb = p.tile_count[tile];
        run_tile(tile, __builtin_amdgcn_readfirstlane(tb.count), __builtin_amdgcn_readfirstlane(tb.head), false);
    }
    // 2. Busy tiles of this XCD's list: the first by workgroup index, the rest
    //    from the XCD's counter (heavy ones are done).
    for (uint32_t k = w; k < nbusy;) {
        uint32_t tile;
        if (BIG && nbig) {
            uint32_t b0, b1;
            xcd_range(ntiles, xcd, b0, b1);
            tile = b0 + k;
        } else if (BATCH && k != w) {
            tile = __builtin_amdgcn_readfirstlane(sh.next_tile);  // loaded during the previous tile
        } else {
            tile = k == w ? first : __builtin_amdgcn_readfirstlane(list[k]);
        }
        TileBin tb = k == w && !nbig ? first_bin : TileBin{0u, 0u};
        if (BATCH && k != w) tb = TileBin{sh.next_count, sh.next_head};
        else if (k != w || nbig) tb = p.tile_count[tile];
        const uint32_t cnt = __builtin_amdgcn_readfirstlane(tb.count);
        if (!(p.heavy_cap && cnt >= kHeavyTile)) {
            run_tile(tile, cnt, __builtin_amdgcn_readfirstlane(tb.head), true);
        } else {
            if (tid == 0u) sh.next = per_xcd + atomicAdd(next, 1u);
            __syncthreads();
        }
        k = __builtin_amdgcn_readfirstlane(sh.next);
        __syncthreads();
    }
    // 3. Empty tiles: groups of 8 consecutive tiles of the XCD's contiguous
    //    range taken from its counter (the workgroups that finish their busy
    //    tiles first take more of them), the clears, stores only.
    if ((BIG && nbig) || (!BATCH && !p.v0.clear_color_en && !p.v0.clear_depth_en)) return;
    uint32_t t0, t1;
    xcd_range(ntiles, xcd, t0, t1);
    uint32_t *empty = &p.ctr->empty[par][xcd][0];
    for (;;) {
        if (tid == 0u) sh.next = atomicAdd(empty, 8u);
        __syncthreads();
        const uint32_t base = t0 + __builtin_amdgcn_readfirstlane(sh.next);
        __syncthreads();  // sh.next is written again next turn
        if (base >= t1) break;
        uint32_t c[8];
#pragma unroll
        for (uint32_t j = 0; j < 8u; j++) {
            const uint32_t t = base + j;
            c[j] = t < t1 ? p.tile_count[t].count : 1u;
        }
#pragma unroll
        for (uint32_t j = 0; j < 8u; j++) {
            if (c[j] != 0u) continue;
            const uint32_t tile = base + j;
            uint32_t r = 0;
            if constexpr (BATCH) r = __builtin_amdgcn_readfirstlane(p.tile_req[tile]);
            const View &v = view_of<BATCH>(p, r);
            const uint32_t lt = tile - v.tile_base;
            // one-sample kernels: the divisor made opaque here (an empty asm),
            // so that its reciprocal is computed in this loop and not kept live
            // through the whole kernel (where it took a register the shading needs)
            uint32_t ntx = v.ntx;
            if constexpr (!MS && !BIG) {
                ntx = __builtin_amdgcn_readfirstlane(ntx);
                asm volatile("" : "+s"(ntx));
            }
            const uint32_t ty = lt / ntx, tx = lt - ty * ntx;
            const uint32_t px = (tx << sx) + (tid & ((1u << sx) - 1u)), py = (ty << sy) + (tid >> sx);
            if (px < v.W && py < v.H) {
                const size_t pidx = (size_t)py * v.W + px;
                if (v.pixel && v.clear_color_en) store_stream(&v.pixel[pidx], v.clear_color);
                if (v.depth && v.clear_depth_en) store_stream(&v.depth[pidx], v.clear_depth);
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *dst, uint32_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) dst[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline uint32_t blocks_of(size_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

template <uint32_t SH>
static void launch_setup_t(const FrameParams &p, hipStream_t s, bool batched, uint32_t blocks) {
    if (batched) {
        hipLaunchKernelGGL((k_setup<SH, true, false>), dim3(blocks), dim3(256), 0, s, p);
    } else {
        // triangle blocks, then vertex blocks; checkpoint frames: blocks of
        // triangles and vertices, then span blocks
        FrameParams q = p;
        uint32_t span_blocks = 0;
        size_t lds = 0;
        if (p.ck_on && p.v0.H) {  // span blocks: 4 items of kSpanRows rows per triangle and block
            const uint32_t sy = 8u - p.tile_sx;
            q.span_parts = (p.v0.H + (1u << sy) + kSpanRows - 1u) / kSpanRows;
            span_blocks = p.v0.n_tris * ((q.span_parts + 3u) / 4u);
            const size_t tables = 4u * (size_t)span_table_bytes(sy, p.v0.ntx);  // in the static LDS when it fits
            lds = tables <= sizeof(SetupShared) ? 0u : tables;
        }
        const size_t nv = (size_t)p.v0.n_instances * p.v0.n_vertices;
        if (span_blocks) {
            q.tri_blocks = q.setup_blocks = blocks_of(std::max<size_t>(p.v0.n_tris, nv), 256);
        } else {
            q.tri_blocks = blocks_of(p.v0.n_tris, 256);
            q.setup_blocks = q.tri_blocks + (vd_count(SH) > 0 ? blocks_of(nv, 256) : 0u);
        }
        if (span_blocks)
            hipLaunchKernelGGL((k_setup<SH, false, true>), dim3(q.setup_blocks + span_blocks), dim3(256), lds, s, q);
        else
            hipLaunchKernelGGL((k_setup<SH, false, false>), dim3(q.setup_blocks), dim3(256), 0, s, q);
    }
}

hipError_t launch_setup(const FrameParams &p, hipStream_t s, bool batched, uint32_t blocks) {
    if (batched && blocks == 0) return hipSuccess;
    switch (p.shader) {
    case FLAT: launch_setup_t<FLAT>(p, s, batched, blocks); break;
    case COLOR: launch_setup_t<COLOR>(p, s, batched, blocks); break;
    case GOURAUD: launch_setup_t<GOURAUD>(p, s, batched, blocks); break;
    case PHONG: launch_setup_t<PHONG>(p, s, batched, blocks); break;
    case DEPTH_ONLY: launch_setup_t<DEPTH_ONLY>(p, s, batched, blocks); break;
    case TEXTURED: launch_setup_t<TEXTURED>(p, s, batched, blocks); break;
    case PHONG_SHADOW: launch_setup_t<PHONG_SHADOW>(p, s, batched, blocks); break;
    case TEX_SHADOW: launch_setup_t<TEX_SHADOW>(p, s, batched, blocks); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Workgroups of a raster kernel that fit on the device at once (`cus` compute
// units); the busy-tile loop gives every workgroup a first tile, so workgroups
// that could only start when others finish would hold theirs back.
template <uint32_t SH, bool BATCH, bool BIG, bool GS, bool MS>
static uint32_t resident_groups(uint32_t cus) {
    static int per_cu = 0;
    if (per_cu == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(&k_raster<SH, BATCH, BIG, GS, MS>),
                                                         256, 0) != hipSuccess ||
            n < 1)
            n = 8;
        per_cu = n;
    }
    return (uint32_t)per_cu * cus;
}

template <uint32_t SH, bool GS, bool MS>
static void launch_raster_m(const FrameParams &p, hipStream_t s, bool batched, uint32_t groups, uint32_t cus) {
    if (batched) {
        groups = std::min(groups, std::max<uint32_t>(8u, resident_groups<SH, true, false, GS, MS>(cus) / 8u * 8u));
        hipLaunchKernelGGL((k_raster<SH, true, false, GS, MS>), dim3(groups), dim3(256), 0, s, p);
    } else if (p.big_cap) {
        groups = std::min(groups, std::max<uint32_t>(8u, resident_groups<SH, false, true, GS, MS>(cus) / 8u * 8u));
        hipLaunchKernelGGL((k_raster<SH, false, true, GS, MS>), dim3(groups), dim3(256), 0, s, p);
    } else {
        groups = std::min(groups, std::max<uint32_t>(8u, resident_groups<SH, false, false, GS, MS>(cus) / 8u * 8u));
        hipLaunchKernelGGL((k_raster<SH, false, false, GS, MS>), dim3(groups), dim3(256), 0, s, p);
    }
}

template <uint32_t SH, bool GS>
static void launch_raster_g(const FrameParams &p, hipStream_t s, bool batched, uint32_t groups, uint32_t cus) {
    if (p.msaa_level) launch_raster_m<SH, GS, true>(p, s, batched, groups, cus);
    else launch_raster_m<SH, GS, false>(p, s, batched, groups, cus);
}

// Textured shaders whose sampler is not their own take the kernels that read
// textures through any sampler kind.
template <uint32_t SH>
static void launch_raster_t(const FrameParams &p, hipStream_t s, bool batched, uint32_t groups, uint32_t cus) {
    if constexpr (SH == TEXTURED || SH == TEX_SHADOW) {
        const uint32_t own = SH == TEX_SHADOW ? (uint32_t)SAMPLER_LINEAR_CLAMPED : (uint32_t)SAMPLER_LINEAR;
        if (p.sampler != own) {
            launch_raster_g<SH, true>(p, s, batched, groups, cus);
            return;
        }
    }
    launch_raster_g<SH, false>(p, s, batched, groups, cus);
}

// ---------------------------------------------------------------------------
// Span blocks of a checkpoint frame's setup launch: block b walks rows of
// triangle t = b % n_tris, its wave w the item 4 * (b / n_tris) + w (kSpanRows
// rows from the triangle's first tile row).  Wave 0 sets the triangle up
// (tri_setup, as the setup blocks do, which write its records and bin none of
// a checkpoint frame's triangles).  A lane takes a row: the row span
// (triangles.rs:219-254, with the row's band), a bit in the wave's LDS mask
// table for every tile column the span touches, then the walk from x0
// (triangles.rs:257-301, the same w += dx additions as the raster's walk)
// storing the weights where the span enters each column; a second table
// keeps a bit per row whose span holds the whole column.  The wave then
// appends the entries of its tiles to their lists, with their key-order
// bounds, and raises the occluder bound of each tile the triangle covers.
// Slots sit at fixed places (layout in internal.hpp), so no block waits for
// another.
// ---------------------------------------------------------------------------

// Key-order bounds of a big triangle over one tile, for the occlusion test
// (write modes).  Every weight the raster computes in the tile comes from the
// walk of its row from x0 (triangles.rs:257-301): the start value and each of
// at most (xmax - xmin) additions round by half an ulp of a partial sum, which
// M bounds (|o| + |dy| y + |dx| x over the triangle's box), so each weight
// stays within err = (xmax - xmin + 8) 2^-23 M of its plane, u2 within the sum
// of those and its two subtractions, and z (three products, two sums) within
// ez of the plane hz . (w0, w1, u2).  The planes are affine, so their extremes
// over the tile's pixels lie at its corners, evaluated in f32 with the
// evaluation's own rounding added to the margins (all terms doubled).
struct TileBounds {
    uint32_t klo;  // key order no fragment of the triangle in the tile goes below
    uint32_t khi;  // key order none goes above (valid where `full`)
    bool full;     // every pixel of the tile is a fragment of the triangle
};

EUCHIP_DEV TileBounds tile_bounds(const FrameParams &p, const TriRec &R, uint32_t mode, uint32_t X0, uint32_t Y0,
                                  uint32_t TW, uint32_t TH) {
    const float Xm = (float)R.xmax + 32.0f, Ym = (float)R.ymax;
    float M[3], e[3];
    const float K = (float)(R.xmax - R.xmin + 8u) * 0x1p-22f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        M[i] = (fabsf(R.o[i]) + fabsf(R.dy[i]) * Ym + fabsf(R.dx[i]) * Xm) * 1.001f;
        e[i] = K * M[i] + 0x1p-125f;
    }
    const float S = M[0] + M[1] + M[2];
    const float eu = (e[0] + e[1] + e[2]) + 0x1p-21f * S;
    const float h0 = fabsf(R.hz[0]), h1 = fabsf(R.hz[1]), h2 = fabsf(R.hz[2]);
    const float P = h0 * M[0] + h1 * M[1] + h2 * S;
    const float ez = 2.0f * ((h0 * e[0] + h1 * e[1] + h2 * eu) + 0x1p-18f * P) + 0x1p-124f;
    // the corners' weights and z (planes of the reference's setup values)
    const float xs[2] = {(float)X0, (float)(X0 + TW - 1u)}, ys[2] = {(float)Y0, (float)(Y0 + TH - 1u)};
    float zmin = INFINITY, zmax = -INFINITY, wmin[3] = {INFINITY, INFINITY, INFINITY};
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
            float w[3];
#pragma unroll
            for (int i = 0; i < 3; i++) w[i] = (R.o[i] + R.dy[i] * ys[c]) + R.dx[i] * xs[a];
            const float u2 = (w[2] - w[0]) - w[1];
            const float z = (R.hz[0] * w[0] + R.hz[1] * w[1]) + R.hz[2] * u2;
            zmin = fminf(zmin, z), zmax = fmaxf(zmax, z);
            wmin[0] = fminf(wmin[0], w[0]), wmin[1] = fminf(wmin[1], w[1]), wmin[2] = fminf(wmin[2], u2);
        }
    TileBounds b;
    const bool finite = (zmin - zmin == 0.0f) & (zmax - zmax == 0.0f) & (ez - ez == 0.0f);
    const float zlo = zmin - ez, zhi = zmax + ez;
    // order values, as the walk's keys form them (-0 and +0 alike)
    auto ord = [](float z) {
        const uint32_t u = __float_as_uint(z + 0.0f);
        return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
    };
    b.klo = !finite ? 0u : mode == 1u ? ord(zlo) : ~ord(zhi);
    b.khi = mode == 1u ? ord(zhi) : ~ord(zlo);
    const bool clip = !(R.flags & TRI_ALL_IN) && p.z_clip;
    b.full = finite & (wmin[0] > e[0] + 0x1p-21f * M[0]) & (wmin[1] > e[1] + 0x1p-21f * M[1]) &
             (wmin[2] > eu + 0x1p-20f * S) & (!clip | ((p.z_min <= zlo) & (zhi <= p.z_max)));
    return b;
}

EUCHIP_DEV void span_item(const FrameParams &p, uint32_t b, void *static_lds) {
    extern __shared__ uint32_t span_lds[];  // the span tables when the static LDS is too small
    __shared__ TriRec span_tri;              // the block's triangle, set up once by wave 0
    __shared__ uint32_t span_keep;
    const View &v = p.v0;
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t t = b % v.n_tris, part = (b / v.n_tris) * 4u + wv;
    if (wv == 0u) {
        TriRec R0;
        const bool keep = tri_setup(p, v, t, R0);
        if (lane == 0u) span_tri = R0, span_keep = keep ? 1u : 0u;
    }
    __syncthreads();
    if (!span_keep || part >= p.span_parts) return;
    const TriRec R = span_tri;
    const uint32_t sx = p.tile_sx, sy = 8u - sx, TH = 1u << sy, TW = 1u << sx;
    const uint32_t col0 = R.xmin >> sx;
    const uint32_t cols = ((R.xmax - 1u) >> sx) - col0 + 1u;
    if (part == 0u && lane == 0u) {  // what the raster's row tasks read besides the slot
        uint4 *dst = p.big_rec + (size_t)t * kBigRecQuads;
        dst[0] = make_uint4(t, R.flags << 30, __float_as_uint(R.dx[0]), __float_as_uint(R.dx[1]));
        dst[1] = make_uint4(__float_as_uint(R.dx[2]), __float_as_uint(R.hz[0]), __float_as_uint(R.hz[1]),
                            __float_as_uint(R.hz[2]));
    }
    const uint32_t ybase = ((R.ymin >> sy) << sy) + part * kSpanRows;
    if (ybase >= R.ymax) return;
    const uint32_t groups = kSpanRows >> sy, ne = groups * cols;
    // this wave's tables of (tile row, column) bytes: rows reaching into the
    // tile, rows whose span holds all of its columns
    const uint32_t tb = span_table_bytes(sy, v.ntx);
    uint8_t *base = 4u * tb <= sizeof(SetupShared) ? (uint8_t *)static_lds : (uint8_t *)span_lds;
    uint8_t *const tab = base + wv * tb;
    const uint32_t nb = (ne + 3u) & ~3u;
    uint32_t *mw = (uint32_t *)tab, *fw = (uint32_t *)(tab + nb);
    const uint8_t *mb = tab, *fb = tab + nb;
    for (uint32_t i = lane; i < (nb >> 2); i += 64u) mw[i] = 0u, fw[i] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t y = ybase + lane;
    uint32_t x0 = 0, x1 = 0;
    if (y >= R.ymin && y < R.ymax) {
        uint32_t by0, by1;
        band_of(p, v, y, by0, by1);
        row_span(R, y, by0, by1, x0, x1);
    }
    if (x0 < x1) {
        const uint32_t at0 = (lane >> sy) * cols - col0, bit = 1u << (y & (TH - 1u));
        for (uint32_t c = x0 >> sx; c <= ((x1 - 1u) >> sx); c++) {
            const uint32_t at = at0 + c, sh = (at & 3u) * 8u;
            atomicOr(&mw[at >> 2], bit << sh);
            if (x0 <= (c << sx) && (c << sx) + TW <= x1) atomicOr(&fw[at >> 2], bit << sh);
        }
        const float Y = (float)y, X0 = (float)x0;
        float w0 = (R.o[0] + R.dy[0] * Y) + R.dx[0] * X0;
        float w1 = (R.o[1] + R.dy[1] * Y) + R.dx[1] * X0;
        float w2 = (R.o[2] + R.dy[2] * Y) + R.dx[2] * X0;
        uint4 *row = p.ck + (size_t)t * v.ntx * v.H + y;  // column c's slot at row + c * H
        for (uint32_t x = x0;;) {
            const uint32_t c = x >> sx, xc = c << sx;
            const uint32_t xe = xc + TW < x1 ? xc + TW : x1;
            row[(size_t)c * v.H] = make_uint4(__float_as_uint(w0), __float_as_uint(w1), __float_as_uint(w2),
                                              (x - xc) | ((xe - xc) << 16));
            if (xe >= x1) break;
            for (; x + 16u <= xe; x += 16u) {
#pragma unroll
                for (uint32_t q = 0; q < 16u; q++) {
                    w0 = w0 + R.dx[0];
                    w1 = w1 + R.dx[1];
                    w2 = w2 + R.dx[2];
                }
            }
            for (; x < xe; x++) {
                w0 = w0 + R.dx[0];
                w1 = w1 + R.dx[1];
                w2 = w2 + R.dx[2];
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // the entries of this item's tiles (every tile row of it is this item's)
    // with their bounds; a tile whose every row's span holds it and whose
    // pixels the triangle covers (tile_bounds) raises its occluder bound.
    // Key order: Less + write takes zord, Greater + write ~zord; other depth
    // modes keep every entry (klo 0) and record no occluder.
    const uint32_t mode = (p.depth_write && p.depth_test == DT_LESS) ? 1u
                        : (p.depth_write && p.depth_test == DT_GREATER) ? 2u : 0u;
    const uint32_t ty0 = ybase >> sy, all_rows = (1u << TH) - 1u;
    for (uint32_t i = lane; i < ne; i += 64u) {
        const uint32_t m = mb[i];
        if (m == 0u) continue;
        const uint32_t gi = i / cols, c = i - gi * cols;
        const size_t tile = (size_t)(ty0 + gi) * v.ntx + col0 + c;
        TileBounds tb{0u, 0u, false};
        if (mode != 0u) tb = tile_bounds(p, R, mode, (col0 + c) << sx, (ty0 + gi) << sy, TW, TH);
        const uint32_t k = atomicAdd(&p.ck_count[tile].x, 1u);  // below mstride: one entry per triangle
        p.cktab[tile * p.mstride + k] = make_uint2((t << 8) | m, tb.klo);
        if (tb.full && fb[i] == all_rows) atomicMax(&p.ck_count[tile].y, ~tb.khi);
    }
}

// Workgroups of the tile loop: a multiple of the 8 XCDs, no more than the
// tiles need (rounded up to 8) and no more than `max_groups` (compute units x 8).
hipError_t launch_raster(const FrameParams &p, hipStream_t s, bool batched, uint32_t max_groups) {
    if (p.ntiles == 0) return hipSuccess;
    uint32_t g = ((p.ntiles + 7u) / 8u) * 8u;
    const uint32_t cap = std::max<uint32_t>(8u, (max_groups / 8u) * 8u);
    if (g > cap) g = cap;
    const uint32_t cus = std::max<uint32_t>(1u, max_groups / 8u);
    switch (p.shader) {
    case FLAT: launch_raster_t<FLAT>(p, s, batched, g, cus); break;
    case COLOR: launch_raster_t<COLOR>(p, s, batched, g, cus); break;
    case GOURAUD: launch_raster_t<GOURAUD>(p, s, batched, g, cus); break;
    case PHONG: launch_raster_t<PHONG>(p, s, batched, g, cus); break;
    case DEPTH_ONLY: launch_raster_t<DEPTH_ONLY>(p, s, batched, g, cus); break;
    case TEXTURED: launch_raster_t<TEXTURED>(p, s, batched, g, cus); break;
    case PHONG_SHADOW: launch_raster_t<PHONG_SHADOW>(p, s, batched, g, cus); break;
    case TEX_SHADOW: launch_raster_t<TEX_SHADOW>(p, s, batched, g, cus); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}


// ---------------------------------------------------------------------------
// lines (lines.rs:15-121): one-request frames of LineList / LineTriangleList
// ---------------------------------------------------------------------------

// Clip position and shaded-vertex id of stream element e of request v.
EUCHIP_DEV void element_clip(const FrameParams &p, const View &v, uint32_t e, float c[4], uint32_t &sid) {
    const uint32_t inst = e / v.count;
    uint32_t vi = e - inst * v.count;
    if (v.idx) vi = v.idx[vi];
    const float *P = v.pos + (size_t)v.pos_comps * vi;
    if (p.xform == 0u) {
        c[0] = P[0], c[1] = P[1], c[2] = P[2], c[3] = v.pos_comps == 4u ? P[3] : 1.0f;
    } else {
        const float p4[4] = {P[0], P[1], P[2], 1.0f};
        float w[4];
        mat4_vec4(v.models + 16u * inst, p4, w);
        mat4_vec4(v.u + U_VP, w, c);
    }
    sid = v.sv_base + inst * v.n_vertices + vi;
}

// `f as isize` clamped to what the record keeps: NaN -> 0, otherwise truncation,
// saturated (coordinates past kLineCoordLimit mark the line undrawn).
EUCHIP_DEV int32_t line_coord(float f, bool &far) {
    if (f != f) return 0;
    if (!(f > -(float)kLineCoordLimit && f < (float)kLineCoordLimit)) {
        far = true;
        return 0;
    }
    return (int32_t)f;
}

// f32::clamp then `as isize` for a bound within [lo, hi] (lo, hi integral, >= 0).
EUCHIP_DEV int32_t clamp_coord(float f, float lo, float hi) {
    if (f < lo) f = lo;
    if (f > hi) f = hi;
    if (f != f) return 0;
    return (int32_t)f;
}

// k_line_setup: the vertex stage (lane t = shaded vertex t) and one line record
// per lane (lane t = line t): lines.rs:41-86 in f32 over the whole target (the
// band clamp of the clip rectangle only cuts rows other bands draw).
template <uint32_t SH>
__global__ __launch_bounds__(256) void k_line_setup(FrameParams p) {
    const View &v = p.v0;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if constexpr (vd_count(SH) > 0) {
        if (t < v.n_instances * v.n_vertices) {
            const uint32_t inst = t / v.n_vertices;
            float vd[VDN<SH>::N];
            vertex_data<SH>(p, v, inst, t - inst * v.n_vertices, vd);
#pragma unroll
            for (uint32_t q = 0; q < vd_count(SH); q++) p.vary[((size_t)v.sv_base + t) * vd_count(SH) + q] = vd[q];
        }
    }
    if (t >= p.n_lines) return;
    uint32_t e0, e1;
    if (p.primitives == PRIMS_LINE_LIST) {
        e0 = 2u * t, e1 = 2u * t + 1u;
    } else {  // edges a-b, b-c, c-a of stage-output triangle t / 3 (primitives.rs:71-84)
        const uint32_t tri = t / 3u, j = t - 3u * tri;
        e0 = corner_element(p.geometry, tri, (int)j);
        e1 = corner_element(p.geometry, tri, j == 2u ? 0 : (int)j + 1);
    }
    float hom[2][4];
    uint32_t sid[2];
    element_clip(p, v, e0, hom[0], sid[0]);
    element_clip(p, v, e1, hom[1], sid[1]);
    const float flip1 = p.y_up ? -1.0f : 1.0f;
    const float size0 = (float)v.W, size1 = (float)v.H;
    float ez[2], vs[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const float x = hom[i][0] * 1.0f, y = hom[i][1] * flip1, z = hom[i][2];
        const float w = rmax(hom[i][3], 0.0001f);
        const float ex = x / w, ey = y / w;
        ez[i] = z / w;
        vs[i][0] = size0 * (ex * 0.5f + 0.5f);
        vs[i][1] = size1 * (ey * -0.5f + 0.5f);
    }
    bool far = false;
    const int32_t x1 = line_coord(vs[0][0], far), y1 = line_coord(vs[0][1], far);
    const int32_t x2 = line_coord(vs[1][0], far), y2 = line_coord(vs[1][1], far);
    const int32_t wx1 = clamp_coord(rmin(vs[0][0], vs[1][0]) + 0.0f, 0.0f, size0);
    const int32_t wy1 = clamp_coord(rmin(vs[0][1], vs[1][1]) + 0.0f, 0.0f, size1);
    const int32_t wx2 = clamp_coord(rmax(vs[0][0], vs[1][0]) + 1.0f, 0.0f, size0);
    const int32_t wy2 = clamp_coord(rmax(vs[0][1], vs[1][1]) + 1.0f, 0.0f, size1);
    const int32_t adx = x1 > x2 ? x1 - x2 : x2 - x1, ady = y1 > y2 ? y1 - y2 : y2 - y1;
    const bool use_x = adx > ady;
    const float norm = 1.0f / (use_x ? vs[1][0] - vs[0][0] : vs[1][1] - vs[0][1]);
    uint32_t *r = p.lrec + (size_t)t * kLineRecWords;
    const uint4 q0 = make_uint4(__float_as_uint(vs[0][0]), __float_as_uint(vs[0][1]), __float_as_uint(norm),
                                __float_as_uint(ez[0]));
    const uint4 q1 = make_uint4(__float_as_uint(ez[1]), (uint32_t)x1, (uint32_t)y1, (uint32_t)x2);
    const uint4 q2 = make_uint4((uint32_t)y2, (uint32_t)wx1, (uint32_t)(wx2 - 1), (uint32_t)wy1);
    const uint4 q3 = make_uint4((uint32_t)(wy2 - 1), (use_x ? 1u : 0u) | (far ? 0u : 2u), sid[0], sid[1]);
    uint4 *rq = reinterpret_cast<uint4 *>(r);
    rq[0] = q0, rq[1] = q1, rq[2] = q2, rq[3] = q3;
}

// k_line_init: every pixel's winner key from the depth before the pass.
__global__ __launch_bounds__(256) void k_line_init(FrameParams p) {
    const View &v = p.v0;
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= (size_t)v.W * v.H) return;
    const bool uses_depth = p.depth_test != DT_NONE || p.depth_write;
    const uint32_t mode = (p.depth_write && p.depth_test == DT_LESS) ? 1u
                        : (p.depth_write && p.depth_test == DT_GREATER) ? 2u : 0u;
    float D0 = 0.0f;
    if (uses_depth) D0 = v.clear_depth_en ? v.clear_depth : v.depth[i];
    p.lkey[i] = init_key(mode, D0);
}

// k_line_raster: one wave per line, its lanes on the pixels k, k + 64, ... of
// the clipped Bresenham walk (clipline 0.2's contract: the unclipped line's
// pixels inside the rectangle), each updating its pixel's winner key.
__global__ __launch_bounds__(256) void k_line_raster(FrameParams p) {
    const View &v = p.v0;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6), nwaves = gridDim.x * 4u;
    const uint32_t mode = (p.depth_write && p.depth_test == DT_LESS) ? 1u
                        : (p.depth_write && p.depth_test == DT_GREATER) ? 2u : 0u;
    for (uint32_t l = wave; l < p.n_lines; l += nwaves) {
        const uint4 *rq = reinterpret_cast<const uint4 *>(p.lrec + (size_t)l * kLineRecWords);
        const uint4 q0 = rq[0], q1 = rq[1], q2 = rq[2], q3 = rq[3];
        if (!(q3.y & 2u)) continue;  // an endpoint past the coordinate limit
        const float vs0x = __uint_as_float(q0.x), vs0y = __uint_as_float(q0.y), norm = __uint_as_float(q0.z);
        const float e0z = __uint_as_float(q0.w), e1z = __uint_as_float(q1.x);
        const int64_t x1 = (int32_t)q1.y, y1 = (int32_t)q1.z, x2 = (int32_t)q1.w, y2 = (int32_t)q2.x;
        const int64_t cx0 = (int32_t)q2.y, cx1 = (int32_t)q2.z, cy0 = (int32_t)q2.w, cy1 = (int32_t)q3.x;
        const bool use_x = q3.y & 1u;
        if (cx1 < cx0 || cy1 < cy0) continue;
        const int64_t dx = x2 - x1, dy = y2 - y1;
        const int64_t adx = dx < 0 ? -dx : dx, ady = dy < 0 ? -dy : dy;
        const bool xmajor = adx >= ady;
        const int64_t M1 = xmajor ? x1 : y1, m1 = xmajor ? y1 : x1;
        const int64_t dM = xmajor ? adx : ady, dm = xmajor ? ady : adx;
        const int64_t sM = (xmajor ? dx : dy) < 0 ? -1 : 1, sm = (xmajor ? dy : dx) < 0 ? -1 : 1;
        const int64_t c0 = xmajor ? cx0 : cy0, c1 = xmajor ? cx1 : cy1;
        const int64_t n0 = xmajor ? cy0 : cx0, n1 = xmajor ? cy1 : cx1;
        int64_t klo = sM > 0 ? c0 - M1 : M1 - c1, khi = sM > 0 ? c1 - M1 : M1 - c0;
        if (klo < 0) klo = 0;
        if (khi > dM) khi = dM;
        for (int64_t k = klo + lane; k <= khi; k += 64) {
            const int64_t j = dM ? (2 * k * dm + dM - 1) / (2 * dM) : 0;
            const int64_t M = M1 + sM * k, m = m1 + sm * j;
            if (m < n0 || m > n1) continue;
            const uint32_t x = (uint32_t)(xmajor ? M : m), y = (uint32_t)(xmajor ? m : M);
            // lines.rs:88-95
            const float frac = (use_x ? (float)x - vs0x : (float)y - vs0y) * norm;
            const float z = e0z + frac * (e1z - e0z);
            if (!passes_z_clip(p, z)) continue;
            const size_t pi = (size_t)y * v.W + x;
            float D0 = 0.0f;
            if (mode == 0u && p.depth_test != DT_NONE) D0 = v.clear_depth_en ? v.clear_depth : v.depth[pi];
            update_key(p.lkey + pi, mode, p.depth_test, z, D0, l);
        }
    }
}

// get_v_data of a line (lines.rs:99-112): weighted_sum2(v0, v1, 1 - frac, frac).
template <uint32_t SH>
EUCHIP_DEV void v_data_line(const float (*V)[VDN<SH>::N], bool use_x, float vs0x, float vs0y, float norm, float X,
                            float Y, float *vd) {
    const float frac = (use_x ? X - vs0x : Y - vs0y) * norm;
    const float w0 = 1.0f - frac;
    if constexpr (SH == FLAT) {
        vd[0] = V[0][0];
    } else {
#pragma unroll
        for (uint32_t c = 0; c < vd_count(SH); c++) vd[c] = V[0][c] * w0 + V[1][c] * frac;
    }
}

// k_line_resolve: every pixel of the target: the winning line's fragment (shaded
// once, with the coarse-shading combine under MSAA), its depth, or the clears.
template <uint32_t SH>
__global__ __launch_bounds__(256) void k_line_resolve(FrameParams p) {
    const View &v = p.v0;
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= (size_t)v.W * v.H) return;
    const uint32_t px = (uint32_t)(i % v.W), py = (uint32_t)(i / v.W);
    const uint32_t mode = (p.depth_write && p.depth_test == DT_LESS) ? 1u
                        : (p.depth_write && p.depth_test == DT_GREATER) ? 2u : 0u;
    const uint64_t kk = p.lkey[i];
    bool has;
    uint32_t prim;
    float z;
    if (mode == 0u) {
        has = kk != 0ull;
        prim = (uint32_t)(kk >> 32) - 1u;
        z = __uint_as_float((uint32_t)kk);
    } else {
        has = ((uint32_t)kk) != 0u;
        prim = ((uint32_t)kk >> 1) - 1u;
        const uint32_t o = (uint32_t)(kk >> 32);
        z = zunord(mode == 1u ? o : ~o, (uint32_t)kk & 1u);
    }
    if (v.pixel) {
        if (p.pixel_write && has && SH != DEPTH_ONLY) {
            const uint4 *rq = reinterpret_cast<const uint4 *>(p.lrec + (size_t)prim * kLineRecWords);
            const uint4 q0 = rq[0], q3 = rq[3];
            const float vs0x = __uint_as_float(q0.x), vs0y = __uint_as_float(q0.y), norm = __uint_as_float(q0.z);
            const bool use_x = q3.y & 1u;
            constexpr uint32_t NV = VDN<SH>::N;
            float V[2][NV];
            const uint32_t sid[2] = {q3.z, q3.w};
#pragma unroll
            for (int k = 0; k < 2; k++)
#pragma unroll
                for (uint32_t c = 0; c < vd_count(SH); c++) V[k][c] = p.vary[(size_t)sid[k] * vd_count(SH) + c];
            if constexpr (vd_count(SH) == 0) V[0][0] = V[1][0] = 0.0f;
            float frag[4], vd[NV];
            const uint32_t L = p.msaa_level;
            if (L == 0u) {
                v_data_line<SH>(V, use_x, vs0x, vs0y, norm, (float)px, (float)py, vd);
                fragment_shader<SH, true>(v, p.sampler, vd, frag);
            } else {
                uint32_t by0, by1;
                band_of(p, v, py, by0, by1);
                const float div = 1.0f / (float)(1u << L);
                const uint32_t rx = px, ry = py - by0;
                const float fx = rfract((float)rx * div);
                const float fy = rfract((float)ry * div);
                const uint32_t cx = rx >> L, cy = ry >> L;
                float t[4][4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t ix = cx + (uint32_t)(q & 1), iy = cy + (uint32_t)(q >> 1);
                    v_data_line<SH>(V, use_x, vs0x, vs0y, norm, (float)(ix << L), (float)(by0 + (iy << L)), vd);
                    fragment_shader<SH, true>(v, p.sampler, vd, t[q]);
                }
                if constexpr (SH == FLAT) {
                    frag[0] = t[0][0];
                } else {
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        const float t0 = t[0][c] * (1.0f - fy) + t[2][c] * fy;
                        const float t1 = t[1][c] * (1.0f - fy) + t[3][c] * fy;
                        frag[c] = t0 * (1.0f - fx) + t1 * fx;
                    }
                }
            }
            v.pixel[i] = blend_shader<SH>(v.clear_color_en ? v.clear_color : v.pixel[i], frag, p.bgra);
        } else if (v.clear_color_en) {
            v.pixel[i] = v.clear_color;
        }
    }
    if (v.depth) {
        if (has && p.depth_write) v.depth[i] = z;
        else if (v.clear_depth_en) v.depth[i] = v.clear_depth;
    }
}

template <uint32_t SH>
static void launch_lines_t(const FrameParams &p, hipStream_t s) {
    const size_t nsv = (size_t)p.v0.n_instances * p.v0.n_vertices;
    const size_t npx = (size_t)p.v0.W * p.v0.H;
    hipLaunchKernelGGL((k_line_setup<SH>), dim3(blocks_of(std::max<size_t>(nsv, p.n_lines), 256)), dim3(256), 0, s, p);
    hipLaunchKernelGGL(k_line_init, dim3(blocks_of(npx, 256)), dim3(256), 0, s, p);
    if (p.n_lines) {
        const uint32_t waves = std::min<uint32_t>(p.n_lines, 4u * 2048u);
        hipLaunchKernelGGL(k_line_raster, dim3((waves + 3u) / 4u), dim3(256), 0, s, p);
    }
    hipLaunchKernelGGL((k_line_resolve<SH>), dim3(blocks_of(npx, 256)), dim3(256), 0, s, p);
}

hipError_t launch_lines(const FrameParams &p, hipStream_t s) {
    if ((size_t)p.v0.W * p.v0.H == 0) return hipSuccess;
    switch (p.shader) {
    case FLAT: launch_lines_t<FLAT>(p, s); break;
    case COLOR: launch_lines_t<COLOR>(p, s); break;
    case GOURAUD: launch_lines_t<GOURAUD>(p, s); break;
    case PHONG: launch_lines_t<PHONG>(p, s); break;
    case DEPTH_ONLY: launch_lines_t<DEPTH_ONLY>(p, s); break;
    case TEXTURED: launch_lines_t<TEXTURED>(p, s); break;
    case PHONG_SHADOW: launch_lines_t<PHONG_SHADOW>(p, s); break;
    case TEX_SHADOW: launch_lines_t<TEX_SHADOW>(p, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_of(n, 256)), dim3(256), 0, s, dst, value, n);
    return hipGetLastError();
}

}  // namespace euchip
