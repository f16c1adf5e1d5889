// kernels.hip -- gfx950 kernels of the triangle path.
//
// Frame = vertex -> setup -> scan -> scatter -> raster:
//   k_vertex  one lane per (instance, vertex): the fixed vertex shader, clip
//             positions (float4) and VertexData written once per vertex.
//   k_setup   one lane per triangle: euc's triangle setup (triangles.rs:55-171)
//             in IEEE f32, cull, compaction of survivors into 128-byte TriRec
//             records, per-tile triangle counts.
//   k_scan    per-tile counts -> bin offsets (single workgroup).
//   k_scatter one lane per record (a whole wave for records spanning many
//             tiles): bin lists of record ids per 8x8 tile.
//   k_raster  one wave per 8x8 tile, one pixel per lane: coverage, z, z-clip
//             and the depth rule resolved as an order-independent reduction
//             (the winner of each pixel is a function of the set of its
//             fragments with primitive ids as tie-breaks), then the winner is
//             shaded once and colour + depth are written once.
#include "internal.hpp"
#include "shade.hpp"

namespace euchip {

EUCHIP_DEV uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// ---------------------------------------------------------------------------
// vertex shading (Pipeline::vertex of the fixed shader set)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_vertex(FrameParams p) {
    const uint32_t sid = blockIdx.x * 256u + threadIdx.x;
    const uint32_t total = p.n_instances * p.n_vertices;
    if (sid >= total) return;
    const uint32_t inst = sid / p.n_vertices;
    const uint32_t vi = sid - inst * p.n_vertices;
    const float *P = p.pos + (size_t)p.pos_comps * vi;
    const float *M = p.models + 16u * inst;
    float clip[4], wpos[4] = {0.f, 0.f, 0.f, 0.f}, wn[4] = {0.f, 0.f, 0.f, 0.f};
    if (p.xform == 0) {
        clip[0] = P[0], clip[1] = P[1], clip[2] = P[2];
        clip[3] = p.pos_comps == 4 ? P[3] : 1.0f;
        wpos[0] = clip[0], wpos[1] = clip[1], wpos[2] = clip[2];
    } else {
        float p4[4] = {P[0], P[1], P[2], 1.0f};
        mat4_vec4(M, p4, wpos);
        mat4_vec4(p.u + U_VP, wpos, clip);
    }
    if (p.nrm) {
        const float *N = p.nrm + 3u * vi;
        float n4[4] = {N[0], N[1], N[2], 0.0f};
        if (p.shader == PHONG_SHADOW) n4[0] = -N[0], n4[1] = -N[1], n4[2] = -N[2];
        if (p.xform == 0) wn[0] = n4[0], wn[1] = n4[1], wn[2] = n4[2], wn[3] = n4[3];
        else mat4_vec4(M, n4, wn);
    }
    p.clip[sid] = make_float4(clip[0], clip[1], clip[2], clip[3]);
    float *vd = p.vary + (size_t)sid * p.nvd;
    switch (p.shader) {
    case FLAT: vd[0] = __uint_as_float(p.rgba8[vi]); break;
    case COLOR:
        for (int c = 0; c < 4; c++) vd[c] = p.col[4u * vi + c];
        break;
    case GOURAUD: {
        float o[4];
        lighting2(p.u, wpos, wn, o);
        vd[0] = o[0], vd[1] = o[1], vd[2] = o[2], vd[3] = o[3];
        break;
    }
    case PHONG:
        vd[0] = wpos[0], vd[1] = wpos[1], vd[2] = wpos[2];
        vd[3] = wn[0], vd[4] = wn[1], vd[5] = wn[2];
        break;
    case TEXTURED:
        vd[0] = p.uv[2u * vi], vd[1] = p.uv[2u * vi + 1];
        break;
    case PHONG_SHADOW:
    case TEX_SHADOW: {
        float w1[4] = {wpos[0], wpos[1], wpos[2], 1.0f}, l4[4];
        mat4_vec4(p.u + U_LVP, w1, l4);
        vd[0] = wpos[0], vd[1] = wpos[1], vd[2] = wpos[2];
        vd[3] = wn[0], vd[4] = wn[1], vd[5] = wn[2];
        int o = 6;
        if (p.shader == TEX_SHADOW) {
            vd[6] = p.uv[2u * vi], vd[7] = p.uv[2u * vi + 1];
            o = 8;
        }
        vd[o + 0] = l4[0] / l4[3];
        vd[o + 1] = l4[1] / l4[3];
        vd[o + 2] = l4[2] / l4[3];
        break;
    }
    default: break;
    }
}

// ---------------------------------------------------------------------------
// Visit every tile of each active lane's tile rectangle.  Small rectangles are
// walked by their own lane; rectangles with more than 8 tiles are spread over
// the whole wave so one huge triangle does not serialise a lane.  f(tile, id)
// receives the `id` of the lane that owns the rectangle.
// ---------------------------------------------------------------------------
template <class F>
EUCHIP_DEV void for_each_tile(bool act, uint32_t tx0, uint32_t ty0, uint32_t nx, uint32_t ny, uint32_t ntx,
                              uint32_t id, F f) {
    const uint32_t n = act ? nx * ny : 0u;
    const bool big = n > 8u;
    if (!big) {
        for (uint32_t j = 0; j < n; j++) {
            uint32_t jy = j / nx, jx = j - jy * nx;
            f((ty0 + jy) * ntx + tx0 + jx, id);
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
        const uint32_t bid = __builtin_amdgcn_readlane(id, l);
        for (uint32_t j = lane; j < bn; j += 64u) {
            uint32_t jy = j / bnx, jx = j - jy * bnx;
            f((bty0 + jy) * ntx + btx0 + jx, bid);
        }
    }
}

EUCHIP_DEV bool passes_z_clip(const FrameParams &p, float z) {
    return !p.z_clip || (p.z_min <= z && z <= p.z_max);
}

// ---------------------------------------------------------------------------
// triangle setup (triangles.rs:55-171), one lane per triangle
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_setup(FrameParams p) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    bool keep = false;
    TriRec R;
    if (t < p.n_tris) {
        float hom[3][4], euc[3][3];
        uint32_t sid[3];
        const float flip1 = p.y_up ? -1.0f : 1.0f;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const uint32_t e = 3u * t + (uint32_t)i;
            const uint32_t inst = e / p.count;
            const uint32_t k = e - inst * p.count;
            const uint32_t vi = p.idx ? p.idx[k] : k;
            sid[i] = inst * p.n_vertices + vi;
            const float4 c = p.clip[sid[i]];
            hom[i][0] = c.x * 1.0f;
            hom[i][1] = c.y * flip1;
            hom[i][2] = c.z;
            hom[i][3] = c.w;
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
        bool culled = false;
        if (p.cull == 1u) culled = winding * 1.0f < 0.0f;
        else if (p.cull == 2u) culled = winding * -1.0f < 0.0f;
        if (!culled) {
            if (winding >= 0.0f) {
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    float tmp = hom[0][c];
                    hom[0][c] = hom[2][c];
                    hom[2][c] = tmp;
                }
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    float tmp = euc[0][c];
                    euc[0][c] = euc[2][c];
                    euc[2][c] = tmp;
                }
                uint32_t ts = sid[0];
                sid[0] = sid[2];
                sid[2] = ts;
            }
            const float sx = (float)p.W, sy = (float)p.H;
            const float T[3][3] = {{2.0f / sx, 0.0f, -1.0f}, {0.0f, -2.0f / sy, 1.0f}, {0.0f, 0.0f, 1.0f}};
            // coords_to_weights (triangles.rs:86-102)
            const float c3[3] = {hom[2][0], hom[2][1], hom[2][3]};
            const float ca[3] = {hom[0][0] - c3[0], hom[0][1] - c3[1], hom[0][3] - c3[2]};
            const float cb[3] = {hom[1][0] - c3[0], hom[1][1] - c3[1], hom[1][3] - c3[2]};
            const float n[3] = {ca[1] * cb[2] - ca[2] * cb[1], ca[2] * cb[0] - ca[0] * cb[2],
                                ca[0] * cb[1] - ca[1] * cb[0]};
            const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
            const float nc = n[0] * c3[0] + n[1] * c3[1] + n[2] * c3[2];
            const float rd = nn > 0.0f ? 1.0f / rmin(nc, -__FLT_EPSILON__) : 1.0f;
            float rows[3][3] = {{cb[1] * c3[2] - cb[2] * c3[1], cb[2] * c3[0] - cb[0] * c3[2],
                                 cb[0] * c3[1] - cb[1] * c3[0]},
                                {c3[1] * ca[2] - c3[2] * ca[1], c3[2] * ca[0] - c3[0] * ca[2],
                                 c3[0] * ca[1] - c3[1] * ca[0]},
                                {n[0], n[1], n[2]}};
            float M[3][3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
#pragma unroll
                for (int j = 0; j < 3; j++) rows[i][j] = rows[i][j] * rd;
#pragma unroll
                for (int j = 0; j < 3; j++) M[i][j] = rows[i][0] * T[0][j] + rows[i][1] * T[1][j] + rows[i][2] * T[2][j];
            }
            // screen coordinates and bounds (triangles.rs:110-139), band clamp = whole target
            float vs[3][2];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                vs[i][0] = sx * (euc[i][0] * 0.5f + 0.5f);
                vs[i][1] = sy * (euc[i][1] * -0.5f + 0.5f);
            }
            R.xmin = sat_clamp(rmin(rmin(vs[0][0], vs[1][0]), vs[2][0]) + 0.0f, 0u, p.W);
            R.ymin = sat_clamp(rmin(rmin(vs[0][1], vs[1][1]), vs[2][1]) + 0.0f, 0u, p.H);
            R.xmax = sat_clamp(rmax(rmax(vs[0][0], vs[1][0]), vs[2][0]) + 1.0f, 0u, p.W);
            R.ymax = sat_clamp(rmax(rmax(vs[0][1], vs[1][1]), vs[2][1]) + 1.0f, 0u, p.H);
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
            // verts_by_y (triangles.rs:148-171), ties resolved as the reference does
            const float min_y = rmin(rmin(vs[0][1], vs[1][1]), vs[2][1]);
            int ia, ib, ic;
            if (vs[0][1] == min_y) {
                if (vs[1][1] < vs[2][1]) ia = 0, ib = 1, ic = 2;
                else ia = 0, ib = 2, ic = 1;
            } else if (vs[1][1] == min_y) {
                if (vs[0][1] < vs[2][1]) ia = 1, ib = 0, ic = 2;
                else ia = 1, ib = 2, ic = 0;
            } else {
                if (vs[0][1] < vs[1][1]) ia = 2, ib = 0, ic = 1;
                else ia = 2, ib = 1, ic = 0;
            }
            R.vby[0] = vs[ia][0], R.vby[1] = vs[ia][1];
            R.vby[2] = vs[ib][0], R.vby[3] = vs[ib][1];
            R.vby[4] = vs[ic][0], R.vby[5] = vs[ic][1];
            const bool allin = passes_z_clip(p, euc[0][2]) && passes_z_clip(p, euc[1][2]) && passes_z_clip(p, euc[2][2]);
            R.flags = allin ? 1u : 0u;
            R.prim = t;
            R.pad[0] = R.pad[1] = R.pad[2] = R.pad[3] = R.pad[4] = 0u;
            keep = R.xmin < R.xmax && R.ymin < R.ymax;
        }
    }
    // wave-aggregated compaction of the surviving records
    const uint64_t mask = __ballot(keep);
    const uint32_t lane = lane_id();
    uint32_t base = 0;
    if (mask) {
        const int leader = __ffsll((unsigned long long)mask) - 1;
        if ((int)lane == leader) base = atomicAdd(p.nrec, (uint32_t)__popcll(mask));
        base = __builtin_amdgcn_readlane(base, leader);
    }
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    uint32_t tx0 = 0, ty0 = 0, nx = 0, ny = 0;
    if (keep) {
        const uint32_t r = base + rank;
        uint4 *dst = reinterpret_cast<uint4 *>(p.recs + r);
        const uint4 *src = reinterpret_cast<const uint4 *>(&R);
#pragma unroll
        for (int q = 0; q < 8; q++) dst[q] = src[q];
        tx0 = R.xmin / kTileW;
        ty0 = R.ymin / kTileH;
        nx = (R.xmax - 1u) / kTileW - tx0 + 1u;
        ny = (R.ymax - 1u) / kTileH - ty0 + 1u;
    }
    uint32_t *cnt = p.tile_count;
    for_each_tile(keep, tx0, ty0, nx, ny, p.ntx, 0u, [&](uint32_t tile, uint32_t) { atomicAdd(cnt + tile, 1u); });
}

// ---------------------------------------------------------------------------
// per-tile counts -> bin offsets; counts are reset for the next frame
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_scan(FrameParams p) {
    __shared__ uint32_t part[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t n = p.ntx * p.nty;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t beg = tid * per < n ? tid * per : n;
    const uint32_t end = beg + per < n ? beg + per : n;
    uint32_t sum = 0;
    for (uint32_t i = beg; i < end; i++) sum += p.tile_count[i];
    part[tid] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        uint32_t v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = part[tid] - sum;
    for (uint32_t i = beg; i < end; i++) {
        const uint32_t c = p.tile_count[i];
        p.tile_off[i] = run;
        p.tile_cursor[i] = run;
        p.tile_count[i] = 0u;
        run += c;
    }
    if (tid == 1023u) {
        p.tile_off[n] = part[1023];
        if (part[1023] > p.bin_cap) atomicOr(p.flags, FLAG_BIN_OVERFLOW);
    }
}

// ---------------------------------------------------------------------------
// bin lists: record ids per tile (order inside a bin is irrelevant, see k_raster)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scatter(FrameParams p) {
    if (p.flags[0] & FLAG_BIN_OVERFLOW) return;
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    const uint32_t nrec = p.nrec[0];
    const bool act = r < nrec;
    uint32_t tx0 = 0, ty0 = 0, nx = 0, ny = 0;
    if (act) {
        const TriRec &R = p.recs[r];
        tx0 = R.xmin / kTileW;
        ty0 = R.ymin / kTileH;
        nx = (R.xmax - 1u) / kTileW - tx0 + 1u;
        ny = (R.ymax - 1u) / kTileH - ty0 + 1u;
    }
    uint32_t *cur = p.tile_cursor;
    uint32_t *bins = p.bins;
    const uint32_t cap = p.bin_cap;
    for_each_tile(act, tx0, ty0, nx, ny, p.ntx, r, [&](uint32_t tile, uint32_t owner) {
        const uint32_t pos = atomicAdd(cur + tile, 1u);
        if (pos < cap) bins[pos] = owner;
    });
}

// ---------------------------------------------------------------------------
// deferred shading of one pixel's winning fragment (pipeline.rs:529-578)
// ---------------------------------------------------------------------------
EUCHIP_DEV void v_data(const FrameParams &p, const TriRec &R, const float *V0, const float *V1, const float *V2,
                       float X, float Y, float *vd) {
    // get_v_data (triangles.rs:274-294) and weighted_sum3 (math.rs:37-40)
    float w[3];
#pragma unroll
    for (int i = 0; i < 3; i++) w[i] = (R.o[i] + R.dy[i] * Y) + R.dx[i] * X;
    const float u2 = (w[2] - w[0]) - w[1];
    const float r = 1.0f / w[2];
    const float n0 = w[0] * r, n1 = w[1] * r, n2 = u2 * r;
    if (p.shader == FLAT) {
        vd[0] = V0[0];
        return;
    }
    for (uint32_t c = 0; c < p.nvd; c++) vd[c] = (V0[c] * n0 + V1[c] * n1) + V2[c] * n2;
}

EUCHIP_DEV uint32_t shade_pixel(const FrameParams &p, const TriRec &R, uint32_t px, uint32_t py, uint32_t by0,
                                uint32_t old) {
    float V[3][kMaxVD];
    const uint32_t nvd = p.nvd;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float *src = p.vary + (size_t)R.sid[i] * nvd;
        for (uint32_t c = 0; c < nvd; c++) V[i][c] = src[c];
    }
    float frag[4], vd[kMaxVD];
    const uint32_t L = p.msaa_level;
    if (L == 0u) {
        v_data(p, R, V[0], V[1], V[2], (float)px, (float)py, vd);
        fragment_shader(p, vd, frag);
    } else {
        // coarse shading on the band-relative 2^L grid, bilinear reconstruction
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
            v_data(p, R, V[0], V[1], V[2], (float)(ix << L), (float)(by0 + (iy << L)), vd);
            fragment_shader(p, vd, t[q]);
        }
        if (p.shader == FLAT) {
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
    return blend_shader(p, old, frag);
}

// ---------------------------------------------------------------------------
// raster: one 8x8 tile per wave64 workgroup
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_raster(FrameParams p) {
    if (p.flags[0] & FLAG_BIN_OVERFLOW) return;
    const uint32_t tile = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    if (tile == 0u && lane == 0u) p.nrec[0] = 0u;  // record counter of the next frame
    const uint32_t ty = tile / p.ntx, tx = tile - ty * p.ntx;
    const uint32_t px = tx * kTileW + (lane & 7u);
    const uint32_t py = ty * kTileH + (lane >> 3);
    const bool inside = px < p.W && py < p.H;
    const size_t pidx = (size_t)py * p.W + px;
    uint32_t by0 = 0, by1 = p.H;
    if (p.band_par) {
        by0 = (py / p.G) * p.G;
        by1 = by0 + p.G < p.H ? by0 + p.G : p.H;
    }
    const bool uses_depth = p.depth_test != DT_NONE || p.depth_write;
    float D0 = 0.0f;
    if (inside && uses_depth) D0 = p.clear_depth_en ? p.clear_depth : p.depth[pidx];
    // 0: winner = last fragment passing the test against D0
    // 1: LESS + write: winner = first fragment of minimum z below D0
    // 2: GREATER + write: winner = first fragment of maximum z above D0
    const uint32_t mode = (p.depth_write && p.depth_test == DT_LESS) ? 1u
                        : (p.depth_write && p.depth_test == DT_GREATER) ? 2u : 0u;
    float best_z = D0;
    uint32_t best_prim = 0u, best_rec = 0u;
    bool has = false;
    const uint32_t beg = p.tile_off[tile], end = p.tile_off[tile + 1u];
    for (uint32_t k = beg; k < end; k++) {
        const uint32_t r = p.bins[k];
        const TriRec &R = p.recs[r];
        if (!inside || py < R.ymin || py >= R.ymax) continue;
        const uint32_t bw = R.xmax - R.xmin;
        const uint32_t yl = R.ymin < by0 ? by0 : (R.ymin > by1 ? by1 : R.ymin);
        const uint32_t yh = R.ymax < by0 ? by0 : (R.ymax > by1 ? by1 : R.ymax);
        const uint32_t bh = yh - yl;
        uint32_t x0, x1;
        if (bw < 128u && bh < 128u && bw * bh < 128u) {
            x0 = R.xmin;
            x1 = R.xmax;
        } else {
            // scanline row range (triangles.rs:228-253)
            const float Y = (float)py;
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
        if (px < x0 || px >= x1) continue;
        // w_hom at the row start, then the reference's sequential w += dx walk
        const float Y = (float)py, X0 = (float)x0;
        float w0 = (R.o[0] + R.dy[0] * Y) + R.dx[0] * X0;
        float w1 = (R.o[1] + R.dy[1] * Y) + R.dx[1] * X0;
        float w2 = (R.o[2] + R.dy[2] * Y) + R.dx[2] * X0;
        for (uint32_t x = x0; x < px; x++) {
            w0 = w0 + R.dx[0];
            w1 = w1 + R.dx[1];
            w2 = w2 + R.dx[2];
        }
        const float u2 = (w2 - w0) - w1;
        if (!(w0 >= 0.0f && w1 >= 0.0f && u2 >= 0.0f)) continue;
        const float z = (R.hz[0] * w0 + R.hz[1] * w1) + R.hz[2] * u2;
        if (!((R.flags & 1u) || passes_z_clip(p, z))) continue;
        bool acc;
        if (mode == 1u) {
            acc = z < best_z || (z == best_z && has && R.prim < best_prim);
        } else if (mode == 2u) {
            acc = z > best_z || (z == best_z && has && R.prim < best_prim);
        } else {
            bool pass;
            switch (p.depth_test) {
            case DT_LESS: pass = z < D0; break;
            case DT_EQUAL: pass = z == D0; break;
            case DT_GREATER: pass = z > D0; break;
            default: pass = true; break;
            }
            acc = pass && (!has || R.prim > best_prim);
        }
        if (acc) {
            has = true;
            best_z = z;
            best_prim = R.prim;
            best_rec = r;
        }
    }
    if (!inside) return;
    if (p.pixel) {
        if (p.pixel_write && has) {
            if (p.shader != DEPTH_ONLY) {
                const uint32_t old = p.clear_color_en ? p.clear_color : 0u;
                p.pixel[pidx] = shade_pixel(p, p.recs[best_rec], px, py, by0, old);
            } else if (p.clear_color_en) {
                p.pixel[pidx] = p.clear_color;
            }
        } else if (p.clear_color_en) {
            p.pixel[pidx] = p.clear_color;
        }
    }
    if (p.depth) {
        if (has && p.depth_write) p.depth[pidx] = best_z;
        else if (p.clear_depth_en) p.depth[pidx] = p.clear_depth;
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *dst, uint32_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) dst[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline uint32_t blocks(size_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

hipError_t launch_vertex(const FrameParams &p, hipStream_t s) {
    const size_t n = (size_t)p.n_instances * p.n_vertices;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_vertex, dim3(blocks(n, 256)), dim3(256), 0, s, p);
    return hipGetLastError();
}
hipError_t launch_setup(const FrameParams &p, hipStream_t s) {
    if (p.n_tris == 0) return hipSuccess;
    hipLaunchKernelGGL(k_setup, dim3(blocks(p.n_tris, 256)), dim3(256), 0, s, p);
    return hipGetLastError();
}
hipError_t launch_scan(const FrameParams &p, hipStream_t s) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, p);
    return hipGetLastError();
}
hipError_t launch_scatter(const FrameParams &p, hipStream_t s) {
    if (p.n_tris == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3(blocks(p.n_tris, 256)), dim3(256), 0, s, p);
    return hipGetLastError();
}
hipError_t launch_raster(const FrameParams &p, hipStream_t s) {
    hipLaunchKernelGGL(k_raster, dim3(p.ntx * p.nty), dim3(64), 0, s, p);
    return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks(n, 256)), dim3(256), 0, s, dst, value, n);
    return hipGetLastError();
}

}  // namespace euchip
