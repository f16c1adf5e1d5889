"""A second, independent restatement of euc 0.6.0's triangle path in pure
Python with numpy float32 scalars, for tiny targets only (test infrastructure).

It transliterates src/rasterizer/triangles.rs:15-367 and
src/pipeline.rs:396-614 (render_inner, BlitterImpl) line by line, for vertex
streams given directly in clip space and two fragment models: the FLAT
provoking colour and Rgba<f32> vertex colours (examples/triangle.rs:7-25),
with PAR/SEQ row bands.  It shares no code with oracle/euc_oracle.c; the
tests compare the two.
"""
from __future__ import annotations

import numpy as np

F = np.float32
EPS = F(np.finfo(np.float32).eps)


def f32(x):
    return F(x)


def sat(x) -> int:
    """Rust `f32 as usize` (saturating, NaN -> 0)."""
    x = float(x)
    if not (x > 0.0):
        return 0
    if x >= 2.0**64:
        return 2**64 - 1
    return int(x)


def sat_u8(x) -> int:
    x = float(x)
    if not (x > 0.0):
        return 0
    if x >= 255.0:
        return 255
    return int(x)


def fmin(a, b):
    """f32::min: the non-NaN operand wins."""
    if a != a:
        return b
    if b != b:
        return a
    return a if a < b else b


def fmax(a, b):
    if a != a:
        return b
    if b != b:
        return a
    return a if a > b else b


def fract(x):
    return F(x - F(np.trunc(x)))


def cross(a, b):
    return [F(a[1] * b[2] - a[2] * b[1]), F(a[2] * b[0] - a[0] * b[2]), F(a[0] * b[1] - a[1] * b[0])]


def sub(a, b):
    return [F(a[i] - b[i]) for i in range(3)]


def add(a, b):
    return [F(a[i] + b[i]) for i in range(3)]


def dot(a, b):
    return F(F(F(a[0] * b[0]) + F(a[1] * b[1])) + F(a[2] * b[2]))


def matmul(a, b):
    return [[F(F(F(a[i][0] * b[0][j]) + F(a[i][1] * b[1][j])) + F(a[i][2] * b[2][j])) for j in range(3)]
            for i in range(3)]


def mat3_mul_vec3(m, v):
    return [F(F(F(m[i][0] * v[0]) + F(m[i][1] * v[1])) + F(m[i][2] * v[2])) for i in range(3)]


def lerp(a, b, t):
    return F(a + F(t * F(b - a)))


class Blitter:
    def __init__(self, cfg, tgt_min, tgt_max, tgt_size, pixel, depth):
        self.cfg = cfg
        self.tgt_min, self.tgt_max, self.tgt_size = tgt_min, tgt_max, tgt_size
        self.pixel, self.depth = pixel, depth
        self.primitive_count = 0
        L = cfg["msaa_level"]
        self.msaa_level = L
        self.msaa_div = F(F(1.0) / F(1 << L))
        if L > 0:
            self.msaa_w = ((tgt_max[0] - tgt_min[0]) >> L) + 3
            self.msaa = {}
        self.frags = 0

    def test_fragment(self, x, y, z):
        t = self.cfg["depth_test"]
        if t == 0:
            return True
        old = self.depth[y, x]
        if t == 1:
            return z < old
        if t == 2:
            return z == old
        return z > old

    def fragment(self, vd):
        return vd  # FLAT: colour bits; COLOR: Rgba<f32>

    def ws2(self, a, b, w0, w1):
        if self.cfg["shader"] == "flat":
            return a
        return [F(F(a[c] * w0) + F(b[c] * w1)) for c in range(4)]

    def blend(self, frag):
        if self.cfg["shader"] == "flat":
            return frag
        bts = [sat_u8(F(fmin(fmax(F(frag[c]), F(0.0)), F(1.0)) * F(255.0))) for c in range(4)]
        return bts[0] | (bts[1] << 8) | (bts[2] << 16) | (bts[3] << 24)

    def msaa_fragment(self, x, y, gv):
        k = (x + 1, y + 1)
        ent = self.msaa.get(k)
        if ent is None or ent[0] != self.primitive_count:
            v = gv(F(self.tgt_min[0] + (x << self.msaa_level)), F(self.tgt_min[1] + (y << self.msaa_level)))
            ent = (self.primitive_count, self.fragment(v))
            self.msaa[k] = ent
        return ent[1]

    def emit_fragment(self, x, y, gv, z):
        self.frags += 1
        if self.cfg["depth_write"]:
            self.depth[y, x] = z
        if self.cfg["pixel_write"]:
            if self.msaa_level == 0:
                frag = self.fragment(gv(F(x), F(y)))
            else:
                fx = fract(F(F(x - self.tgt_min[0]) * self.msaa_div))
                fy = fract(F(F(y - self.tgt_min[1]) * self.msaa_div))
                px = (x - self.tgt_min[0]) >> self.msaa_level
                py = (y - self.tgt_min[1]) >> self.msaa_level
                t00 = self.msaa_fragment(px, py, gv)
                t10 = self.msaa_fragment(px + 1, py, gv)
                t01 = self.msaa_fragment(px, py + 1, gv)
                t11 = self.msaa_fragment(px + 1, py + 1, gv)
                t0 = self.ws2(t00, t01, F(F(1.0) - fy), fy)
                t1 = self.ws2(t10, t11, F(F(1.0) - fy), fy)
                frag = self.ws2(t0, t1, F(F(1.0) - fx), fx)
            self.pixel[y, x] = self.blend(frag)


def passes_z_clip(cfg, z):
    if cfg["z_clip"] is None:
        return True
    lo, hi = cfg["z_clip"]
    return F(lo) <= z and z <= F(hi)


def rasterize(cfg, verts, blit):
    """Triangles::rasterize (triangles.rs:15-306); verts: list of (hom[4], vd)."""
    W, H = blit.tgt_size
    size_x, size_y = F(W), F(H)
    to_ndc = [[F(F(2.0) / size_x), F(0.0), F(-1.0)], [F(0.0), F(F(-2.0) / size_y), F(1.0)],
              [F(0.0), F(0.0), F(1.0)]]
    cull_dir = {0: None, 1: F(1.0), 2: F(-1.0)}[cfg["cull"]]
    flip = [F(1.0), F(-1.0) if cfg["y_up"] else F(1.0)]
    for t in range(len(verts) // 3):
        blit.primitive_count += 1
        tri = verts[3 * t:3 * t + 3]
        hom = [[F(h[0] * flip[0]), F(h[1] * flip[1]), F(h[2]), F(h[3])] for h, _ in tri]
        vout = [v for _, v in tri]
        euc = [[F(h[0] / h[3]), F(h[1] / h[3]), F(h[2] / h[3])] for h in hom]
        winding = cross(sub(euc[1], euc[0]), sub(euc[2], euc[0]))[2]
        if cull_dir is not None and F(winding * cull_dir) < 0.0:
            continue
        if winding >= 0.0:
            hom, euc, vout = hom[::-1], euc[::-1], vout[::-1]
        a, b, c = hom
        c3 = [c[0], c[1], c[3]]
        ca = sub([a[0], a[1], a[3]], c3)
        cb = sub([b[0], b[1], b[3]], c3)
        n = cross(ca, cb)
        msq = F(F(F(n[0] * n[0]) + F(n[1] * n[1])) + F(n[2] * n[2]))
        rec_det = F(F(1.0) / fmin(dot(n, c3), -EPS)) if msq > 0.0 else F(1.0)
        rows = [[F(e * rec_det) for e in r] for r in (cross(cb, c3), cross(c3, ca), n)]
        M = matmul(rows, to_ndc)
        vs = [[F(size_x * F(F(e[0] * F(0.5)) + F(0.5))), F(size_y * F(F(e[1] * F(-0.5)) + F(0.5)))] for e in euc]
        smin, smax = blit.tgt_min, blit.tgt_max
        clamp = lambda v, lo, hi: max(lo, min(hi, v))
        bmin = [clamp(sat(F(fmin(fmin(vs[0][i], vs[1][i]), vs[2][i]) + F(0.0))), smin[i], smax[i]) for i in range(2)]
        bmax = [clamp(sat(F(fmax(fmax(vs[0][i], vs[1][i]), vs[2][i]) + F(1.0))), smin[i], smax[i]) for i in range(2)]
        o = mat3_mul_vec3(M, [F(0.0), F(0.0), F(1.0)])
        dx = [F(e * F(F(1.0) / F(1000.0))) for e in sub(mat3_mul_vec3(M, [F(1000.0), F(0.0), F(1.0)]), o)]
        dy = [F(e * F(F(1.0) / F(1000.0))) for e in sub(mat3_mul_vec3(M, [F(0.0), F(1000.0), F(1.0)]), o)]
        ys = [v[1] for v in vs]
        min_y = fmin(fmin(ys[0], ys[1]), ys[2])
        if vs[0][1] == min_y:
            vby = [vs[0], vs[1], vs[2]] if vs[1][1] < vs[2][1] else [vs[0], vs[2], vs[1]]
        elif vs[1][1] == min_y:
            vby = [vs[1], vs[0], vs[2]] if vs[0][1] < vs[2][1] else [vs[1], vs[2], vs[0]]
        else:
            vby = [vs[2], vs[0], vs[1]] if vs[0][1] < vs[1][1] else [vs[2], vs[1], vs[0]]
        allin = all(passes_z_clip(cfg, e[2]) for e in euc)
        hz = [h[2] for h in hom]

        def gv(x, y, o=o, dx=dx, dy=dy, vout=vout):
            wh = add(add(o, [F(e * y) for e in dy]), [F(e * x) for e in dx])
            wu = [wh[0], wh[1], F(F(wh[2] - wh[0]) - wh[1])]
            r = F(F(1.0) / wh[2])
            w = [F(e * r) for e in wu]
            if cfg["shader"] == "flat":
                return vout[0]
            return [F(F(F(vout[0][k] * w[0]) + F(vout[1][k] * w[1])) + F(vout[2][k] * w[2])) for k in range(4)]

        for y in range(bmin[1], bmax[1]):
            ext = (bmax[0] - bmin[0]) * (bmax[1] - bmin[1])
            if ext < 128:
                x0, x1 = bmin[0], bmax[0]
            else:
                A, B, Cc = vby
                Y = F(y)
                ac = lerp(A[0], Cc[0], F(F(Y - A[1]) / F(Cc[1] - A[1])))
                if Y < B[1]:
                    ab = lerp(A[0], B[0], F(F(Y - A[1]) / F(B[1] - A[1])))
                    lo, hi = fmin(ab, ac), fmax(ab, ac)
                else:
                    bc = lerp(B[0], Cc[0], F(F(Y - B[1]) / F(Cc[1] - B[1])))
                    lo, hi = fmin(bc, ac), fmax(bc, ac)
                fl, ce = F(np.floor(lo)), F(np.ceil(hi))
                x0 = sat(fl) if (fl >= F(bmin[0]) and fl < F(bmax[0])) else bmin[0]
                x1 = sat(ce) if (ce >= F(bmin[0]) and ce < F(bmax[0])) else bmax[0]
            w = add(add(o, [F(e * F(y)) for e in dy]), [F(e * F(x0)) for e in dx])
            for x in range(x0, x1):
                wu = [w[0], w[1], F(F(w[2] - w[0]) - w[1])]
                if wu[0] >= 0.0 and wu[1] >= 0.0 and wu[2] >= 0.0:
                    z = dot(hz, wu)
                    if (allin or passes_z_clip(cfg, z)) and blit.test_fragment(x, y, z):
                        blit.emit_fragment(x, y, gv, z)
                w = add(w, dx)


def render(cfg, hom, vd, pixel, depth, group_rows=None):
    with np.errstate(all="ignore"):  # IEEE results (inf, NaN) are the point, not an error
        return _render(cfg, hom, vd, pixel, depth, group_rows)


def _render(cfg, hom, vd, pixel, depth, group_rows=None):
    """pixel: uint32 [H, W] or None; depth: float32 [H, W] or None.  cfg keys:
    shader ('flat'|'color'), depth_test (0..3), depth_write, pixel_write, cull,
    y_up, z_clip ((lo, hi) or None), msaa_level, bands ('par'|'seq')."""
    uses_depth = cfg["depth_test"] != 0 or cfg["depth_write"]
    if not cfg["pixel_write"] and not uses_depth:
        return 0
    psz = (0, 0) if pixel is None else (pixel.shape[1], pixel.shape[0])
    dsz = (0, 0) if depth is None else (depth.shape[1], depth.shape[0])
    if cfg["pixel_write"] and uses_depth:
        assert psz == dsz
        size = psz
    elif cfg["pixel_write"]:
        size = psz
    else:
        size = dsz
    n = len(hom) - len(hom) % 3
    verts = [([F(c) for c in hom[i]], vd[i]) for i in range(n)]
    W, H = size
    frags = 0
    if cfg["bands"] == "seq":
        bands = [(0, H)]
    else:
        G = group_rows if group_rows is not None else 20000 * (1 << cfg["msaa_level"]) // max(W, 1)
        if H // G == 0:
            return 0
        bands = [(r, min(r + G, H)) for r in range(0, H, G)]
    for by0, by1 in bands:
        b = Blitter(cfg, (0, by0), (W, by1), (W, H), pixel, depth)
        rasterize(cfg, verts, b)
        frags += b.frags
    return frags
