"""Pins of the oracle's lines rasterizer (src/rasterizer/lines.rs:15-121,
LineList / LineTriangleList of src/primitives.rs:46-104) -- CPU only.

The pixel walk is `clipline::clipline` (clipline 0.2, not vendored, no tests in
the reference): parity with it is unpinned.  The oracle's stand-in is pinned to
what its contract states and what the mathematics fixes:
  * the pixels equal a textbook incremental Bresenham walk (error term
    e = 2|dm| - |dM|, step when e > 0) from the first endpoint, both ends
    included, on every small line brute-forced;
  * clipping never moves a pixel: the clipped walk is the unclipped one
    restricted to the rectangle;
  * known answers: horizontal, vertical and diagonal lines, the tie of
    (0,0)-(2,1) resolved back towards the start.
And the rasterizer around it: z and the varyings interpolate along the major
axis from the first endpoint (frac = (x - x0) / (x1 - x0) in float64 within a
few ulp), w is clamped to 0.0001, a LineTriangleList is the LineList of its
edges a-b, b-c, c-a, and render_par's bands do not change the result.
"""
import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd.pipeline import (BandPolicy, CullMode, DepthMode, Geometry, PipelineDesc, Primitives, Shader, Xform)
from oracle import oracle as O


def bresenham(x1, y1, x2, y2):
    dx, dy = x2 - x1, y2 - y1
    sx, sy = (1 if dx >= 0 else -1), (1 if dy >= 0 else -1)
    adx, ady = abs(dx), abs(dy)
    pts = []
    if adx >= ady:
        e, y = 2 * ady - adx, y1
        for i in range(adx + 1):
            pts.append((x1 + sx * i, y))
            if e > 0:
                y += sy
                e -= 2 * adx
            e += 2 * ady
    else:
        e, x = 2 * adx - ady, x1
        for i in range(ady + 1):
            pts.append((x, y1 + sy * i))
            if e > 0:
                x += sx
                e -= 2 * ady
            e += 2 * adx
    return pts


def test_walk_equals_incremental_bresenham():
    rng = np.random.default_rng(0)
    for _ in range(3000):
        x1, y1, x2, y2 = (int(v) for v in rng.integers(-20, 21, 4))
        got = O.clip_line((x1, y1), (x2, y2), (-100, -100), (100, 100))
        assert got == bresenham(x1, y1, x2, y2), (x1, y1, x2, y2)


def test_clipping_does_not_move_pixels():
    rng = np.random.default_rng(1)
    for _ in range(3000):
        x1, y1, x2, y2 = (int(v) for v in rng.integers(-30, 31, 4))
        cx0, cy0 = (int(v) for v in rng.integers(-20, 10, 2))
        cx1, cy1 = cx0 + int(rng.integers(-2, 25)), cy0 + int(rng.integers(-2, 25))
        want = [(x, y) for x, y in bresenham(x1, y1, x2, y2) if cx0 <= x <= cx1 and cy0 <= y <= cy1]
        assert O.clip_line((x1, y1), (x2, y2), (cx0, cy0), (cx1, cy1)) == want


def test_long_lines_clip_cheaply():
    pts = O.clip_line((-10 ** 8, 5), (10 ** 8, 6), (0, 0), (9, 9))
    # a 2e8-long, 1-high line is halfway up at x = 0: that half goes back to y = 5
    assert pts == [(0, 5)] + [(x, 6) for x in range(1, 10)]


def test_known_answers():
    assert O.clip_line((1, 2), (5, 2), (0, 0), (9, 9)) == [(x, 2) for x in range(1, 6)]
    assert O.clip_line((3, 7), (3, 4), (0, 0), (9, 9)) == [(3, y) for y in (7, 6, 5, 4)]
    assert O.clip_line((0, 0), (3, 3), (0, 0), (9, 9)) == [(i, i) for i in range(4)]
    assert O.clip_line((0, 0), (2, 1), (0, 0), (9, 9)) == [(0, 0), (1, 0), (2, 1)]
    assert O.clip_line((2, 1), (0, 0), (0, 0), (9, 9)) == [(2, 1), (1, 1), (0, 0)]
    assert O.clip_line((4, 4), (4, 4), (0, 0), (9, 9)) == [(4, 4)]
    assert O.clip_line((4, 4), (4, 4), (5, 5), (9, 9)) == []


def line_mesh(pos, col=None, rgba8=None):
    pos = np.asarray(pos, np.float32)
    return S.Mesh(pos=pos, count=len(pos), col=None if col is None else np.asarray(col, np.float32),
                  rgba8=None if rgba8 is None else np.asarray(rgba8, np.uint32))


def render(desc, mesh, w, h, depth0=1.0, color0=0):
    c = np.full((h, w), color0, np.uint32)
    d = np.full((h, w), depth0, np.float32)
    rc, frags = O.render(desc, S.uniforms(), mesh, c, d)
    assert rc == 0
    return frags, c, d


def lines_desc(prims=Primitives.LINE_LIST, shader=Shader.FLAT, dm=DepthMode.LESS_WRITE, bands=BandPolicy.SEQ, **kw):
    return PipelineDesc.make(shader, Xform.CLIP, depth_mode=dm, band_policy=bands, primitives=prims,
                             cull=CullMode.NONE, **kw)


def test_horizontal_line_pixels_and_depth():
    """A line from screen (2.5, 3.5) to (13.5, 3.5) on a 16 x 8 target: pixels x = 2..13
    of row 3; z from 0.25 to 0.75 at frac = (x - 2.5) / 11."""
    w, h = 16, 8
    to_ndc = lambda x, y: (2 * x / w - 1, 1 - 2 * y / h)
    p0, p1 = to_ndc(2.5, 3.5), to_ndc(13.5, 3.5)
    mesh = line_mesh([[p0[0], p0[1], 0.25, 1], [p1[0], p1[1], 0.75, 1]], rgba8=[0xAABBCCDD, 0x11223344])
    frags, c, d = render(lines_desc(), mesh, w, h)
    assert frags == 12
    drawn = np.argwhere(c != 0)
    assert drawn.tolist() == [[3, x] for x in range(2, 14)]
    assert np.all(c[3, 2:14] == 0xAABBCCDD)  # FLAT keeps the first vertex's colour
    frac = (np.arange(2, 14) - 2.5) / 11.0
    np.testing.assert_allclose(d[3, 2:14], 0.25 + frac * 0.5, rtol=0, atol=4e-7)


def test_color_interpolates_along_the_major_axis():
    w, h = 40, 30
    mesh = line_mesh([[-0.9, -0.8, 0.5, 1], [0.8, 0.7, 0.5, 1]], col=[[1, 0, 0, 1], [0, 0, 1, 1]])
    frags, c, _ = render(lines_desc(shader=Shader.COLOR, dm=DepthMode.NONE), mesh, w, h)
    ys, xs = np.nonzero(c)
    assert frags == len(xs) and frags > 20
    x0 = (-0.9 * 0.5 + 0.5) * w
    x1 = (0.8 * 0.5 + 0.5) * w
    frac = (xs - x0) / (x1 - x0)
    r = (c[ys, xs] & 0xFF).astype(np.int64)
    b = ((c[ys, xs] >> 16) & 0xFF).astype(np.int64)
    assert np.all(np.abs(r - np.floor(np.clip(1 - frac, 0, 1) * 255)) <= 1)
    assert np.all(np.abs(b - np.floor(np.clip(frac, 0, 1) * 255)) <= 1)


def test_w_is_clamped_to_a_small_positive_value():
    """w <= 0 divides by 0.0001 (lines.rs:51): a corner at (0.0001 x, ...) with w = -3
    lands where (x, ...) with w = 1 would land."""
    w, h = 64, 48
    a = line_mesh([[-0.5, 0.1, 0.5, 1], [0.00005, -0.00002, 0.00005, -3.0]], rgba8=[7, 7])
    b = line_mesh([[-0.5, 0.1, 0.5, 1], [0.5, -0.2, 0.5, 1.0]], rgba8=[7, 7])
    fa, ca, _ = render(lines_desc(dm=DepthMode.NONE), a, w, h)
    fb, cb, _ = render(lines_desc(dm=DepthMode.NONE), b, w, h)
    assert fa == fb > 0 and np.array_equal(ca, cb)


@pytest.mark.parametrize("geometry", [Geometry.IDENTITY, Geometry.BOTH_SIDES])
def test_line_triangle_list_is_the_line_list_of_its_edges(geometry):
    rng = np.random.default_rng(2)
    n = 30
    pos = np.empty((3 * n, 4), np.float32)
    pos[:, :2] = rng.uniform(-1.2, 1.2, (3 * n, 2))
    pos[:, 2] = rng.uniform(0, 1, 3 * n)
    pos[:, 3] = rng.uniform(0.5, 2.0, 3 * n)
    col = rng.uniform(0, 1, (3 * n, 4)).astype(np.float32)
    tri = line_mesh(pos, col=col)
    order = []
    for t in range(n):
        a, b, c = 3 * t, 3 * t + 1, 3 * t + 2
        outs = [(a, b, c)] + ([(a, c, b)] if geometry == Geometry.BOTH_SIDES else [])
        for (p, q, r) in outs:
            order += [p, q, q, r, r, p]
    edges = line_mesh(pos[order], col=col[order])
    ga = render(lines_desc(Primitives.LINE_TRIANGLE_LIST, Shader.COLOR, geometry=geometry), tri, 50, 40)
    gb = render(lines_desc(Primitives.LINE_LIST, Shader.COLOR), edges, 50, 40)
    assert ga[0] == gb[0] > 0
    assert np.array_equal(ga[1], gb[1]) and np.array_equal(ga[2].view(np.uint32), gb[2].view(np.uint32))


def test_bands_do_not_change_lines():
    rng = np.random.default_rng(3)
    pos = np.empty((200, 4), np.float32)
    pos[:, :2] = rng.uniform(-1.3, 1.3, (200, 2))
    pos[:, 2] = rng.uniform(0, 1, 200)
    pos[:, 3] = 1.0
    mesh = line_mesh(pos, rgba8=rng.integers(0, 2 ** 32, 200, dtype=np.uint64).astype(np.uint32))
    seq = render(lines_desc(), mesh, 70, 53)
    par = lines_desc(bands=BandPolicy.PAR)
    par.threads = 4
    par.band_rows_override = 5
    got = render(par, mesh, 70, 53)
    assert got[0] == seq[0] and np.array_equal(got[1], seq[1])


def test_trailing_partial_line_is_dropped():
    pos = [[-0.5, 0, 0.5, 1], [0.5, 0, 0.5, 1], [0, 0.5, 0.5, 1]]
    a = render(lines_desc(dm=DepthMode.NONE), line_mesh(pos, rgba8=[1, 2, 3]), 20, 20)
    b = render(lines_desc(dm=DepthMode.NONE), line_mesh(pos[:2], rgba8=[1, 2]), 20, 20)
    assert a[0] == b[0] and np.array_equal(a[1], b[1])
