"""Closed-form pins of the oracle's fixed shader models and of its coarse-shading
(MSAA) combine -- CPU only.

The reference's example shaders (examples/teapot.rs:52-141,
examples/texture_mapping.rs:12-35) are written against `vek`, which is not
vendored and has no tests, so the oracle's restatement of them is pinned here by
what the mathematics fixes:
  * PHONG_SHADOW (teapot.rs:99-132): a point light at angle phi from the surface
    normal and a camera at angle theta give light = ambient + kd*cos(phi) +
    ks*cos(phi + theta)^30 in light, ambient in shadow; the shadow-map lookup
    (teapot.rs:123-128) maps light-space (x, y) to uv = (x/2 + 1/2, -y/2 + 1/2)
    and is strict against depth + bias;
  * TEX_SHADOW: a constant texture and shadow map give texel * light / 255;
  * the vertex chain (teapot.rs:82-93): clip = VP * (model * (p, 1)) against a
    float64 product, the negated normal of the teapot model, and the light-space
    perspective divide;
  * the MSAA combine (pipeline.rs:541-570): bilinear reconstruction reproduces an
    affine colour field exactly (so MSAA equals per-pixel shading to one LSB), and
    for a perspective (non-affine) field the coarse grid sits at the band-relative
    origin (0, by0) -- a float64 model of that grid, with band heights that are
    not multiples of 2^L, agrees with the oracle to one LSB while the model with
    an absolute grid origin does not.
"""
import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd.pipeline import AaMode, BandPolicy, CullMode, DepthMode, PipelineDesc, Shader, Xform
from oracle import oracle as O

EPS = float(np.finfo(np.float32).eps)


def desc(shader, **kw):
    return PipelineDesc.make(shader, Xform.MVP, **kw)


def dir3(angle):
    """Unit vector at `angle` (radians) from +z towards +x."""
    return np.array([np.sin(angle), 0.0, np.cos(angle)])


# --------------------------------------------------------------------------- #
# PHONG_SHADOW
# --------------------------------------------------------------------------- #

def phong_shadow_vd(wpos, normal, lvp):
    return np.concatenate([wpos, normal, lvp]).astype(np.float32)


@pytest.mark.parametrize("phi,theta", [(0.0, 0.0), (0.3, -0.3), (0.6, -0.5), (0.25, 0.2), (1.2, -0.1),
                                       (2.0, -2.0)])
def test_phong_shadow_point_light_closed_form(phi, theta):
    """n = +z at the origin, light at 10 * dir(phi), camera at 7 * dir(theta):
    diffuse = kd * max(cos phi, 0), specular = ks * max(cos(phi + theta), 0)^30."""
    amb, kd, ks = 0.125, 0.5, 2.0
    surf = (1.0, 0.5, 0.25, 1.0)
    u = S.uniforms(cam=tuple(7.0 * dir3(theta)), surf=surf, ambient=amb, k_diffuse=kd, k_specular=ks,
                   shadow_bias=2.0 ** -10, light_pos=tuple(10.0 * dir3(phi)))
    lit = np.ones((4, 4), np.float32)        # depth 1 everywhere: z = 0.5 is lit
    dark = np.zeros((4, 4), np.float32)      # depth 0: z = 0.5 is in shadow
    vd = phong_shadow_vd((0, 0, 0), (0, 0, 2.0), (0.1, -0.2, 0.5))
    light = amb + kd * max(np.cos(phi), 0.0) + ks * max(np.cos(phi + theta), 0.0) ** 30
    out, px = O.fragment(desc(Shader.PHONG_SHADOW), u, vd, shadow=lit)
    np.testing.assert_allclose(out, np.array(surf) * light, rtol=3e-5, atol=1e-7)
    # blend: clamp01 * 255 as u8 (teapot.rs:135-140), one LSB for the f32 rounding
    want = np.floor(np.clip(np.array(surf) * light, 0, 1) * 255)
    got = np.array([(px >> (8 * c)) & 0xFF for c in range(4)])
    assert np.all(np.abs(got - want) <= 1), (got, want)
    out, _ = O.fragment(desc(Shader.PHONG_SHADOW), u, vd, shadow=dark)
    np.testing.assert_allclose(out, np.array(surf) * amb, rtol=1e-6)


def test_phong_shadow_mirror_direction_is_full_lobe():
    """Camera exactly on the mirror ray (theta = -phi) sees the whole specular
    coefficient; a sign slip in the reflection would see cos(2 phi)^30 instead."""
    phi = 0.7
    u = S.uniforms(cam=tuple(5.0 * dir3(-phi)), surf=(1, 1, 1, 1), ambient=0.0, k_diffuse=0.0, k_specular=1.0,
                   light_pos=tuple(3.0 * dir3(phi)))
    vd = phong_shadow_vd((0, 0, 0), (0, 0, 1.0), (0, 0, 0.5))
    out, _ = O.fragment(desc(Shader.PHONG_SHADOW), u, vd, shadow=np.ones((2, 2), np.float32))
    np.testing.assert_allclose(out[0], 1.0, rtol=1e-5)
    assert np.cos(2 * phi) ** 30 < 0.01


def test_shadow_lookup_bias_boundary():
    """in_light = z < depth + bias, strict: with depth 0.5 and bias 2^-10 the sum
    0.5 + 2^-10 is exact, so z equal to it is shadowed and z just below it is lit."""
    bias = 2.0 ** -10
    u = S.uniforms(cam=(0, 0, 5.0), surf=(1, 1, 1, 1), ambient=0.25, k_diffuse=0.5, k_specular=0.0,
                   shadow_bias=bias, light_pos=(0, 0, 5.0))
    smap = np.full((8, 8), 0.5, np.float32)
    for z, lit in [(0.5 + bias, False), (0.5 + bias / 2, True), (0.5 + 2 * bias, False), (0.5, True)]:
        vd = phong_shadow_vd((0, 0, 0), (0, 0, 1.0), (0.3, 0.3, z))
        out, _ = O.fragment(desc(Shader.PHONG_SHADOW), u, vd, shadow=smap)
        assert out[0] == pytest.approx(0.75 if lit else 0.25, rel=1e-6), (z, out)


@pytest.mark.parametrize("x,y,lit", [(-0.5, 0.5, True), (-0.5, -0.5, False), (0.5, 0.5, False), (0.5, -0.5, False)])
def test_shadow_lookup_uv_mapping(x, y, lit):
    """uv = (x * 0.5 + 0.5, -y * 0.5 + 0.5): light-space (-0.5, 0.5) lands on texel
    (1, 1) of a 4x4 map (fractions 0), inside the far (depth 1) top-left quadrant;
    the other quadrants hold depth 0.  A missing y flip or x sign would swap them."""
    smap = np.zeros((4, 4), np.float32)
    smap[:2, :2] = 1.0
    u = S.uniforms(cam=(0, 0, 5.0), surf=(1, 1, 1, 1), ambient=0.25, k_diffuse=0.5, k_specular=0.0,
                   light_pos=(0, 0, 5.0))
    vd = phong_shadow_vd((0, 0, 0), (0, 0, 1.0), (x, y, 0.5))
    out, _ = O.fragment(desc(Shader.PHONG_SHADOW), u, vd, shadow=smap)
    assert out[0] == pytest.approx(0.75 if lit else 0.25, rel=1e-6)


# --------------------------------------------------------------------------- #
# TEX_SHADOW and TEXTURED
# --------------------------------------------------------------------------- #

@pytest.mark.parametrize("angle,lit", [(0.0, True), (np.pi / 3, True), (np.pi / 3, False), (2.0, True)])
def test_tex_shadow_constant_inputs(angle, lit):
    """Constant RGBA8 texture t and constant shadow map: out.rgb = t.rgb * light / 255,
    out.a = t.a / 255, light = ambient + [lit] kd * max(n.l0, 0)."""
    amb, kd = 0.2, 0.7
    texel = (200, 100, 50, 255)
    tex = np.full((8, 8), texel[0] | texel[1] << 8 | texel[2] << 16 | texel[3] << 24, np.uint32)
    smap = np.full((8, 8), 1.0 if lit else 0.0, np.float32)
    u = S.uniforms(lights=[((0.0, 0.0, 1.0), 1.0), ((1.0, 0.0, 0.0), 0.0)], ambient=amb, k_diffuse=kd)
    n = 3.0 * dir3(angle)
    vd = np.concatenate([(0, 0, 0), n, (0.4, 0.6), (0.1, 0.1, 0.5)]).astype(np.float32)
    out, px = O.fragment(desc(Shader.TEX_SHADOW), u, vd, texture=tex, shadow=smap)
    light = amb + (kd * max(np.cos(angle), 0.0) if lit else 0.0)
    want = np.array([texel[0] * light / 255, texel[1] * light / 255, texel[2] * light / 255, 1.0])
    np.testing.assert_allclose(out, want, rtol=2e-6)
    got = np.array([(px >> (8 * c)) & 0xFF for c in range(4)])
    assert np.all(np.abs(got - np.floor(np.clip(want, 0, 1) * 255)) <= 1)


def test_textured_constant_texture_blend():
    """TEXTURED's blend is `e as u8` on 0..255 channel values (texture_mapping.rs:31-34):
    a constant texture comes back unchanged."""
    tex = np.full((5, 3), 0x80FF1002, np.uint32)
    out, px = O.fragment(desc(Shader.TEXTURED), S.uniforms(), (0.3, 0.7), texture=tex)
    assert list(out) == [2.0, 16.0, 255.0, 128.0]
    assert px == 0x80FF1002


# --------------------------------------------------------------------------- #
# the vertex chain
# --------------------------------------------------------------------------- #

def f64_clip_bound(m, v):
    """Worst-case rounding of a left-to-right f32 evaluation of m @ v."""
    return 8 * EPS * (np.abs(m) @ np.abs(v)) + 1e-30


@pytest.mark.parametrize("shader", [Shader.PHONG, Shader.PHONG_SHADOW, Shader.TEX_SHADOW, Shader.GOURAUD])
def test_vertex_mvp_chain(shader):
    model = S.f32(S.translation(1.5, -2.0, 0.25) @ S.rotation(0.3, -0.4, 0.2) @ S.scaling(2.0))
    vp = S.f32(S.perspective_fov_lh_zo(1.1, 640, 480, 0.1, 100.0) @ S.look_at_lh((0, 1, -6), (0, 0, 0), (0, 1, 0)))
    lvp = S.f32(S.perspective_fov_lh_zo(0.75, 512, 512, 0.5, 50.0) @ S.look_at_lh((4, 6, -3), (0, 0, 0), (0, 1, 0)))
    rng = np.random.default_rng(7)
    pos = rng.uniform(-1, 1, (16, 3)).astype(np.float32)
    nrm = rng.normal(size=(16, 3)).astype(np.float32)
    uv = rng.uniform(0, 1, (16, 2)).astype(np.float32)
    mesh = S.Mesh(pos=pos, nrm=nrm, uv=uv, count=16, models=np.stack([np.eye(4, dtype=np.float32), model]))
    u = S.uniforms(view_proj=vp, light_vp=lvp, cam=(0.5, 1.0, -4.0))
    d = desc(shader)
    M64, VP64, LVP64 = model.astype(np.float64), vp.astype(np.float64), lvp.astype(np.float64)
    for vi in range(16):
        clip, vd = O.vertex(d, u, mesh, 1, vi)
        p4 = np.append(pos[vi].astype(np.float64), 1.0)
        world = M64 @ p4
        want = VP64 @ world
        # the world position rounds once per row; carry that bound through VP
        tol = f64_clip_bound(VP64, world) + np.abs(VP64) @ f64_clip_bound(M64, p4)
        assert np.all(np.abs(clip - want) <= tol), (vi, clip, want)
        sign = -1.0 if shader == Shader.PHONG_SHADOW else 1.0   # teapot.rs:87 negates the normal
        n4 = np.append(sign * nrm[vi].astype(np.float64), 0.0)
        wn = M64 @ n4
        if shader == Shader.GOURAUD:
            want_c = O.lighting2(u, world[:3].astype(np.float32), wn[:3].astype(np.float32))
            np.testing.assert_allclose(vd, want_c, rtol=1e-4, atol=1e-5)
            continue
        np.testing.assert_allclose(vd[:3], world[:3], rtol=0, atol=float(f64_clip_bound(M64, p4)[:3].max()))
        np.testing.assert_allclose(vd[3:6], wn[:3], rtol=0, atol=float(f64_clip_bound(M64, n4)[:3].max()))
        if shader in (Shader.PHONG_SHADOW, Shader.TEX_SHADOW):
            l4 = LVP64 @ np.append(world[:3], 1.0)
            o = 8 if shader == Shader.TEX_SHADOW else 6
            np.testing.assert_allclose(vd[o:o + 3], l4[:3] / l4[3], rtol=2e-5, atol=2e-6)
            if shader == Shader.TEX_SHADOW:
                assert list(vd[6:8]) == list(uv[vi])


def test_vertex_clip_input_is_passed_through():
    pos = np.array([[0.25, -0.5, 0.75, 2.0], [1, 2, 3, 4]], np.float32)
    mesh = S.Mesh(pos=pos, count=2, rgba8=np.array([7, 9], np.uint32))
    clip, vd = O.vertex(PipelineDesc.make(Shader.FLAT, Xform.CLIP), S.uniforms(), mesh, 0, 1)
    assert list(clip) == [1, 2, 3, 4] and vd.view(np.uint32)[0] == 9


# --------------------------------------------------------------------------- #
# MSAA coarse shading (pipeline.rs:541-570)
# --------------------------------------------------------------------------- #

def color_tri(pos, col):
    pos = np.asarray(pos, np.float32)
    return S.Mesh(pos=pos, count=len(pos), col=np.asarray(col, np.float32))


def render_color(mesh, w, h, level, group_rows=0, bands=BandPolicy.SEQ):
    d = PipelineDesc.make(Shader.COLOR, Xform.CLIP, cull=CullMode.NONE, band_policy=bands,
                          aa_mode=AaMode.Msaa(level) if level else AaMode.NONE, threads=3)
    if group_rows:
        d.band_rows_override = group_rows
    c = np.full((h, w), 0x01010101, np.uint32)
    rc, _ = O.render(d, S.uniforms(), mesh, c, None)
    assert rc == 0
    return c


def channels(c):
    return c.view(np.uint8).reshape(c.shape + (4,)).astype(np.int32)


@pytest.mark.parametrize("level", [1, 2, 3])
def test_msaa_reproduces_an_affine_field(level):
    """w = 1 everywhere: colours are affine in screen space and bilinear
    reconstruction of an affine function is exact, so coarse shading matches
    per-pixel shading to one LSB (f32 rounding) at every pixel, bands included."""
    big = [[-3, -3, 0.5, 1], [5, -3, 0.5, 1], [-3, 5, 0.5, 1]]  # covers the target
    col = [[0.0, 0.1, 0.9, 1.0], [1.0, 0.2, 0.1, 0.5], [0.3, 1.0, 0.0, 0.2]]
    mesh = color_tri(big, col)
    ref = channels(render_color(mesh, 61, 47, 0))
    for bands, g in [(BandPolicy.SEQ, 0), (BandPolicy.PAR, 7)]:
        got = channels(render_color(mesh, 61, 47, level, group_rows=g, bands=bands))
        assert np.abs(got - ref).max() <= 1


def f64_color(mesh, w, h, X, Y):
    """Perspective-correct colour of the (extended) triangle plane at screen point
    (X, Y): screen-space barycentrics from the NDC vertices, divided by each w."""
    p = mesh.pos.astype(np.float64)
    ndc = p[:, :2] / p[:, 3:4]
    sx = (ndc[:, 0] * 0.5 + 0.5) * w
    sy = (ndc[:, 1] * -0.5 + 0.5) * h
    A = np.array([[sx[0], sx[1], sx[2]], [sy[0], sy[1], sy[2]], [1, 1, 1]])
    lam = np.linalg.solve(A, np.array([X, Y, 1.0]))
    b = lam / p[:, 3]
    b = b / b.sum()
    return b @ mesh.col.astype(np.float64)


def msaa_model(mesh, w, h, X, Y, level, by0):
    s = 1 << level
    rx, ry = X, Y - by0
    fx, fy = (rx % s) / s, (ry % s) / s
    cx, cy = (rx // s) * s, (ry // s) * s
    t = {(i, j): f64_color(mesh, w, h, cx + i * s, by0 + cy + j * s) for i in (0, 1) for j in (0, 1)}
    t0 = t[0, 0] * (1 - fy) + t[0, 1] * fy
    t1 = t[1, 0] * (1 - fy) + t[1, 1] * fy
    return t0 * (1 - fx) + t1 * fx


@pytest.mark.parametrize("level,group_rows", [(1, 7), (2, 10), (3, 13)])
def test_msaa_grid_origin_is_band_relative(level, group_rows):
    """A perspective colour field (vertex w from 0.3 to 3) is not affine, so the
    coarse grid's origin shows: with bands of G rows (G not a multiple of 2^L)
    every shaded pixel agrees to one LSB with the float64 model whose grid starts
    at the band's first row, and not with one whose grid starts at row 0."""
    w, h = 40, 41
    tri = [[-0.95 * 0.3, -0.9 * 0.3, 0.5, 0.3], [0.95 * 3.0, -0.95 * 3.0, 0.5, 3.0], [-0.9, 0.98, 0.5, 1.0]]
    col = [[1.0, 0.0, 0.0, 1.0], [0.0, 1.0, 0.0, 1.0], [0.0, 0.0, 1.0, 1.0]]
    mesh = color_tri(tri, col)
    got = render_color(mesh, w, h, level, group_rows=group_rows, bands=BandPolicy.PAR)
    shaded = got != 0x01010101
    assert shaded.sum() > 300
    gc = channels(got)
    worst_rel = worst_abs = 0
    for Y, X in zip(*np.nonzero(shaded)):
        by0 = (Y // group_rows) * group_rows
        rel = np.floor(np.clip(msaa_model(mesh, w, h, X, Y, level, by0), 0, 1) * 255)
        ab = np.floor(np.clip(msaa_model(mesh, w, h, X, Y, level, 0), 0, 1) * 255)
        worst_rel = max(worst_rel, np.abs(gc[Y, X] - rel).max())
        worst_abs = max(worst_abs, np.abs(gc[Y, X] - ab).max())
    assert worst_rel <= 1
    assert worst_abs >= 3
