"""Edge cases of the reference's float semantics on the GPU, against the oracle,
bit-exact (colour u32, depth f32 bits, status):

  * clip vertices with w <= 0, w = +-0, tiny w, NaN and +-Inf coordinates, and
    zero-area triangles: the homogeneous divide, winding, cull and the saturating
    `as usize` bounds of triangles.rs:55-139 (NaN -> 0, +Inf -> the target edge);
  * depth targets loaded with NaN, -0, +0, +-Inf and z landing exactly on +-0,
    under every DepthMode: `partial_cmp(z, old) == Some(ord)` (pipeline.rs:519-526),
    where NaN never passes and -0 == +0, and the written depth keeps its own sign;
  * a trailing partial triangle with several instances (triangles straddling
    instances), indexed and not (primitives.rs:23-44, index.rs:44-55);
  * MSAA levels 4-6 (pipeline.rs:291-294, 541-570);
  * the texture coordinate u = 1.0, which fract() wraps to texel 0
    (sampler/linear.rs:38-41), for the unclamped and the Clamped sampler.
"""
import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd.pipeline import (AaMode, BandPolicy, CoordinateMode, CullMode, DepthMode, PipelineDesc, PixelMode,
                              Shader, Xform)
from tests.parity import assert_bit_exact, run_pair

pytestmark = pytest.mark.gpu

NAN, INF = np.float32(np.nan), np.float32(np.inf)


def special_clip_mesh(seed, n_tris=600):
    """Triangles whose corners mix ordinary clip positions with special values."""
    rng = np.random.default_rng(seed)
    pos = np.empty((3 * n_tris, 4), np.float32)
    pos[:, :2] = rng.uniform(-1.3, 1.3, (3 * n_tris, 2))
    pos[:, 2] = rng.uniform(-0.2, 1.2, 3 * n_tris)
    pos[:, 3] = rng.uniform(0.5, 1.5, 3 * n_tris)
    special_w = np.array([0.0, -0.0, -1.0, -0.5, 1e-30, -1e-30, 1e-7, NAN, INF, -INF], np.float32)
    special_xy = np.array([NAN, INF, -INF, 0.0, -0.0, 1e30, -1e30, 3e38], np.float32)
    for v in range(3 * n_tris):
        r = rng.random()
        if r < 0.12:
            pos[v, 3] = rng.choice(special_w)
        elif r < 0.20:
            pos[v, int(rng.integers(0, 2))] = rng.choice(special_xy)
        elif r < 0.24:
            pos[v, 2] = rng.choice([NAN, INF, -INF, -0.0, 0.0])
    # zero-area triangles: repeated corners and collinear corners
    for t in range(0, n_tris, 17):
        pos[3 * t + 1] = pos[3 * t]
    for t in range(5, n_tris, 23):
        pos[3 * t + 2, :2] = 0.5 * (pos[3 * t, :2] + pos[3 * t + 1, :2])
        pos[3 * t + 2, 3] = pos[3 * t, 3] = pos[3 * t + 1, 3] = 1.0
    rgba8 = rng.integers(0, 2 ** 32, 3 * n_tris, dtype=np.uint64).astype(np.uint32)
    return S.Mesh(pos=pos, count=3 * n_tris, rgba8=rgba8)


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("cull", [CullMode.NONE, CullMode.BACK])
@pytest.mark.parametrize("cm", [CoordinateMode.VULKAN, CoordinateMode.OPENGL.without_z_clip()])
def test_nonfinite_and_degenerate_clip_vertices(gpu_ctx, seed, cull, cm):
    mesh = special_clip_mesh(seed)
    for dm, bands in [(DepthMode.LESS_WRITE, BandPolicy.PAR), (DepthMode.NONE, BandPolicy.SEQ)]:
        d = PipelineDesc.make(Shader.FLAT, Xform.CLIP, PixelMode.WRITE, dm, cm, cull, AaMode.NONE, bands)
        r = run_pair(gpu_ctx, d, S.uniforms(), mesh, 173, 151)
        assert r["frags"] > 0
        assert_bit_exact(r, f"special clip seed={seed} {cull.name} {dm} {bands.name}")


def test_nonfinite_vertices_with_interpolation(gpu_ctx):
    """COLOR interpolates with r = 1 / w_hom[2]: w = 0 corners make it infinite."""
    mesh = special_clip_mesh(9, 300)
    rng = np.random.default_rng(4)
    mesh.col = rng.uniform(-0.5, 1.5, (mesh.count, 4)).astype(np.float32)
    mesh.rgba8 = None
    d = PipelineDesc.make(Shader.COLOR, Xform.CLIP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE)
    r = run_pair(gpu_ctx, d, S.uniforms(), mesh, 160, 140)
    assert_bit_exact(r, "special clip, interpolated colour")


def depth_pattern(h, w, seed):
    rng = np.random.default_rng(seed)
    vals = np.array([NAN, -0.0, 0.0, INF, -INF, 0.5, 1.0, 0.25, -0.5], np.float32)
    return vals[rng.integers(0, len(vals), (h, w))]


def zero_z_mesh(seed):
    """Overlapping triangles whose corners all have z = -0.0, +0.0, 0.5, NaN or +-Inf:
    the interpolated z is then exactly -0.0 (all weights >= 0), +0.0, ... ."""
    rng = np.random.default_rng(seed)
    n = 240
    pos = np.empty((3 * n, 4), np.float32)
    pos[:, :2] = rng.uniform(-1.2, 1.2, (3 * n, 2))
    pos[:, 3] = 1.0
    zs = np.array([-0.0, 0.0, 0.5, 0.25, NAN, INF, -INF, 1.0], np.float32)
    pos[:, 2] = np.repeat(zs[rng.integers(0, len(zs), n)], 3)
    rgba8 = rng.integers(0, 2 ** 32, 3 * n, dtype=np.uint64).astype(np.uint32)
    return S.Mesh(pos=pos, count=3 * n, rgba8=rgba8)


@pytest.mark.parametrize("dm", [DepthMode.LESS_WRITE, DepthMode.LESS_PASS, DepthMode.GREATER_WRITE,
                                DepthMode.GREATER_PASS, DepthMode(2, True), DepthMode(2, False), DepthMode.NONE,
                                DepthMode(0, True)])
@pytest.mark.parametrize("zclip", [True, False])
def test_loaded_depth_special_values(gpu_ctx, dm, zclip):
    cm = CoordinateMode.VULKAN if zclip else CoordinateMode.VULKAN.without_z_clip()
    mesh = zero_z_mesh(11)
    w, h = 297, 203  # render_par: bands of 67 rows
    d = PipelineDesc.make(Shader.FLAT, Xform.CLIP, PixelMode.WRITE, dm, cm, CullMode.NONE, AaMode.NONE,
                          BandPolicy.PAR)
    r = run_pair(gpu_ctx, d, S.uniforms(), mesh, w, h, depth_init=depth_pattern(h, w, 5),
                 color_init=np.full((h, w), 0x55AA55AA, np.uint32))
    assert_bit_exact(r, f"loaded depth {dm} zclip={zclip}")
    if dm.write and dm.test == 1 and not zclip:
        # the winner's own zero sign is written: some -0.0 depths come from fragments
        z = r["o_depth"]
        assert np.any((z == 0) & np.signbit(z))


@pytest.mark.parametrize("indexed", [False, True])
@pytest.mark.parametrize("count", [7, 8, 10])
def test_trailing_partial_triangle_with_instances(gpu_ctx, indexed, count):
    """count % 3 != 0 with three instances: stream element e is vertex e % count of
    instance e / count, so triangles straddle instances, and the last partial one
    of the whole stream is dropped."""
    rng = np.random.default_rng(count)
    nv = 12
    pos = rng.uniform(-1.5, 1.5, (nv, 3)).astype(np.float32)
    nrm = rng.normal(size=(nv, 3)).astype(np.float32)
    idx = rng.integers(0, nv, count).astype(np.uint32) if indexed else None
    models = np.stack([S.f32(S.translation(0.3 * i, -0.2 * i, 5.0 + i) @ S.rotation(0.3 * i, 0.5, 0.1)).reshape(16)
                       for i in range(3)])
    mesh = S.Mesh(pos=pos, nrm=nrm, idx=idx, count=count, models=models)
    u = S.uniforms(view_proj=S.f32(S.perspective_fov_lh_zo(1.2, 128, 96, 0.1, 50.0)), cam=(0, 0, 0))
    for shader in (Shader.GOURAUD, Shader.PHONG):  # render_seq: render_par leaves a 128-px-wide target blank
        d = PipelineDesc.make(shader, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE,
                              band_policy=BandPolicy.SEQ)
        r = run_pair(gpu_ctx, d, u, mesh, 128, 96)
        assert r["frags"] > 0
        assert_bit_exact(r, f"partial triangle count={count} indexed={indexed} {shader.name}")


@pytest.mark.parametrize("level", [4, 5, 6])
@pytest.mark.parametrize("bands", [BandPolicy.PAR, BandPolicy.SEQ])
def test_msaa_high_levels(gpu_ctx, level, bands):
    # render_par's bands are G = 20000 * 2^L / W rows: 1000 rows need W >= 1280 at L = 6
    w, h = (1600, 1000) if bands == BandPolicy.PAR else (300, 200)
    sc = S.torus_scene(width=w, height=h, nu=32, nv=16, instances=2)
    for shader in (Shader.GOURAUD, Shader.PHONG):
        d = PipelineDesc.make(shader, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, aa_mode=AaMode.Msaa(level),
                              band_policy=bands)
        r = run_pair(gpu_ctx, d, sc.uniforms, sc.mesh, sc.width, sc.height)
        assert r["frags"] > 1000
        assert_bit_exact(r, f"msaa level {level} {bands.name} {shader.name}")


def wrap_texture():
    """Texel 0 of every row is red, the last texel blue, the rest green."""
    tex = np.full((4, 8), 0xFF00FF00, np.uint32)
    tex[:, 0] = 0xFF0000FF
    tex[:, -1] = 0xFFFF0000
    return tex


@pytest.mark.parametrize("shader", [Shader.TEXTURED, Shader.TEX_SHADOW])
def test_texture_u_one_wraps(gpu_ctx, shader):
    """Every corner has u = 1.0: the interpolated u is 1.0 exactly at many pixels
    (fract -> texel 0, red) and 1 - ulp at others (the last texel, blue)."""
    rng = np.random.default_rng(3)
    n = 200
    pos = rng.uniform(-1.2, 1.2, (3 * n, 3)).astype(np.float32)
    pos[:, 2] = rng.uniform(0.1, 0.9, 3 * n)
    uv = np.stack([np.ones(3 * n), rng.uniform(0, 1, 3 * n)], axis=1).astype(np.float32)
    nrm = np.tile(np.array([[0, 0, -1]], np.float32), (3 * n, 1))
    mesh = S.Mesh(pos=pos, nrm=nrm, uv=uv, count=3 * n, models=np.eye(4, dtype=np.float32).reshape(1, 16))
    d = PipelineDesc.make(shader, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE,
                          band_policy=BandPolicy.SEQ)
    u = S.uniforms(lights=[((0, 0, -1.0), 1.0), ((1, 0, 0), 0.0)], ambient=1.0, k_diffuse=0.0)
    shadow = np.ones((16, 16), np.float32) if shader == Shader.TEX_SHADOW else None
    r = run_pair(gpu_ctx, d, u, mesh, 120, 90, texture=wrap_texture(), shadow=shadow)
    assert_bit_exact(r, f"u = 1.0 {shader.name}")
    red = ((r["o_color"] & 0xFF) > 0x80) & (((r["o_color"] >> 16) & 0xFF) < 0x80)
    blue = ((r["o_color"] >> 16) & 0xFF) > 0x80
    assert red.sum() > 20 and blue.sum() > 20, (int(red.sum()), int(blue.sum()))
