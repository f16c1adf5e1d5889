"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bar: colour and depth targets bit-identical (u32 / f32 bit patterns) and the
same status, for every fixed shader whose arithmetic is exact IEEE f32 (all of
them: the fixed shaders use sqrt and division only, which are correctly
rounded on both sides, and an exact x^30 chain instead of powf).
"""
import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd.pipeline import (AaMode, BandPolicy, CoordinateMode, CullMode, DepthMode, PipelineDesc, PixelMode,
                              Shader, Xform)
from tests.parity import assert_bit_exact, run_pair

pytestmark = pytest.mark.gpu


def flat(depth_mode=DepthMode.LESS_WRITE, cm=CoordinateMode.VULKAN, cull=CullMode.BACK, bands=BandPolicy.PAR,
         pixel_mode=PixelMode.WRITE, aa=AaMode.NONE):
    return PipelineDesc.make(Shader.FLAT, Xform.CLIP, pixel_mode, depth_mode, cm, cull, aa, bands)


@pytest.fixture(scope="module")
def plain():
    return S.plain_scene()


@pytest.mark.parametrize("bands", [BandPolicy.PAR, BandPolicy.SEQ])
def test_plain_640x480(gpu_ctx, plain, bands):
    """The CPU-runnable workload: 1024 random clip-space triangles, flat, LESS_WRITE."""
    r = run_pair(gpu_ctx, flat(bands=bands), plain.uniforms, plain.mesh, 640, 480)
    assert r["frags"] > 100000
    assert_bit_exact(r, f"plain {bands.name}")


@pytest.mark.parametrize("dm", [DepthMode.NONE, DepthMode.LESS_PASS, DepthMode.GREATER_WRITE,
                                DepthMode.GREATER_PASS, DepthMode(2, True), DepthMode(2, False)])
def test_plain_depth_modes(gpu_ctx, plain, dm):
    init = 0.5 if dm.test in (2, 3) else 1.0
    r = run_pair(gpu_ctx, flat(depth_mode=dm), plain.uniforms, plain.mesh, 640, 480, depth_init=init)
    assert_bit_exact(r, f"depth {dm}")


@pytest.mark.parametrize("cull", [CullMode.NONE, CullMode.FRONT])
@pytest.mark.parametrize("cm", [CoordinateMode.OPENGL, CoordinateMode.DIRECTX, CoordinateMode.VULKAN.without_z_clip()])
def test_plain_cull_coords(gpu_ctx, plain, cull, cm):
    r = run_pair(gpu_ctx, flat(cull=cull, cm=cm), plain.uniforms, plain.mesh, 640, 480)
    assert_bit_exact(r, f"{cull.name} {cm}")


def test_async_frames_changing_sizes(plain):
    """Back-to-back frames without a host sync in between, on a fresh context,
    with the tile count growing and shrinking from frame to frame: every frame
    bins into counts the previous frames left zeroed (two alternating halves,
    each re-zeroed by the frame after the one that filled it)."""
    from euc_amd.renderer import Context
    from oracle import oracle as O
    sizes = [(1920, 1080), (64, 48), (1, 1), (2048, 64), (640, 480), (1920, 1080), (33, 900), (640, 480)]
    desc = flat()
    ctx = Context(0)
    try:
        gm = ctx.mesh(plain.mesh)
        outs = []
        for w, h in sizes:
            c, d = ctx.color_target(w, h), ctx.depth_target(w, h)
            ctx.render(desc, plain.uniforms, gm, c, d, clear_color=0x01020304, clear_depth=1.0, sync=False)
            outs.append((c, d))
        ctx.synchronize()
        for (w, h), (c, d) in zip(sizes, outs):
            oc = np.full((h, w), 0x01020304, np.uint32)
            od = np.full((h, w), 1.0, np.float32)
            rc, _ = O.render(desc, plain.uniforms, plain.mesh, oc, od)
            assert_bit_exact({"o_rc": rc, "g_rc": 0, "o_color": oc, "o_depth": od, "g_color": c.download(),
                              "g_depth": d.download()}, f"async frame {w}x{h}")
    finally:
        ctx.close()


def test_big_triangle_list_full():
    """No more triangles than tiles: triangles covering more than 64 tiles go to
    the big-triangle list instead of the bins, up to 4096 of them; the rest of the
    big ones are binned as usual.  6000 triangles at 1600x960 (6000 tiles of 32x8),
    about 5250 of them with boxes past 64 tiles, overflow the list."""
    from euc_amd.renderer import Context
    rng = np.random.default_rng(0xB16)
    n = 6000
    c = rng.uniform(-1.0, 1.0, (n, 1, 2))
    off = rng.uniform(-0.35, 0.35, (n, 3, 2))
    pos = np.empty((n * 3, 4), np.float32)
    pos[:, :2] = (c + off).reshape(-1, 2)
    pos[:, 2] = rng.uniform(0.0, 1.0, n * 3)
    pos[:, 3] = 1.0
    col = (rng.integers(0, 1 << 24, n, dtype=np.uint64) | np.uint64(0xFF000000)).astype(np.uint32)
    mesh = S.Mesh(pos=pos, count=n * 3, rgba8=np.repeat(col, 3))
    desc = flat(cull=CullMode.NONE)
    ctx = Context(0)
    try:
        r = run_pair(ctx, desc, S.uniforms(), mesh, 1600, 960)
        assert r["frags"] > 1_000_000
        assert_bit_exact(r, "big-triangle list past its capacity")
    finally:
        ctx.close()


def test_big_triangle_mode_after_other_frames(plain):
    """Big-triangle mode, then a frame without it, then big-triangle mode on the
    same frame parity with another mesh: the last frame must scan only its own
    big triangles."""
    from euc_amd.renderer import Context
    from oracle import oracle as O
    rng = np.random.default_rng(0xB17)
    n = 16
    pos = np.empty((n * 3, 4), np.float32)
    pos[:, :2] = rng.uniform(-0.6, 0.6, (n * 3, 2))
    pos[:, 2] = rng.uniform(0.0, 1.0, n * 3)
    pos[:, 3] = 1.0
    col = (rng.integers(0, 1 << 24, n, dtype=np.uint64) | np.uint64(0xFF000000)).astype(np.uint32)
    other = S.Mesh(pos=pos, count=n * 3, rgba8=np.repeat(col, 3))
    desc = flat(cull=CullMode.NONE)
    ctx = Context(0)
    try:
        frames = [(plain.mesh, 640, 480), (plain.mesh, 64, 48), (other, 640, 480)]  # big, not big, big
        for mesh, w, h in frames:
            gm = ctx.mesh(mesh)
            c, d = ctx.color_target(w, h), ctx.depth_target(w, h)
            ctx.render(desc, S.uniforms(), gm, c, d, clear_color=0, clear_depth=1.0)
            oc = np.zeros((h, w), np.uint32)
            od = np.ones((h, w), np.float32)
            rc, _ = O.render(desc, S.uniforms(), mesh, oc, od)
            assert_bit_exact({"o_rc": rc, "g_rc": 0, "o_color": oc, "o_depth": od, "g_color": c.download(),
                              "g_depth": d.download()}, f"{len(mesh.pos) // 3} triangles at {w}x{h}")
    finally:
        ctx.close()


@pytest.mark.parametrize("wh", [(160, 120), (141, 141), (317, 211), (1000, 37), (33, 900), (1, 1), (2048, 64)])
@pytest.mark.parametrize("bands", [BandPolicy.PAR, BandPolicy.SEQ])
def test_plain_odd_sizes(gpu_ctx, plain, wh, bands):
    """Ragged sizes; PAR leaves targets with H / G == 0 untouched (pipeline.rs:330)."""
    w, h = wh
    r = run_pair(gpu_ctx, flat(bands=bands), plain.uniforms, plain.mesh, w, h, color_init=0x11223344)
    assert_bit_exact(r, f"{wh} {bands.name}")


def test_plain_loaded_targets(gpu_ctx, plain):
    """Targets with prior contents (LOAD, not a fused clear)."""
    rng = np.random.default_rng(7)
    c0 = rng.integers(0, 2**32, (480, 640), dtype=np.uint64).astype(np.uint32)
    d0 = rng.random((480, 640), dtype=np.float32)
    r = run_pair(gpu_ctx, flat(), plain.uniforms, plain.mesh, 640, 480, color_init=c0, depth_init=d0)
    assert_bit_exact(r, "loaded")


@pytest.fixture(scope="module")
def torus_small():
    return S.torus_scene(width=640, height=360, nu=64, nv=32)


@pytest.mark.parametrize("shader", [Shader.GOURAUD, Shader.PHONG])
@pytest.mark.parametrize("bands", [BandPolicy.PAR, BandPolicy.SEQ])
def test_torus_small(gpu_ctx, torus_small, shader, bands):
    d = PipelineDesc.make(shader, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, band_policy=bands)
    sc = torus_small
    r = run_pair(gpu_ctx, d, sc.uniforms, sc.mesh, sc.width, sc.height)
    assert r["frags"] > 1000
    assert_bit_exact(r, f"torus {shader.name} {bands.name}")


@pytest.mark.parametrize("level", [1, 2, 3])
@pytest.mark.parametrize("shader", [Shader.GOURAUD, Shader.PHONG])
def test_torus_msaa(gpu_ctx, torus_small, level, shader):
    d = PipelineDesc.make(shader, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, aa_mode=AaMode.Msaa(level))
    sc = torus_small
    r = run_pair(gpu_ctx, d, sc.uniforms, sc.mesh, sc.width, sc.height)
    assert_bit_exact(r, f"msaa {level} {shader.name}")


@pytest.fixture(scope="module")
def torus_1080p():
    return S.torus_scene()


@pytest.mark.parametrize("shader", [Shader.GOURAUD, Shader.PHONG])
def test_torus_1080p(gpu_ctx, torus_1080p, shader):
    """The metric workload at full size: 4 x 65,536-triangle tori at 1920x1080."""
    sc = torus_1080p
    d = PipelineDesc.make(shader, Xform.MVP, depth_mode=DepthMode.LESS_WRITE)
    r = run_pair(gpu_ctx, d, sc.uniforms, sc.mesh, sc.width, sc.height)
    assert r["frags"] > 400000
    assert_bit_exact(r, f"torus1080 {shader.name}")


def test_color_vertex_triangle(gpu_ctx):
    """examples/triangle.rs: one Gouraud triangle with vertex colours, Empty depth."""
    pos = np.array([[-1, -1, 0, 1], [1, -1, 0, 1], [0, 1, 0, 1]], np.float32)
    col = np.array([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1]], np.float32)
    mesh = S.Mesh(pos=pos, col=col, count=3)
    for cull in (CullMode.BACK, CullMode.NONE):
        d = PipelineDesc.make(Shader.COLOR, Xform.CLIP, cull=cull)
        r = run_pair(gpu_ctx, d, S.uniforms(), mesh, 640, 480, depth=False)
        assert_bit_exact(r, f"triangle.rs cull={cull.name}")


# --------------------------------------------------------------------------- #
# two-pass shadow mapping + texture (the heightfield workload)
# --------------------------------------------------------------------------- #

def shadow_then_main(gpu_ctx, sc, msaa_level):
    """Depth-only pass into the light's shadow map, then the textured, lit,
    shadowed main pass sampling it (examples/teapot.rs:205-222 structure)."""
    S_ = sc.extra["shadow_size"]
    sd = PipelineDesc.make(Shader.DEPTH_ONLY, Xform.MVP, pixel_mode=PixelMode.PASS,
                           depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE)
    r1 = run_pair(gpu_ctx, sd, sc.extra["shadow_uniforms"], sc.mesh, S_, S_, pixel=False)
    assert_bit_exact(r1, "shadow pass")
    shadow = r1["o_depth"]
    md = PipelineDesc.make(Shader.TEX_SHADOW, Xform.MVP, depth_mode=DepthMode.LESS_WRITE,
                           aa_mode=AaMode.Msaa(msaa_level) if msaa_level else AaMode.NONE)
    r2 = run_pair(gpu_ctx, md, sc.uniforms, sc.mesh, sc.width, sc.height, texture=sc.texture, shadow=shadow)
    return r1, r2


@pytest.mark.parametrize("msaa", [0, 1, 2])
def test_heightfield_small(gpu_ctx, msaa):
    sc = S.heightfield_scene(width=480, height=270, grid=81, tex_size=256, shadow_size=512)
    r1, r2 = shadow_then_main(gpu_ctx, sc, msaa)
    assert r1["frags"] > 1000 and r2["frags"] > 1000
    assert_bit_exact(r2, f"heightfield main msaa={msaa}")


def test_heightfield_full(gpu_ctx):
    """200k-triangle heightfield, 2048^2 shadow map, 1024^2 texture, 1080p, Msaa { level: 1 }."""
    sc = S.heightfield_scene()
    r1, r2 = shadow_then_main(gpu_ctx, sc, 1)
    assert r2["frags"] > 100000
    assert_bit_exact(r2, "heightfield full")


def test_textured_cube(gpu_ctx):
    """examples/texture_mapping.rs: a textured cube, `e as u8` blend, OpenGL-style projection."""
    rng = np.random.default_rng(1)
    tex = rng.integers(0, 2**32, (64, 64), dtype=np.uint64).astype(np.uint32)
    pos, nrm, idx = S.box(3)
    uv = (pos[:, :2] * 0.5 + 0.5).astype(np.float32)
    model = S.f32(S.translation(0, 0, 4.0) @ S.rotation(0.6, 0.8, 0.2)).reshape(1, 16)
    mesh = S.Mesh(pos=pos, uv=uv, idx=idx, count=int(idx.size), models=model)
    u = S.uniforms(view_proj=S.f32(S.perspective_fov_lh_zo(1.2, 320, 240, 0.1, 50.0)))
    for cull in (CullMode.BACK, CullMode.NONE):
        d = PipelineDesc.make(Shader.TEXTURED, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=cull)
        r = run_pair(gpu_ctx, d, u, mesh, 320, 240, texture=tex, color_init=180)
        assert r["frags"] > 1000
        assert_bit_exact(r, f"textured cull={cull.name}")


def test_phong_shadow_point_light(gpu_ctx):
    """examples/teapot.rs Teapot shader: point light, camera position, shadow lookup."""
    sc = S.torus_scene(width=400, height=300, nu=48, nv=24, instances=2)
    lvp = S.f32(S.perspective_fov_lh_zo(0.9, 512, 512, 0.1, 100.0) @ S.look_at_lh((6, 8, -2), (0, 0, 4), (0, 1, 0)))
    su = S.uniforms(view_proj=lvp)
    sd = PipelineDesc.make(Shader.DEPTH_ONLY, Xform.MVP, pixel_mode=PixelMode.PASS,
                           depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE)
    r1 = run_pair(gpu_ctx, sd, su, sc.mesh, 512, 512, pixel=False)
    assert_bit_exact(r1, "teapot-style shadow pass")
    u = sc.uniforms.copy()
    u[S.U_LPOS:S.U_LPOS + 3] = (6, 8, -2)
    u[S.U_LVP:S.U_LVP + 16] = lvp.reshape(16)
    d = PipelineDesc.make(Shader.PHONG_SHADOW, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.FRONT)
    r = run_pair(gpu_ctx, d, u, sc.mesh, sc.width, sc.height, shadow=r1["o_depth"])
    assert r["frags"] > 1000
    assert_bit_exact(r, "phong shadow")
