"""The samplers, pixel layouts and geometry stages on the GPU, against the
oracle, bit-exact: every Linear / Nearest sampler bare and wrapped in Clamped,
Tiled and Mirrored (src/sampler/*.rs) with texture coordinates far outside
[0, 1]; the BGRA blend of the teapot example (examples/teapot.rs:135-140); the
fixed geometry stages (src/pipeline.rs:221-229) under every cull mode, in single
renders and in a batch."""
import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd.pipeline import (AaMode, BandPolicy, CullMode, DepthMode, Geometry, PipelineDesc, PixelFormat, PixelMode,
                              Sampler, Shader, Xform)
from euc_amd.renderer import Request
from oracle import oracle as O
from tests.parity import assert_bit_exact, assert_same, run_pair

pytestmark = pytest.mark.gpu


def uv_cube(scale, offset):
    pos, nrm, idx = S.box(4)
    uv = (pos[:, :2] * scale + offset).astype(np.float32)
    model = S.f32(S.translation(0, 0, 4.0) @ S.rotation(0.6, 0.8, 0.2)).reshape(1, 16)
    return S.Mesh(pos=pos, nrm=nrm, uv=uv, idx=idx, count=int(idx.size), models=model)


@pytest.mark.parametrize("sampler", list(Sampler))
def test_textured_samplers(gpu_ctx, sampler):
    rng = np.random.default_rng(int(sampler) + 10)
    tex = rng.integers(0, 2 ** 32, (37, 23), dtype=np.uint64).astype(np.uint32)
    mesh = uv_cube(2.0, 0.25)  # uv from -1.75 to 2.25
    u = S.uniforms(view_proj=S.f32(S.perspective_fov_lh_zo(1.2, 160, 120, 0.1, 50.0)))
    d = PipelineDesc.make(Shader.TEXTURED, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.BACK,
                          band_policy=BandPolicy.SEQ, sampler=sampler)
    r = run_pair(gpu_ctx, d, u, mesh, 160, 120, texture=tex, color_init=7)
    assert r["frags"] > 1000
    assert_bit_exact(r, f"sampler {sampler.name}")


@pytest.mark.parametrize("sampler", [Sampler.NEAREST_TILED, Sampler.LINEAR_MIRRORED, Sampler.NEAREST_CLAMPED])
def test_tex_shadow_sampler_and_bgra(gpu_ctx, sampler):
    sc = S.heightfield_scene(width=240, height=160, grid=41, tex_size=64, shadow_size=128)
    SS = sc.extra["shadow_size"]
    sd = PipelineDesc.make(Shader.DEPTH_ONLY, Xform.MVP, pixel_mode=PixelMode.PASS,
                           depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE, band_policy=BandPolicy.SEQ)
    r1 = run_pair(gpu_ctx, sd, sc.extra["shadow_uniforms"], sc.mesh, SS, SS, pixel=False)
    assert r1["frags"] > 1000
    assert_bit_exact(r1, "shadow pass")
    m = sc.mesh
    m2 = S.Mesh(pos=m.pos, nrm=m.nrm, uv=(m.uv * 3.0 - 1.0).astype(np.float32), idx=m.idx, count=m.count,
                models=m.models)
    md = PipelineDesc.make(Shader.TEX_SHADOW, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, aa_mode=AaMode.Msaa(1),
                           band_policy=BandPolicy.SEQ, sampler=sampler, pixel_format=PixelFormat.BGRA8)
    r2 = run_pair(gpu_ctx, md, sc.uniforms, m2, sc.width, sc.height, texture=sc.texture, shadow=r1["o_depth"])
    assert r2["frags"] > 1000
    assert_bit_exact(r2, f"tex_shadow {sampler.name} bgra")


def test_phong_shadow_bgra(gpu_ctx):
    sc = S.torus_scene(width=200, height=150, nu=32, nv=16, instances=2)
    lvp = S.f32(S.perspective_fov_lh_zo(0.9, 128, 128, 0.1, 100.0) @ S.look_at_lh((6, 8, -2), (0, 0, 4), (0, 1, 0)))
    shadow = np.full((128, 128), 1.0, np.float32)
    O.render(PipelineDesc.make(Shader.DEPTH_ONLY, Xform.MVP, pixel_mode=PixelMode.PASS,
                               depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE, band_policy=BandPolicy.SEQ),
             S.uniforms(view_proj=lvp), sc.mesh, None, shadow)
    assert (shadow < 1).sum() > 100
    u = sc.uniforms.copy()
    u[S.U_LPOS:S.U_LPOS + 3] = (6, 8, -2)
    u[S.U_LVP:S.U_LVP + 16] = lvp.reshape(16)
    for fmt in (PixelFormat.RGBA8, PixelFormat.BGRA8):
        d = PipelineDesc.make(Shader.PHONG_SHADOW, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.FRONT,
                              pixel_format=fmt)
        r = run_pair(gpu_ctx, d, u, sc.mesh, sc.width, sc.height, shadow=shadow)
        assert_bit_exact(r, f"phong shadow {fmt.name}")


@pytest.mark.parametrize("geometry", [Geometry.REVERSED, Geometry.BOTH_SIDES])
@pytest.mark.parametrize("cull", [CullMode.NONE, CullMode.BACK, CullMode.FRONT])
def test_geometry_stages(gpu_ctx, geometry, cull):
    sc = S.torus_scene(width=200, height=150, nu=40, nv=20, instances=3)
    for shader, dm in [(Shader.PHONG, DepthMode.LESS_WRITE), (Shader.GOURAUD, DepthMode.NONE)]:
        d = PipelineDesc.make(shader, Xform.MVP, depth_mode=dm, cull=cull, geometry=geometry)
        r = run_pair(gpu_ctx, d, sc.uniforms, sc.mesh, sc.width, sc.height)
        assert r["frags"] > 1000
        assert_bit_exact(r, f"geometry {geometry.name} {cull.name} {shader.name}")


def test_geometry_in_a_batch(gpu_ctx):
    desc = PipelineDesc.make(Shader.PHONG, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.FRONT,
                             band_policy=BandPolicy.SEQ, geometry=Geometry.BOTH_SIDES)
    scenes = [S.icon_request(0xB0 + i, max_tris=3000, max_side=96) for i in range(6)]
    meshes = [gpu_ctx.mesh(sc.mesh) for sc in scenes]
    tg = [(gpu_ctx.color_target(sc.width, sc.height), gpu_ctx.depth_target(sc.width, sc.height)) for sc in scenes]
    b = gpu_ctx.batch().set(desc, [Request(m, c, d, sc.uniforms) for sc, m, (c, d) in zip(scenes, meshes, tg)],
                            clear_color=0xFF202020, clear_depth=1.0)
    b.run(graph=False)
    for i, (sc, (c, d)) in enumerate(zip(scenes, tg)):
        oc = np.full((sc.height, sc.width), 0xFF202020, np.uint32)
        od = np.full((sc.height, sc.width), 1.0, np.float32)
        rc, _ = O.render(desc, sc.uniforms, sc.mesh, oc, od)
        assert rc == 0
        assert_same(c.download(), d.download(), oc, od, f"batched geometry request {i}")
    b.close()
