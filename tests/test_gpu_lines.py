"""The lines rasterizer on the GPU (LineList and LineTriangleList,
src/rasterizer/lines.rs:15-121, src/primitives.rs:46-104) against the oracle,
bit-exact: random clip-space lines under every depth mode and both band
policies, MSAA, special clip coordinates (w <= 0 clamped to 0.0001, NaN, far
endpoints), torus wireframes with the geometry stages, and the reference's
wireframes example (teapot.obj as a LineTriangleList, BGRA red)."""
import os

import numpy as np
import pytest

from euc_amd import io
from euc_amd import scenes as S
from euc_amd.pipeline import (AaMode, BandPolicy, CoordinateMode, CullMode, DepthMode, Geometry, PipelineDesc,
                              PixelMode, Primitives, Shader, Xform)
from tests.parity import assert_bit_exact, run_pair

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")


def random_lines(seed, n, special=False):
    rng = np.random.default_rng(seed)
    pos = np.empty((2 * n, 4), np.float32)
    pos[:, :2] = rng.uniform(-1.3, 1.3, (2 * n, 2))
    pos[:, 2] = rng.uniform(-0.1, 1.1, 2 * n)
    pos[:, 3] = rng.uniform(0.5, 1.5, 2 * n)
    if special:
        k = rng.random(2 * n)
        pos[k < 0.05, 3] = 0.0
        pos[(k >= 0.05) & (k < 0.1), 3] = -1.0
        pos[(k >= 0.1) & (k < 0.13), 0] = np.nan
        pos[(k >= 0.13) & (k < 0.16), 1] = 1e30
        pos[(k >= 0.16) & (k < 0.2), 2] = np.nan
    rgba8 = rng.integers(0, 2 ** 32, 2 * n, dtype=np.uint64).astype(np.uint32)
    col = rng.uniform(-0.2, 1.2, (2 * n, 4)).astype(np.float32)
    return S.Mesh(pos=pos, count=2 * n, rgba8=rgba8, col=col)


@pytest.mark.parametrize("dm", [DepthMode.LESS_WRITE, DepthMode.NONE, DepthMode.GREATER_WRITE, DepthMode.LESS_PASS,
                                DepthMode(2, True)])
@pytest.mark.parametrize("bands", [BandPolicy.PAR, BandPolicy.SEQ])
def test_line_list(gpu_ctx, dm, bands):
    mesh = random_lines(1, 700)
    for shader in (Shader.FLAT, Shader.COLOR):
        d = PipelineDesc.make(shader, Xform.CLIP, depth_mode=dm, band_policy=bands, primitives=Primitives.LINE_LIST)
        init = 0.5 if dm.test in (2, 3) else 1.0
        r = run_pair(gpu_ctx, d, S.uniforms(), mesh, 331, 241, depth_init=init)
        assert r["frags"] > 1000 or dm.test == 2
        assert_bit_exact(r, f"lines {shader.name} {dm} {bands.name}")


@pytest.mark.parametrize("level", [1, 2])
def test_line_list_msaa(gpu_ctx, level):
    mesh = random_lines(2, 400)
    d = PipelineDesc.make(Shader.COLOR, Xform.CLIP, depth_mode=DepthMode.LESS_WRITE, aa_mode=AaMode.Msaa(level),
                          primitives=Primitives.LINE_LIST)
    r = run_pair(gpu_ctx, d, S.uniforms(), mesh, 600, 400)
    assert r["frags"] > 1000
    assert_bit_exact(r, f"lines msaa {level}")


@pytest.mark.parametrize("cm", [CoordinateMode.VULKAN, CoordinateMode.OPENGL.without_z_clip()])
def test_line_list_special_coordinates(gpu_ctx, cm):
    mesh = random_lines(3, 600, special=True)
    d = PipelineDesc.make(Shader.COLOR, Xform.CLIP, PixelMode.WRITE, DepthMode.LESS_WRITE, cm, CullMode.NONE,
                          AaMode.NONE, BandPolicy.PAR, primitives=Primitives.LINE_LIST)
    r = run_pair(gpu_ctx, d, S.uniforms(), mesh, 200, 150)
    assert_bit_exact(r, f"lines special {cm}")


@pytest.mark.parametrize("geometry", [Geometry.IDENTITY, Geometry.BOTH_SIDES])
def test_torus_wireframe(gpu_ctx, geometry):
    sc = S.torus_scene(width=320, height=240, nu=24, nv=12, instances=2)
    for shader in (Shader.GOURAUD, Shader.PHONG):
        d = PipelineDesc.make(shader, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, geometry=geometry,
                              primitives=Primitives.LINE_TRIANGLE_LIST)
        r = run_pair(gpu_ctx, d, sc.uniforms, sc.mesh, sc.width, sc.height)
        assert r["frags"] > 1000
        assert_bit_exact(r, f"torus wireframe {geometry.name} {shader.name}")


def test_wireframes_example(gpu_ctx):
    """examples/wireframes.rs: teapot.obj as a LineTriangleList, a constant red in the
    BGRA u32 layout, no depth, 1280 x 960, camera 6 units away."""
    m = io.load_obj(os.path.join(DATA, "teapot.obj"))
    w, h = 1280, 960
    mesh = S.Mesh(pos=m.pos, count=m.count, rgba8=np.full(m.count, 0xFFFF0000, np.uint32), n_instances=1,
                  models=S.f32(S.rotation(np.pi, 0, 0)).reshape(1, 16))
    vp = S.f32(S.perspective_fov_lh_zo(1.3, w, h, 0.01, 100.0) @ S.translation(0, 0, 6.0))
    d = PipelineDesc.make(Shader.FLAT, Xform.MVP, primitives=Primitives.LINE_TRIANGLE_LIST)
    r = run_pair(gpu_ctx, d, S.uniforms(view_proj=vp), mesh, w, h, depth=False)
    assert r["frags"] > 20000
    assert_bit_exact(r, "wireframes example")
