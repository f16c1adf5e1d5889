"""The reference's own workloads on the GPU, from its own input files (loaded by
euc_amd/io.py), against the oracle, bit-exact:

  * benches/teapot.rs:143-205 -- teapot.obj, a 512^2 depth-only shadow pass then
    the Phong + shadow-lookup colour pass with the BGRA blend, at 640x480 and at
    2048x1600 (two of the benchmark's sizes);
  * examples/texture_mapping.rs -- the textured cube with rust.png through the
    bare Linear sampler, and checkerboard.png through Nearest.
"""
import os

import numpy as np
import pytest

from euc_amd import io
from euc_amd import scenes as S
from euc_amd.pipeline import CullMode, DepthMode, PipelineDesc, PixelFormat, PixelMode, Sampler, Shader, Xform
from tests.parity import assert_bit_exact, run_pair

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")


@pytest.fixture(scope="module")
def teapot():
    return io.load_obj(os.path.join(DATA, "teapot.obj"))


@pytest.mark.parametrize("wh", [(640, 480), (2048, 1600)])
def test_teapot_benchmark_frame(gpu_ctx, teapot, wh):
    sc = S.teapot_scene(teapot, *wh)
    SS = sc.extra["shadow_size"]
    sd = PipelineDesc.make(Shader.DEPTH_ONLY, Xform.MVP, pixel_mode=PixelMode.PASS, depth_mode=DepthMode.LESS_WRITE,
                           cull=CullMode.NONE)
    r1 = run_pair(gpu_ctx, sd, sc.extra["shadow_uniforms"], sc.mesh, SS, SS, pixel=False)
    assert r1["frags"] > 40000
    assert_bit_exact(r1, "teapot shadow pass")
    md = PipelineDesc.make(Shader.PHONG_SHADOW, Xform.MVP, depth_mode=DepthMode.LESS_WRITE,
                           pixel_format=PixelFormat.BGRA8)
    r2 = run_pair(gpu_ctx, md, sc.uniforms, sc.mesh, sc.width, sc.height, shadow=r1["o_depth"])
    assert r2["frags"] > 0.2 * sc.width * sc.height
    assert_bit_exact(r2, f"teapot colour pass {wh}")


@pytest.mark.parametrize("frame", [0, 300, 1234])
def test_texture_mapping_example(gpu_ctx, frame):
    tex = io.load_png(os.path.join(DATA, "rust.png"))
    sc = S.texture_cube_scene(tex, frame=frame)
    d = PipelineDesc.make(Shader.TEXTURED, Xform.MVP)
    r = run_pair(gpu_ctx, d, sc.uniforms, sc.mesh, sc.width, sc.height, texture=tex, color_init=180, depth=False)
    assert r["frags"] > 10000
    assert_bit_exact(r, f"texture_mapping frame {frame}")


def test_checkerboard_nearest(gpu_ctx):
    tex = io.load_png(os.path.join(DATA, "checkerboard.png"))
    sc = S.texture_cube_scene(tex, frame=77, width=320, height=240)
    for sampler in (Sampler.NEAREST, Sampler.NEAREST_MIRRORED):
        d = PipelineDesc.make(Shader.TEXTURED, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE,
                              sampler=sampler)
        r = run_pair(gpu_ctx, d, sc.uniforms, sc.mesh, sc.width, sc.height, texture=tex, color_init=180)
        assert_bit_exact(r, f"checkerboard {sampler.name}")
