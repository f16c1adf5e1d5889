"""The C++ host mirror (include/euchip.hpp) on the GPU: tests/cpp/pipeline_check
renders each case through euc::Pipeline::render the way a euc user would; its
targets must equal the oracle's bit for bit, and reference panics must come back
with the reference's message."""
import os
import subprocess

import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd.pipeline import (AaMode, BandPolicy, CoordinateMode, CullMode, DepthMode, Geometry, PipelineDesc,
                              PixelFormat, PixelMode, Primitives, Shader, Xform)
from oracle import oracle as O

from parity import assert_same

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "pipeline_check")

DEPTH = {"none": DepthMode.NONE, "less_write": DepthMode.LESS_WRITE, "greater_write": DepthMode.GREATER_WRITE,
         "less_pass": DepthMode.LESS_PASS, "greater_pass": DepthMode.GREATER_PASS}
COORDS = {"vulkan": CoordinateMode.VULKAN, "opengl": CoordinateMode.OPENGL, "metal": CoordinateMode.METAL,
          "directx": CoordinateMode.DIRECTX, "vulkan_noclip": CoordinateMode.VULKAN.without_z_clip()}
CULL = {"none": CullMode.NONE, "back": CullMode.BACK, "front": CullMode.FRONT}


def random_clip_tris(seed, n):
    rng = np.random.default_rng(seed)
    pos = np.empty((3 * n, 4), np.float32)
    pos[:, :2] = rng.uniform(-1.2, 1.2, (3 * n, 2))
    pos[:, 2] = rng.uniform(-0.1, 1.1, 3 * n)
    pos[:, 3] = rng.uniform(0.8, 1.4, 3 * n)
    col = rng.uniform(0.0, 1.0, (3 * n, 4)).astype(np.float32)
    rgba8 = rng.integers(0, 2 ** 32, 3 * n, dtype=np.uint64).astype(np.uint32)
    return pos, col, rgba8


CASES = [
    # name, shader, depth_mode, coords, cull, msaa, bands, pixel_write, (pw, ph), (dw, dh), pixel, depth
    ("flat_less", Shader.FLAT, "less_write", "vulkan", "back", -1, "par", True, (320, 200), (320, 200), 1, 1),
    ("color_greater_opengl", Shader.COLOR, "greater_write", "opengl", "none", -1, "seq", True, (96, 64), (96, 64), 1, 1),
    ("color_msaa2_front", Shader.COLOR, "less_write", "vulkan_noclip", "front", 2, "seq", True, (80, 72), (80, 72), 1, 1),
    ("depth_only_empty_pixel", Shader.FLAT, "less_write", "metal", "none", -1, "seq", False, (0, 0), (64, 48), 0, 1),
    ("pixel_only_empty_depth", Shader.COLOR, "none", "directx", "back", -1, "seq", True, (64, 48), (0, 0), 1, 0),
    ("size_mismatch_panics", Shader.FLAT, "less_write", "vulkan", "back", -1, "seq", True, (64, 48), (32, 48), 1, 1),
    ("band_zero_panics", Shader.FLAT, "less_write", "vulkan", "back", -1, "par", True, (30000, 2), (30000, 2), 1, 1),
    # LineList, and a LineTriangleList through the both-sides geometry stage with the BGRA layout
    ("line_list", Shader.COLOR, "less_write", "vulkan", "none", -1, "seq", True, (90, 70), (90, 70), 1, 1,
     {"prims": Primitives.LINE_LIST}),
    ("line_triangles_bgra", Shader.COLOR, "none", "opengl", "back", 1, "seq", True, (90, 70), (90, 70), 1, 1,
     {"prims": Primitives.LINE_TRIANGLE_LIST, "geometry": Geometry.BOTH_SIDES, "pixel_format": PixelFormat.BGRA8}),
    ("triangles_reversed_bgra", Shader.COLOR, "less_write", "vulkan", "back", -1, "seq", True, (90, 70), (90, 70),
     1, 1, {"geometry": Geometry.REVERSED, "pixel_format": PixelFormat.BGRA8}),
]


def write_case(d, seed, case):
    name, sh, dm, cm, cull, msaa, bands, pw_, (pw, ph), (dw, dh), has_p, has_d = case[:12]
    extra = case[12] if len(case) > 12 else {}
    pos, col, rgba8 = random_clip_tris(seed, 40)
    os.makedirs(d, exist_ok=True)
    pos.tofile(os.path.join(d, "pos.f32"))
    col.tofile(os.path.join(d, "col.f32"))
    rgba8.tofile(os.path.join(d, "rgba8.u32"))
    S.uniforms().astype(np.float32).tofile(os.path.join(d, "uniforms.f32"))
    kv = {"shader": int(sh), "clip": 1, "depth_mode": dm, "pixel_write": int(pw_), "coords": cm, "cull": cull,
          "msaa": msaa, "bands": bands, "pos_comps": 4, "pw": pw, "ph": ph, "dw": dw, "dh": dh,
          "pixel": has_p, "depth": has_d, "clear_color": "0x11223344", "clear_depth": "0.75",
          **{k: int(v) for k, v in extra.items()}}
    with open(os.path.join(d, "case.txt"), "w") as f:
        for k, v in kv.items():
            f.write(f"{k} {v}\n")
    aa = AaMode.NONE if msaa < 0 else AaMode.Msaa(msaa)
    desc = PipelineDesc.make(sh, Xform.CLIP, PixelMode.WRITE if pw_ else PixelMode.PASS, DEPTH[dm], COORDS[cm],
                             CULL[cull], aa, BandPolicy.SEQ if bands == "seq" else BandPolicy.PAR, 8,
                             pixel_format=extra.get("pixel_format", PixelFormat.RGBA8),
                             geometry=extra.get("geometry", Geometry.IDENTITY),
                             primitives=extra.get("prims", Primitives.TRIANGLE_LIST))
    mesh = S.Mesh(pos=pos, count=pos.shape[0], col=col, rgba8=rgba8)
    return desc, mesh


def test_cpp_mirror_matches_oracle(tmp_path):
    assert os.path.exists(BIN), f"{BIN} is missing: build it with `python -m euc_amd.build`"
    dirs, expect = [], []
    for k, case in enumerate(CASES):
        d = str(tmp_path / case[0])
        desc, mesh = write_case(d, 0xC0DE + k, case)
        (pw, ph), (dw, dh), has_p, has_d = case[8], case[9], case[10], case[11]
        pix = np.full((ph, pw), 0x11223344, np.uint32) if has_p else None
        dep = np.full((dh, dw), 0.75, np.float32) if has_d else None
        rc, _ = O.render(desc, S.uniforms(), mesh, pix, dep)
        dirs.append(d)
        expect.append((rc, pix, dep))
    out = subprocess.run([BIN, *dirs], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    for case, d, (rc, pix, dep) in zip(CASES, dirs, expect):
        panic = os.path.join(d, "panic.txt")
        if rc == 1:
            assert open(panic).read() == ("assertion `left == right` failed: Pixel target size is compatible "
                                          "with depth target size"), case[0]
        elif rc == 2:
            assert open(panic).read() == "attempt to divide by zero", case[0]
        else:
            assert rc == 0 and not os.path.exists(panic), case[0]
        if pix is not None:
            g = np.fromfile(os.path.join(d, "pixel.u32"), np.uint32).reshape(pix.shape)
            assert np.array_equal(g, pix), f"{case[0]}: colour differs from the oracle"
        if dep is not None:
            g = np.fromfile(os.path.join(d, "depth.f32"), np.float32).reshape(dep.shape)
            assert np.array_equal(g.view(np.uint32), dep.view(np.uint32)), f"{case[0]}: depth differs from the oracle"
