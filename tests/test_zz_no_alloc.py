"""euchip_render_noalloc through the C ABI: the same targets as euchip_render
(bit-exact colour and depth) on two calls in a row, the second with a smaller
input, with no device memory taken by the context; inputs beyond the limits of
its static scratch are rejected before anything is enqueued."""
import ctypes as C

import numpy as np
import pytest

from euc_amd import _lib
from euc_amd import scenes as S
from euc_amd.pipeline import CullMode, DepthMode, PipelineDesc, Shader, Xform
from euc_amd.renderer import Context, desc_c

from parity import assert_same

pytestmark = pytest.mark.gpu

GARBAGE_C, GARBAGE_D = 0x5A5A5A5A, np.float32(0.375)


def noalloc(ctx, desc, uniforms, gmesh, pixel, depth, clear_color, clear_depth):
    d = desc_c(desc, None, None, clear_color, clear_depth)
    u = np.ascontiguousarray(uniforms, dtype=np.float32)
    return _lib.lib().euchip_render_noalloc(ctx.handle, C.byref(d), u.ctypes.data, gmesh.handle,
                                            pixel.handle if pixel else None, depth.handle if depth else None)


def both(ctx, desc, sc, clear_color, clear_depth):
    gm = ctx.mesh(sc.mesh)
    w, h = sc.width, sc.height
    ref_c, ref_d = ctx.color_target(w, h), ctx.depth_target(w, h)
    ctx.render(desc, sc.uniforms, gm, ref_c, ref_d, clear_color=clear_color, clear_depth=clear_depth)
    c = ctx.color_target(w, h).upload(np.full((h, w), GARBAGE_C, np.uint32))
    d = ctx.depth_target(w, h).upload(np.full((h, w), GARBAGE_D, np.float32))
    held = ctx.bytes_in_use()
    rc = noalloc(ctx, desc, sc.uniforms, gm, c, d, clear_color, clear_depth)
    assert rc == 0, _lib.lib().euchip_last_error(ctx.handle).decode()
    assert ctx.bytes_in_use() == held
    return (c.download(), d.download()), (ref_c.download(), ref_d.download())


def test_noalloc_matches_render_twice():
    ctx = Context(0)
    try:
        # 1: the 1080p torus frame, Phong (262,144 triangles: the triangle limit)
        sc = S.torus_scene()
        desc = PipelineDesc.make(Shader.PHONG, Xform.MVP, depth_mode=DepthMode.LESS_WRITE)
        (gc, gd), (rc, rd) = both(ctx, desc, sc, 0, 1.0)
        assert (gc != 0).sum() > 100000
        assert_same(gc, gd, rc, rd, "noalloc torus 1080p")
        # 2: a smaller input on the same scratch: 1024 triangles at 640 x 480, flat
        sc = S.plain_scene()
        desc = PipelineDesc.make(Shader.FLAT, Xform.CLIP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.BACK)
        (gc, gd), (rc, rd) = both(ctx, desc, sc, 0x01020304, 1.0)
        assert_same(gc, gd, rc, rd, "noalloc plain 640x480 after the torus")
    finally:
        ctx.close()


def test_noalloc_rejects_inputs_beyond_its_limits():
    ctx = Context(0)
    try:
        sc = S.torus_scene(nu=256, nv=129, instances=4)  # 264,192 triangles
        gm = ctx.mesh(sc.mesh)
        c = ctx.color_target(sc.width, sc.height).upload(np.full((sc.height, sc.width), GARBAGE_C, np.uint32))
        d = ctx.depth_target(sc.width, sc.height)
        desc = PipelineDesc.make(Shader.PHONG, Xform.MVP, depth_mode=DepthMode.LESS_WRITE)
        rc = noalloc(ctx, desc, sc.uniforms, gm, c, d, 0, 1.0)
        assert rc == 5  # EUCHIP_OVER_BUDGET
        assert "at most 262144" in _lib.lib().euchip_last_error(ctx.handle).decode()
        assert np.all(c.download() == GARBAGE_C)  # nothing was enqueued
        big = ctx.color_target(4096, 4096)
        rc = noalloc(ctx, PipelineDesc.make(Shader.FLAT, Xform.CLIP), S.uniforms(), ctx.mesh(S.plain_scene().mesh),
                     big, None, 0, None)
        assert rc == 5 and "tiles" in _lib.lib().euchip_last_error(ctx.handle).decode()
    finally:
        ctx.close()
