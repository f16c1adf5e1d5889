"""Bin storage on the GPU: tile bins of bounded size, per-tile overflow segments,
growth from the raster pass's own report (no host wait), and frames that need
more overflow storage than they were given.  Every frame is compared with the
oracle, bit-exact."""
import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd._lib import EuchipError
from euc_amd.pipeline import BandPolicy, CullMode, DepthMode, PipelineDesc, Shader, Xform
from euc_amd.renderer import Context
from oracle import oracle as O

from parity import assert_same

pytestmark = pytest.mark.gpu

DESC = PipelineDesc.make(Shader.FLAT, Xform.CLIP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE,
                         band_policy=BandPolicy.SEQ)
CLEAR = 0x01020304


def blob_mesh(seed, n_tris, center=(0.0, 0.0), spread=0.1, size=0.02):
    """n_tris small triangles scattered around `center` (clip space, w = 1)."""
    rng = np.random.default_rng(seed)
    c = rng.normal(0.0, spread, (n_tris, 2)) + np.asarray(center)
    pos = np.empty((n_tris, 3, 4), np.float32)
    pos[:, :, :2] = c[:, None, :] + rng.uniform(-size, size, (n_tris, 3, 2))
    pos[:, :, 2] = rng.uniform(0.05, 0.95, (n_tris, 1))
    pos[:, :, 3] = 1.0
    rgba8 = rng.integers(0, 2 ** 32, 3 * n_tris, dtype=np.uint64).astype(np.uint32)
    return S.Mesh(pos=pos.reshape(-1, 4), count=3 * n_tris, rgba8=rgba8)


def oracle_frame(mesh, w, h):
    c = np.full((h, w), CLEAR, np.uint32)
    d = np.full((h, w), 1.0, np.float32)
    rc, frags = O.render(DESC, S.uniforms(), mesh, c, d)
    assert rc == 0
    return c, d


def test_async_density_jump():
    """Sparse frames, then frames whose tiles hold thousands of entries (far past
    the bin capacity the context started with), then sparse again, all enqueued
    without a host wait: the dense ones spill into overflow segments and are
    still whole; the context grows its bins from the raster pass's report."""
    w, h = 256, 192
    sparse = blob_mesh(1, 300, spread=0.5)
    dense = blob_mesh(2, 60000, spread=0.15, size=0.03)
    seq = [sparse, sparse, dense, dense, sparse, dense, sparse]
    ctx = Context(0)
    try:
        gm = {id(m): ctx.mesh(m) for m in (sparse, dense)}
        outs = []
        for m in seq:
            c, d = ctx.color_target(w, h), ctx.depth_target(w, h)
            ctx.render(DESC, S.uniforms(), gm[id(m)], c, d, clear_color=CLEAR, clear_depth=1.0, sync=False)
            outs.append((c, d))
        ctx.synchronize()
        ref = {id(m): oracle_frame(m, w, h) for m in (sparse, dense)}
        for i, (m, (c, d)) in enumerate(zip(seq, outs)):
            rc, rd = ref[id(m)]
            assert_same(c.download(), d.download(), rc, rd, f"frame {i} ({'dense' if m is dense else 'sparse'})")
        # a later frame uses the grown capacity and is still exact
        c, d = ctx.color_target(w, h), ctx.depth_target(w, h)
        ctx.render(DESC, S.uniforms(), gm[id(dense)], c, d, clear_color=CLEAR, clear_depth=1.0)
        assert_same(c.download(), d.download(), *ref[id(dense)], "dense frame after growth")
    finally:
        ctx.close()


def test_bins_stay_bounded_on_a_dense_corner():
    """200k triangles packed into a corner of a 1080p target: one tile holds
    tens of thousands of entries.  Bins sized by that tile for every tile would
    take gigabytes; they stay within the bin budget and the rest goes to the
    tile's overflow segments."""
    w, h = 1920, 1080
    mesh = blob_mesh(3, 200000, center=(-0.9, 0.9), spread=0.02, size=0.01)
    ctx = Context(0)
    try:
        gm = ctx.mesh(mesh)
        c, d = ctx.color_target(w, h), ctx.depth_target(w, h)
        for _ in range(3):
            ctx.render(DESC, S.uniforms(), gm, c, d, clear_color=CLEAR, clear_depth=1.0)
        held = ctx.bytes_in_use()
        assert held < 400 << 20, held
        rc, rd = oracle_frame(mesh, w, h)
        assert_same(c.download(), d.download(), rc, rd, "dense corner")
    finally:
        ctx.close()


def wide_mesh(seed, n_tris, w, h):
    """Triangles about 40 x 12 pixels over a small target: each covers about
    four 32 x 8 tiles, so every tile holds a large share of all triangles."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(-0.8, 0.8, n_tris)
    cy = rng.uniform(-0.8, 0.8, n_tris)
    pos = np.empty((n_tris, 3, 4), np.float32)
    dx, dy = 40.0 / w, 12.0 / h
    pos[:, 0, :2] = np.stack([cx - dx, cy - dy], 1)
    pos[:, 1, :2] = np.stack([cx + dx, cy - dy], 1)
    pos[:, 2, :2] = np.stack([cx, cy + dy], 1)
    pos[:, :, 2] = rng.uniform(0.05, 0.95, (n_tris, 1))
    pos[:, :, 3] = 1.0
    rgba8 = rng.integers(0, 2 ** 32, 3 * n_tris, dtype=np.uint64).astype(np.uint32)
    return S.Mesh(pos=pos.reshape(-1, 4), count=3 * n_tris, rgba8=rgba8)


def test_pool_exhaustion_is_reported_and_recovered():
    """A frame whose overflow entries exceed the pool it was given is not
    written: an asynchronous one is reported by the next synchronize (the later
    frames on the context are unaffected once it has seen the report), and a
    blocking render re-runs it with the pool it asked for."""
    w, h = 64, 64
    mesh = wide_mesh(4, 1200000, w, h)
    ref = oracle_frame(mesh, w, h)
    ctx = Context(0)
    try:
        gm = ctx.mesh(mesh)
        c, d = ctx.color_target(w, h), ctx.depth_target(w, h)
        ctx.render(DESC, S.uniforms(), gm, c, d, clear_color=CLEAR, clear_depth=1.0, sync=False)
        with pytest.raises(EuchipError) as e:
            ctx.synchronize()
        assert e.value.status == 5  # EUCHIP_OVER_BUDGET
        ctx.synchronize()  # reported once
        c.clear(0)
        ctx.render(DESC, S.uniforms(), gm, c, d, clear_color=CLEAR, clear_depth=1.0, sync=False)
        ctx.synchronize()
        assert_same(c.download(), d.download(), *ref, "async frame after the report")
    finally:
        ctx.close()
    ctx = Context(0)
    try:
        gm = ctx.mesh(mesh)
        c, d = ctx.color_target(w, h), ctx.depth_target(w, h)
        ctx.render(DESC, S.uniforms(), gm, c, d, clear_color=CLEAR, clear_depth=1.0)
        assert_same(c.download(), d.download(), *ref, "blocking frame, re-run with a larger pool")
    finally:
        ctx.close()
