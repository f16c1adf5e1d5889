"""Big-triangle mode on the GPU against the oracle, bit-exact: meshes with no more
triangles than the target has tiles, whose triangles cover many tiles each.

Their rows are walked once by the span blocks of the setup launch and resumed
in every tile (triangles.rs:257-301 additions, unchanged); these cases force the
other paths that mode has:
  * large checkpoint frames (2048^2 with triangles spanning the target), and
    frames whose checkpoint tables would pass their limits, whose big triangles
    go on the big list and are walked from the row starts;
  * every tile shape the context takes for big-triangle frames: 64x4, 128x2 and
    256x1 tiles use the checkpoints with 4-, 2- and 1-row masks, 8x32 and 16x16
    tiles (taller than 8 rows) walk from the row starts;
  * every depth mode, both band policies and MSAA through the checkpoint path;
  * occlusion: entries behind a tile's full-cover triangle are skipped, so
    depths within a few ulps of each other (and exact ties), depths that leave
    the clip range and vertices far outside the target (wide margins) must
    still give the oracle's pixels.
"""
import os

import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd.pipeline import (AaMode, BandPolicy, CoordinateMode, CullMode, DepthMode, PipelineDesc, PixelMode,
                              Shader, Xform)
from tests.parity import assert_bit_exact, run_pair

pytestmark = pytest.mark.gpu

CK_MAX_SLOTS = 1 << 26   # checkpoint slots a frame may take (one per triangle, tile column and row)
CKTAB_MAX = 8 << 20      # tile entries a frame may take (one per tile and triangle)


def corner_triangles(seed, n):
    """n triangles with one vertex near each of three screen corners (boxes that
    cover most of the target), z in [0, 1], w = 1, one flat colour each."""
    rng = np.random.default_rng(seed)
    corners = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]], np.float64)
    pos = np.empty((3 * n, 4), np.float32)
    for t in range(n):
        pick = rng.permutation(4)[:3]
        for i, c in enumerate(pick):
            pos[3 * t + i, :2] = corners[c] * rng.uniform(0.85, 1.15, 2)
        pos[3 * t:3 * t + 3, 2] = rng.uniform(0.0, 1.0, 3)
    pos[:, 3] = 1.0
    rgba8 = np.repeat(rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) | np.uint32(0xFF000000), 3)
    return S.Mesh(pos=pos, count=3 * n, rgba8=rgba8)


def mid_triangles(seed, n, half=0.2):
    """n triangles whose vertices lie within `half` (NDC) of a random centre
    (boxes of about 100 tiles at 1024^2), z in [0, 1], w = 1, one flat colour each."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.8, 0.8, (n, 1, 2))
    pos = np.empty((n, 3, 4), np.float32)
    pos[:, :, :2] = c + rng.uniform(-half, half, (n, 3, 2))
    pos[:, :, 2] = rng.uniform(0.0, 1.0, (n, 3))
    pos[:, :, 3] = 1.0
    rgba8 = np.repeat(rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) | np.uint32(0xFF000000), 3)
    return S.Mesh(pos=pos.reshape(-1, 4), count=3 * n, rgba8=rgba8)


def checkpoint_on(n_tris, w, h, sx=5):
    """Whether a big-triangle frame keeps row checkpoints: its slots (triangles x
    tile columns x rows) and tile entries (tiles x triangles, rounded to 4)
    within their limits."""
    ntx, nty = -(-w // (1 << sx)), -(-h // (256 >> sx))
    return n_tris * ntx * h <= CK_MAX_SLOTS and ntx * nty * ((n_tris + 3) & ~3) <= CKTAB_MAX


def flat(dm=DepthMode.LESS_WRITE, bands=BandPolicy.PAR, aa=AaMode.NONE):
    return PipelineDesc.make(Shader.FLAT, Xform.CLIP, PixelMode.WRITE, dm, CoordinateMode.VULKAN, CullMode.NONE, aa,
                             bands)


def test_large_checkpoint_frame():
    from euc_amd.renderer import Context
    w = h = 2048
    mesh = corner_triangles(7, 28)
    assert checkpoint_on(28, w, h)
    ctx = Context(0)
    try:
        for frame in range(2):
            r = run_pair(ctx, flat(), S.uniforms(), mesh, w, h, color_init=0xFF000000)
            assert r["frags"] > 10 ** 6
            assert_bit_exact(r, f"2048^2 corner triangles, frame {frame}")
    finally:
        ctx.close()


def test_big_list_without_checkpoints():
    # 2,100 triangles of about 100 tiles each on 1024^2 (4,096 tiles): big-triangle
    # mode, but 4,096 x 2,100 tile entries pass their limit, so the big
    # triangles go on the big list and every tile walks them from the row starts
    from euc_amd.renderer import Context
    w = h = 1024
    mesh = mid_triangles(17, 2100)
    assert not checkpoint_on(2100, w, h)
    ctx = Context(0)
    try:
        r = run_pair(ctx, flat(), S.uniforms(), mesh, w, h, color_init=0xFF000000)
        assert r["frags"] > 10 ** 6
        assert_bit_exact(r, "1024^2, 2100 mid-size triangles from the big list")
    finally:
        ctx.close()


@pytest.mark.parametrize("sx", [3, 4, 6, 7, 8])
def test_big_tile_shapes(sx):
    from euc_amd.renderer import Context
    plain = S.plain_scene()
    old = os.environ.get("EUCHIP_BIG_TILE_SX")
    os.environ["EUCHIP_BIG_TILE_SX"] = str(sx)
    try:
        ctx = Context(0)
    finally:
        if old is None:
            del os.environ["EUCHIP_BIG_TILE_SX"]
        else:
            os.environ["EUCHIP_BIG_TILE_SX"] = old
    try:
        for bands in (BandPolicy.PAR, BandPolicy.SEQ):
            r = run_pair(ctx, flat(bands=bands), plain.uniforms, plain.mesh, 640, 480, color_init=0x11223344)
            assert r["frags"] > 10 ** 5
            assert_bit_exact(r, f"tiles {1 << sx}x{256 >> sx} {bands.name}")
    finally:
        ctx.close()


@pytest.mark.parametrize("dm", [DepthMode.LESS_WRITE, DepthMode.GREATER_WRITE, DepthMode.LESS_PASS, DepthMode.NONE,
                                DepthMode(2, True)])
def test_checkpoints_depth_modes(gpu_ctx, dm):
    mesh = corner_triangles(11, 40)
    init = 0.5 if dm.test in (2, 3) else 1.0
    r = run_pair(gpu_ctx, flat(dm=dm), S.uniforms(), mesh, 700, 500, depth_init=init)
    assert r["frags"] > 10 ** 5 or dm.test == 2
    assert_bit_exact(r, f"corner triangles {dm}")


@pytest.mark.parametrize("level", [1, 2])
def test_checkpoints_msaa(gpu_ctx, level):
    mesh = corner_triangles(13, 30)
    d = PipelineDesc.make(Shader.COLOR, Xform.CLIP, PixelMode.WRITE, DepthMode.LESS_WRITE, CoordinateMode.VULKAN,
                          CullMode.NONE, AaMode.Msaa(level), BandPolicy.PAR)
    m = S.Mesh(pos=mesh.pos, count=mesh.count, col=np.random.default_rng(level).uniform(0, 1, (mesh.count, 4))
               .astype(np.float32))
    r = run_pair(gpu_ctx, d, S.uniforms(), m, 1600, 1000)
    assert r["frags"] > 10 ** 5
    assert_bit_exact(r, f"corner triangles msaa {level}")


def test_checkpoint_frames_among_other_frames():
    # Checkpoint frames count their tiles' lists into the half of the count
    # buffer their frame parity selects; the next big-triangle frame's raster
    # clears the other half, and a half left over when other frames came between
    # is cleared before it is counted into again.  Frames of every kind in one
    # context, each bit-exact: checkpoint frames of both parities, binned frames
    # between them (one, two, three in a row), and a larger target that starts
    # the count buffers over.
    from euc_amd.renderer import Context
    plain = S.plain_scene()
    binned = S.icon_request(0x8A1, min_tris=6000, max_tris=8000)  # more triangles than tiles: binned
    corner = corner_triangles(23, 20)
    seq = ["ck", "bin", "ck", "ck", "bin", "bin", "ck", "bin", "bin", "bin", "ck", "ck", "big", "ck", "bin", "ck"]
    ctx = Context(0)
    try:
        for i, kind in enumerate(seq):
            if kind == "ck":
                r = run_pair(ctx, flat(), plain.uniforms, plain.mesh, 640, 480, color_init=0x11223344)
            elif kind == "big":
                r = run_pair(ctx, flat(), S.uniforms(), corner, 2048, 2048, color_init=0xFF000000)
            else:
                d = PipelineDesc.make(Shader.PHONG, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.BACK)
                r = run_pair(ctx, d, binned.uniforms, binned.mesh, binned.width, binned.height)
            assert_bit_exact(r, f"frame {i} ({kind})")
    finally:
        ctx.close()


def tie_triangles(seed, n, z0=0.5, spread=4, slope=True):
    """n corner triangles (boxes over most of the target) whose depths lie within
    `spread` ulps of z0 at every vertex: planes that tie exactly, or cross one
    another by a few ulps, so that the winners of neighbouring pixels differ and
    every occlusion bound sits next to the depths it has to separate."""
    rng = np.random.default_rng(seed)
    m = corner_triangles(seed, n)
    pos = m.pos.copy()
    base = np.float32(z0)
    steps = rng.integers(0, spread + 1, (n, 1 if not slope else 3))
    z = np.full((n, 3), base, np.float32)
    for k in range(spread):
        z = np.where(np.broadcast_to(steps, (n, 3)) > k, np.nextafter(z, np.float32(2.0)), z)
    z[rng.random(n) < 0.25] = base  # exact ties (order breaks them)
    pos[:, 2] = z.reshape(-1)
    return S.Mesh(pos=pos, count=m.count, rgba8=m.rgba8)


@pytest.mark.parametrize("dm", [DepthMode.LESS_WRITE, DepthMode.GREATER_WRITE])
@pytest.mark.parametrize("slope", [False, True])
def test_occlusion_bounds_near_ties(gpu_ctx, dm, slope):
    # Entries behind a tile's full-cover triangle are skipped in the raster; the
    # bounds carry margins for the walk's roundings, so triangles within a few
    # ulps of the occluder (or tied with it) must all stay
    mesh = tie_triangles(31 + slope, 48, slope=slope)
    init = 0.0 if dm.test == 3 else 1.0
    r = run_pair(gpu_ctx, flat(dm=dm), S.uniforms(), mesh, 700, 500, depth_init=init)
    assert_bit_exact(r, f"ulp-close corner triangles {dm} slope={slope}")
    assert r["frags"] > 5 * 10 ** 5


def test_occlusion_bounds_clip_and_large_weights(gpu_ctx):
    # occluders whose depth leaves [0, 1] somewhere (z clip) and triangles with
    # vertices far outside the target (large weights, wide margins)
    rng = np.random.default_rng(41)
    m = corner_triangles(41, 40)
    pos = m.pos.copy()
    pos[:, 2] = rng.uniform(-0.4, 1.4, len(pos)).astype(np.float32)
    far = rng.random(40) < 0.5
    pos.reshape(40, 3, 4)[far, :, :2] *= np.float32(40.0)
    mesh = S.Mesh(pos=pos, count=m.count, rgba8=m.rgba8)
    r = run_pair(gpu_ctx, flat(), S.uniforms(), mesh, 640, 480)
    assert r["frags"] > 10 ** 5
    assert_bit_exact(r, "clipped and far corner triangles")
