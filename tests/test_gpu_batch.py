"""Batched rendering on the GPU: a batch of requests must leave every target
exactly as euchip_render of that request alone does (bit-exact colour and
depth), for every tile width, for eager launches and hipGraph replays, and
through bin growth.  Small cases are also checked against the oracle."""
import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd.pipeline import AaMode, BandPolicy, CullMode, DepthMode, PipelineDesc, Shader, Xform
from euc_amd.renderer import EuchipError, Request, estimate_bytes
from oracle import oracle as O

from parity import assert_same

pytestmark = pytest.mark.gpu

GARBAGE_C, GARBAGE_D = 0x5A5A5A5A, np.float32(0.123)


def icon_desc(band=BandPolicy.SEQ, shader=Shader.PHONG, aa=AaMode.NONE, depth=DepthMode.LESS_WRITE):
    return PipelineDesc.make(shader, Xform.MVP, depth_mode=depth, cull=CullMode.BACK, aa_mode=aa, band_policy=band)


def make_requests(ctx, seeds, **kw):
    scenes = [S.icon_request(s, **kw) for s in seeds]
    meshes = [ctx.mesh(sc.mesh) for sc in scenes]
    return scenes, meshes


def targets(ctx, scenes):
    out = []
    for sc in scenes:
        c = ctx.color_target(sc.width, sc.height).upload(np.full((sc.height, sc.width), GARBAGE_C, np.uint32))
        d = ctx.depth_target(sc.width, sc.height).upload(np.full((sc.height, sc.width), GARBAGE_D, np.float32))
        out.append((c, d))
    return out


def singles(ctx, desc, scenes, meshes, clear_color=0xFF202020, clear_depth=1.0, texture=None):
    res = []
    for sc, m, (c, d) in zip(scenes, meshes, targets(ctx, scenes)):
        ctx.render(desc, sc.uniforms, m, c, d, texture=texture, clear_color=clear_color, clear_depth=clear_depth)
        res.append((c.download(), d.download()))
    return res


def check_batch(ctx, desc, scenes, meshes, ref, tile_w=32, graph=False, clear_color=0xFF202020, clear_depth=1.0,
                replays=1, texture=None):
    tg = targets(ctx, scenes)
    reqs = [Request(m, c, d, sc.uniforms) for sc, m, (c, d) in zip(scenes, meshes, tg)]
    b = ctx.batch().set(desc, reqs, texture=texture, clear_color=clear_color, clear_depth=clear_depth)
    b.set_tile_width(tile_w)
    for rep in range(replays):
        if rep:
            for (c, d), sc in zip(tg, scenes):  # stale results must not survive a replay
                c.upload(np.full((sc.height, sc.width), GARBAGE_C, np.uint32))
                d.upload(np.full((sc.height, sc.width), GARBAGE_D, np.float32))
        b.run(graph=graph)
        for i, ((c, d), (rc, rd)) in enumerate(zip(tg, ref)):
            assert_same(c.download(), d.download(), rc, rd, f"request {i} tile {tile_w} replay {rep}")
    b.close()


@pytest.fixture(scope="module")
def icon_set(gpu_ctx):
    scenes, meshes = make_requests(gpu_ctx, range(0x1C0, 0x1C0 + 20))
    return scenes, meshes


@pytest.mark.parametrize("tile_w", [8, 16, 32])
def test_batch_equals_single_renders(gpu_ctx, icon_set, tile_w):
    scenes, meshes = icon_set
    desc = icon_desc()
    ref = singles(gpu_ctx, desc, scenes, meshes)
    check_batch(gpu_ctx, desc, scenes, meshes, ref, tile_w=tile_w)


def test_batch_graph_replays(gpu_ctx, icon_set):
    scenes, meshes = icon_set
    desc = icon_desc()
    ref = singles(gpu_ctx, desc, scenes, meshes)
    check_batch(gpu_ctx, desc, scenes, meshes, ref, tile_w=16, graph=True, replays=4)


def test_batch_matches_oracle(gpu_ctx):
    scenes, meshes = make_requests(gpu_ctx, range(0x2A0, 0x2A6), max_tris=3000, max_side=96)
    desc = icon_desc()
    ref = []
    for sc in scenes:
        pix = np.full((sc.height, sc.width), 0xFF202020, np.uint32)
        dep = np.full((sc.height, sc.width), 1.0, np.float32)
        rc, _ = O.render(desc, sc.uniforms, sc.mesh, pix, dep)
        assert rc == 0
        ref.append((pix, dep))
    check_batch(gpu_ctx, desc, scenes, meshes, ref, tile_w=8)


def test_batch_par_bands_msaa_and_blank_targets(gpu_ctx):
    # render_par: targets with H < 20000 * 2^L / W get no worker threads (left
    # as cleared); larger ones use the band-relative MSAA grid.
    seeds = range(0x3B0, 0x3B8)
    scenes = []
    for i, s in enumerate(seeds):
        sc = S.icon_request(s, max_tris=4000)
        side = [64, 160, 200, 256][i % 4]
        sc = S.Scene(sc.name, sc.mesh, S.uniforms(view_proj=S.f32(S.perspective_fov_lh_zo(0.9, side, side, 0.1, 20.0))),
                     side, side, 0, 1.0)
        scenes.append(sc)
    meshes = [gpu_ctx.mesh(sc.mesh) for sc in scenes]
    desc = icon_desc(band=BandPolicy.PAR, aa=AaMode.Msaa(1))
    ref = singles(gpu_ctx, desc, scenes, meshes)
    check_batch(gpu_ctx, desc, scenes, meshes, ref, tile_w=32)


def test_batch_bin_growth(gpu_ctx):
    # dense meshes shrunk into a corner of a large target: far more triangles per
    # tile than the first bin estimate, so the first launch spills into the
    # overflow list and the bins grow for the next one
    scenes = []
    for s in range(0x4C0, 0x4C3):
        sc = S.icon_request(s, min_tris=15000, max_tris=20000)
        m = sc.mesh
        model = (np.asarray(m.models, np.float64).reshape(4, 4) @ S.scaling(0.12)).astype(np.float32).reshape(1, 16)
        mesh = S.Mesh(pos=m.pos, nrm=m.nrm, idx=m.idx, count=m.count, n_instances=1, models=model)
        vp = S.f32(S.perspective_fov_lh_zo(0.9, 384, 384, 0.1, 20.0))
        scenes.append(S.Scene(sc.name, mesh, S.uniforms(view_proj=vp), 384, 384, 0, 1.0))
    meshes = [gpu_ctx.mesh(sc.mesh) for sc in scenes]
    desc = icon_desc()
    ref = singles(gpu_ctx, desc, scenes, meshes)
    check_batch(gpu_ctx, desc, scenes, meshes, ref, tile_w=32, replays=3)
    check_batch(gpu_ctx, desc, scenes, meshes, ref, tile_w=8, graph=True, replays=2)


def test_batch_textured_per_request(gpu_ctx):
    scenes, meshes = [], []
    texs = []
    for i, s in enumerate(range(0x5D0, 0x5D4)):
        sc = S.icon_request(s, max_tris=3000)
        m = sc.mesh
        uv = np.stack([m.pos[:, 0] * 0.7 + 0.5, m.pos[:, 1] * 0.7 + 0.5], 1).astype(np.float32)
        mesh = S.Mesh(pos=m.pos, nrm=m.nrm, uv=uv, idx=m.idx, count=m.count, n_instances=1, models=m.models)
        scenes.append(S.Scene(sc.name, mesh, sc.uniforms, sc.width, sc.height, 0, 1.0))
        meshes.append(gpu_ctx.mesh(mesh))
        texs.append(gpu_ctx.texture(S.noise_texture(0x77 + i, 64 << (i % 3))))
    desc = icon_desc(shader=Shader.TEXTURED)
    ref = []
    for sc, m, t, (c, d) in zip(scenes, meshes, texs, targets(gpu_ctx, scenes)):
        gpu_ctx.render(desc, sc.uniforms, m, c, d, texture=t, clear_color=0, clear_depth=1.0)
        ref.append((c.download(), d.download()))
    tg = targets(gpu_ctx, scenes)
    reqs = [Request(m, c, d, sc.uniforms, texture=t) for sc, m, t, (c, d) in zip(scenes, meshes, texs, tg)]
    b = gpu_ctx.batch().set(desc, reqs, clear_color=0, clear_depth=1.0)
    b.run()
    for i, ((c, d), (rc, rd)) in enumerate(zip(tg, ref)):
        assert_same(c.download(), d.download(), rc, rd, f"textured request {i}")
    b.close()


def test_batch_autotune_picks_a_width(gpu_ctx, icon_set):
    scenes, meshes = icon_set
    desc = icon_desc()
    ref = singles(gpu_ctx, desc, scenes, meshes)
    tg = targets(gpu_ctx, scenes)
    reqs = [Request(m, c, d, sc.uniforms) for sc, m, (c, d) in zip(scenes, meshes, tg)]
    b = gpu_ctx.batch().set(desc, reqs, clear_color=0xFF202020, clear_depth=1.0)
    w, times = b.autotune(reps=2)
    assert w in (8, 16, 32) and all(t > 0 for t in times.values())
    b.run(graph=True)
    for i, ((c, d), (rc, rd)) in enumerate(zip(tg, ref)):
        assert_same(c.download(), d.download(), rc, rd, f"autotuned request {i}")
    b.close()


def test_batch_request_errors(gpu_ctx):
    sc = S.icon_request(0x600, max_tris=500)
    m = gpu_ctx.mesh(sc.mesh)
    ok = Request(m, gpu_ctx.color_target(32, 32), gpu_ctx.depth_target(32, 32), sc.uniforms)
    bad = Request(m, gpu_ctx.color_target(32, 32), gpu_ctx.depth_target(48, 32), sc.uniforms)
    b = gpu_ctx.batch()
    with pytest.raises(EuchipError) as e:
        b.set(icon_desc(), [ok, bad])
    assert e.value.name == "SIZE_MISMATCH" and "request 1" in str(e.value)
    wide = Request(m, gpu_ctx.color_target(30000, 1), gpu_ctx.depth_target(30000, 1), sc.uniforms)
    with pytest.raises(EuchipError) as e:
        b.set(icon_desc(band=BandPolicy.PAR), [ok, wide])
    assert e.value.name == "BAND_ZERO" and "request 1" in str(e.value)
    with pytest.raises(EuchipError):
        b.set_tile_width(12)
    b.close()


def test_budget_and_accounting(gpu_ctx):
    from euc_amd.renderer import Context
    ctx = Context(0)
    try:
        assert ctx.bytes_in_use() == 0
        c = ctx.color_target(100, 10)
        assert ctx.bytes_in_use() == 4000
        ctx.set_budget(5000)
        with pytest.raises(EuchipError) as e:
            ctx.color_target(100, 10)
        assert e.value.name == "OVER_BUDGET"
        assert ctx.bytes_in_use() == 4000
        ctx.set_budget(0)
        c.close()
        assert ctx.bytes_in_use() == 0
    finally:
        ctx.close()
    sc = S.icon_request(0x601, max_tris=2000)
    m = gpu_ctx.mesh(sc.mesh)
    est = estimate_bytes(icon_desc(), [Request(m, gpu_ctx.color_target(64, 64), gpu_ctx.depth_target(64, 64),
                                               sc.uniforms)])
    assert est > 0


def mixed_descs():
    """Five pipelines, interleaved: shaders, depth modes, culling, MSAA and bands
    differ between them."""
    return [icon_desc(),
            icon_desc(shader=Shader.GOURAUD),
            icon_desc(depth=DepthMode.LESS_PASS),
            icon_desc(band=BandPolicy.PAR, aa=AaMode.Msaa(1)),
            PipelineDesc.make(Shader.PHONG, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE)]


def test_batch_per_request_descs(gpu_ctx):
    kinds = mixed_descs()
    scenes, meshes = make_requests(gpu_ctx, range(0x7A0, 0x7AC), max_tris=6000)
    descs = [kinds[(3 * i) % len(kinds)] for i in range(len(scenes))]
    ref = []
    for d, sc, m, (c, dt) in zip(descs, scenes, meshes, targets(gpu_ctx, scenes)):
        gpu_ctx.render(d, sc.uniforms, m, c, dt, clear_color=0xFF202020, clear_depth=1.0)
        ref.append((c.download(), dt.download()))
    tg = targets(gpu_ctx, scenes)
    reqs = [Request(m, c, d, sc.uniforms) for sc, m, (c, d) in zip(scenes, meshes, tg)]
    b = gpu_ctx.batch().set(descs, reqs, clear_color=0xFF202020, clear_depth=1.0)
    for tile_w, graph in ((32, False), (8, True), (16, True)):
        b.set_tile_width(tile_w)
        for (c, d), sc in zip(tg, scenes):
            c.upload(np.full((sc.height, sc.width), GARBAGE_C, np.uint32))
            d.upload(np.full((sc.height, sc.width), GARBAGE_D, np.float32))
        b.run(graph=graph)
        for i, ((c, d), (rc, rd)) in enumerate(zip(tg, ref)):
            assert_same(c.download(), d.download(), rc, rd, f"mixed request {i} tile {tile_w} graph {graph}")
    w, _ = b.autotune(reps=1)
    assert w in (8, 16, 32)
    # back to one desc for all: the parts go, the batch renders as a plain one
    b.set(kinds[0], reqs, clear_color=0xFF202020, clear_depth=1.0)
    b.run()
    for i, (c, d) in enumerate(tg):
        if descs[i] == kinds[0]:
            assert_same(c.download(), d.download(), ref[i][0], ref[i][1], f"single-desc request {i}")
    b.close()


def test_batch_per_request_descs_match_oracle(gpu_ctx):
    kinds = mixed_descs()
    scenes, meshes = make_requests(gpu_ctx, range(0x7C0, 0x7C5), max_tris=2500, max_side=96)
    ref = []
    for d, sc in zip(kinds, scenes):
        pix = np.full((sc.height, sc.width), 0xFF202020, np.uint32)
        dep = np.full((sc.height, sc.width), 1.0, np.float32)
        rc, _ = O.render(d, sc.uniforms, sc.mesh, pix, dep)
        assert rc == 0
        ref.append((pix, dep))
    tg = targets(gpu_ctx, scenes)
    reqs = [Request(m, c, d, sc.uniforms) for sc, m, (c, d) in zip(scenes, meshes, tg)]
    b = gpu_ctx.batch().set(kinds, reqs, clear_color=0xFF202020, clear_depth=1.0)
    b.run()
    for i, ((c, d), (rc, rd)) in enumerate(zip(tg, ref)):
        assert_same(c.download(), d.download(), rc, rd, f"mixed request {i} vs oracle")
    b.close()


def test_batch_per_request_desc_errors(gpu_ctx):
    sc = S.icon_request(0x610, max_tris=500)
    m = gpu_ctx.mesh(sc.mesh)
    ok = [Request(m, gpu_ctx.color_target(32, 32), gpu_ctx.depth_target(32, 32), sc.uniforms) for _ in range(3)]
    bad = Request(m, gpu_ctx.color_target(32, 32), gpu_ctx.depth_target(48, 32), sc.uniforms)
    b = gpu_ctx.batch()
    with pytest.raises(EuchipError) as e:  # the bad request is the third of its desc, the fourth of the call
        b.set([icon_desc(), icon_desc(shader=Shader.GOURAUD), icon_desc(), icon_desc()], ok + [bad])
    assert e.value.name == "SIZE_MISMATCH" and "request 3" in str(e.value)
    with pytest.raises(ValueError):
        b.set([icon_desc()], ok)
    b.set([icon_desc(), icon_desc(shader=Shader.GOURAUD), icon_desc()], ok)
    b.run()
    b.close()
