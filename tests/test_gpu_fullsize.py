"""The batched-icon and serving workloads at their stated sizes, against the oracle.

  * 4096 icon requests (seeded spheres, tori and boxes of 200-20k log-uniform
    triangles, square viewports 32-256 px, Phong, render_seq bands) in one batch,
    replayed from a hipGraph: a seeded sample of 256 requests equals the oracle
    bit-exact, and every one of the 4096 targets is identical after a second
    replay over garbage-filled targets (the replay rewrites every pixel).
  * Serving at its stated sizes: a 1024^2 target over a 500k-triangle mesh with a
    2048^2 texture, in a Poisson mix with 1k-500k-triangle meshes and 64-1024 px
    targets; that request equals the oracle and, with every other request of the
    mix, the same request rendered alone (tolerance zero).
"""
import concurrent.futures as cf

import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd.pipeline import BandPolicy, CullMode, DepthMode, PipelineDesc, Shader, Xform
from euc_amd.renderer import Context, Request
from euc_amd.serving import ServeItem, Server, run_poisson
from oracle import oracle as O

from parity import assert_same

pytestmark = pytest.mark.gpu

ICON_SEED = 0xC0FFEE00
ICON_CLEAR = 0xFF202020


def oracle_many(jobs, workers=16):
    """jobs: callables returning (colour, depth); run on a thread pool (the oracle
    releases the interpreter lock while it renders)."""
    with cf.ThreadPoolExecutor(workers) as ex:
        return list(ex.map(lambda f: f(), jobs))


def test_icon_batch_4096_vs_oracle():
    n = 4096
    desc = PipelineDesc.make(Shader.PHONG, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.BACK,
                             band_policy=BandPolicy.SEQ)
    scenes = [S.icon_request(ICON_SEED + i) for i in range(n)]
    ctx = Context(0)
    try:
        meshes = [ctx.mesh(sc.mesh) for sc in scenes]
        tg = [(ctx.color_target(sc.width, sc.height), ctx.depth_target(sc.width, sc.height)) for sc in scenes]
        for c, d in tg:
            c.clear(0x5A5A5A5A)
            d.clear(0.5)
        b = ctx.batch().set(desc, [Request(m, c, d, sc.uniforms) for sc, m, (c, d) in zip(scenes, meshes, tg)],
                            clear_color=ICON_CLEAR, clear_depth=1.0)
        b.run(graph=True)
        first = [(c.download(), d.download()) for c, d in tg]
        for c, d in tg:
            c.clear(0x5A5A5A5A)
            d.clear(0.25)
        b.run(graph=True)
        for i, (c, d) in enumerate(tg):
            assert_same(c.download(), d.download(), first[i][0], first[i][1], f"icon {i}, second replay")
        b.close()
    finally:
        ctx.close()
    sample = sorted(np.random.default_rng(0x1C0).choice(n, 256, replace=False).tolist())

    def job(i):
        sc = scenes[i]

        def run():
            c = np.full((sc.height, sc.width), ICON_CLEAR, np.uint32)
            d = np.full((sc.height, sc.width), 1.0, np.float32)
            rc, _ = O.render(desc, sc.uniforms, sc.mesh, c, d)
            assert rc == 0
            return c, d
        return run

    refs = oracle_many([job(i) for i in sample])
    assert sum(int((r[0] != ICON_CLEAR).sum()) for r in refs) > 100000
    for i, (rc, rd) in zip(sample, refs):
        assert_same(first[i][0], first[i][1], rc, rd, f"icon {i} of 4096 vs oracle")


SERVE_DESC = PipelineDesc.make(Shader.TEXTURED, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.BACK,
                               band_policy=BandPolicy.SEQ)


def test_serving_stated_sizes():
    seed = 0x5E7B
    meshes, texs = S.serving_assets(seed, n_meshes=5, n_textures=2, min_tris=1000, max_tris=500000,
                                    tex_sizes=(256, 2048))
    assert max(m.n_triangles for m in meshes) >= 490000 and texs[1].shape[0] == 2048
    big_mesh = int(np.argmax([m.n_triangles for m in meshes]))
    ctx = Context(0)
    try:
        gm = [ctx.mesh(m) for m in meshes]
        gt = [ctx.texture(t) for t in texs]
        items, hosts = [], []
        for i in range(40):
            mi, ti, u, side = S.serving_request(seed * 1000 + i, len(gm), len(gt), min_side=64, max_side=1024)
            if i == 7:  # the largest request: 1024^2 target, 500k triangles, 2048^2 texture
                mi, ti, side = big_mesh, 1, 1024
                _, _, u, _ = S.serving_request(seed * 1000 + i, len(gm), len(gt), min_side=1024, max_side=1024)
            items.append(ServeItem(gm[mi], u, side, side, gt[ti]))
            hosts.append((mi, ti, u, side))
        assert max(it.width for it in items) == 1024 and min(it.width for it in items) < 128

        def alone(it):
            c = ctx.color_target(it.width, it.height)
            d = ctx.depth_target(it.width, it.height)
            ctx.render(SERVE_DESC, it.uniforms, it.mesh, c, d, texture=it.texture, clear_color=0, clear_depth=1.0)
            out = (c.download(), d.download())
            c.close()
            d.close()
            return out

        ref = [alone(it) for it in items]
        srv = Server(ctx, SERVE_DESC, budget_bytes=0, max_batch=16, slots=2)
        res = run_poisson(srv, items, rate_hz=4000.0, seed=7, keep=lambda i: True)
        srv.close()
        assert all(s == 0 for s in res.statuses), res.statuses
        assert res.stats["batches"] >= 3
        for i, (rc, rd) in enumerate(ref):
            gc, gd = res.outputs[i]
            assert_same(gc, gd, rc, rd, f"served request {i} vs alone")
    finally:
        ctx.close()
    mi, ti, u, side = hosts[7]
    pix = np.zeros((side, side), np.uint32)
    dep = np.ones((side, side), np.float32)
    rc, frags = O.render(SERVE_DESC, u, meshes[mi], pix, dep, texture=texs[ti])
    assert rc == 0 and frags > 50000
    gc, gd = res.outputs[7]
    assert_same(gc, gd, pix, dep, "1024^2 / 500k triangles / 2048^2 texture vs oracle")
