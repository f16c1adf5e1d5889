"""Serving loop on the GPU (euchip_server_*): requests arriving as a Poisson
stream are batched continuously under a device-memory budget; every result
must be the same as that request rendered alone (tolerance: zero -- colour and
depth bit-exact), whatever else shared its batch."""
import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd.pipeline import BandPolicy, CullMode, DepthMode, PipelineDesc, Shader, Xform
from euc_amd.renderer import Context
from euc_amd.serving import ServeItem, Server, run_poisson
from oracle import oracle as O

from parity import assert_same

pytestmark = pytest.mark.gpu

DESC = PipelineDesc.make(Shader.TEXTURED, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.BACK,
                         band_policy=BandPolicy.SEQ)


def setup_ctx(n_req, max_tris=30000, max_side=256, seed=0x5E77):
    ctx = Context(0)
    meshes, texs = S.serving_assets(seed, n_meshes=6, n_textures=3, max_tris=max_tris, tex_sizes=(64, 128, 256))
    gm = [ctx.mesh(m) for m in meshes]
    gt = [ctx.texture(t) for t in texs]
    items, hosts = [], []
    for i in range(n_req):
        mi, ti, u, side = S.serving_request(seed * 1000 + i, len(gm), len(gt), max_side=max_side)
        items.append(ServeItem(gm[mi], u, side, side, gt[ti]))
        hosts.append((meshes[mi], texs[ti], u, side))
    return ctx, items, hosts


def alone(ctx, it):
    c = ctx.color_target(it.width, it.height)
    d = ctx.depth_target(it.width, it.height)
    ctx.render(DESC, it.uniforms, it.mesh, c, d, texture=it.texture, clear_color=0, clear_depth=1.0)
    out = (c.download(), d.download())
    c.close()
    d.close()
    return out


def test_serving_alone_vs_mix():
    ctx, items, _ = setup_ctx(48)
    try:
        ref = [alone(ctx, it) for it in items]
        srv = Server(ctx, DESC, budget_bytes=0, max_batch=8, slots=2)
        res = run_poisson(srv, items, rate_hz=20000.0, seed=3, keep=lambda i: True)
        assert all(s == 0 for s in res.statuses)
        assert res.stats["batches"] >= 2 and res.stats["completed"] == len(items)
        for i, (rc, rd) in enumerate(ref):
            gc, gd = res.outputs[i]
            assert_same(gc, gd, rc, rd, f"request {i} in the mix")
        srv.close()
    finally:
        ctx.close()


def test_serving_matches_oracle():
    ctx, items, hosts = setup_ctx(5, max_tris=4000, max_side=96, seed=0x5E78)
    try:
        srv = Server(ctx, DESC, max_batch=4, slots=2)
        res = run_poisson(srv, items, rate_hz=5000.0, seed=4, keep=lambda i: True)
        for i, (mesh, tex, u, side) in enumerate(hosts):
            pix = np.zeros((side, side), np.uint32)
            dep = np.ones((side, side), np.float32)
            rc, _ = O.render(DESC, u, mesh, pix, dep, texture=tex)
            assert rc == 0
            gc, gd = res.outputs[i]
            assert_same(gc, gd, pix, dep, f"served request {i} vs oracle")
        srv.close()
    finally:
        ctx.close()


def test_serving_budget_admission():
    ctx, items, _ = setup_ctx(24, seed=0x5E79)
    try:
        ref = [alone(ctx, it) for it in items]
        resident = ctx.bytes_in_use()
        # room for the fixed batch tables plus a few requests at a time
        budget = resident + 2 * (8 << 20) + 6 * (256 * 256 * 8) + (8 << 20)
        srv = Server(ctx, DESC, budget_bytes=budget, max_batch=16, slots=2)
        res = run_poisson(srv, items, rate_hz=50000.0, seed=5, keep=lambda i: True)
        st = res.stats
        assert all(s == 0 for s in res.statuses), [(i, s) for i, s in enumerate(res.statuses) if s != 0]
        assert st["peak_bytes"] <= budget
        assert st["batches"] >= 3  # the budget split the burst
        for i, (rc, rd) in enumerate(ref):
            gc, gd = res.outputs[i]
            assert_same(gc, gd, rc, rd, f"request {i} under a budget")
        srv.close()
    finally:
        ctx.set_budget(0)
        ctx.close()


def test_serving_rejects_what_cannot_fit():
    ctx, items, _ = setup_ctx(6, seed=0x5E7A)
    try:
        resident = ctx.bytes_in_use()
        budget = resident + 2 * (8 << 20) + 4 * (256 * 256 * 8) + (8 << 20)
        huge = ServeItem(items[0].mesh, items[0].uniforms, 4096, 4096, items[0].texture)  # 128 MiB of targets
        srv = Server(ctx, DESC, budget_bytes=budget, max_batch=4, slots=2)
        res = run_poisson(srv, items[:3] + [huge] + items[3:], rate_hz=10000.0, seed=6)
        assert res.statuses[3] == 5  # EUCHIP_OVER_BUDGET
        assert all(s == 0 for k, s in enumerate(res.statuses) if k != 3)
        assert res.stats["rejected"] == 1
        srv.close()
    finally:
        ctx.set_budget(0)
        ctx.close()


def test_serving_per_request_descs():
    # Requests with pipelines of their own beside the server's: every result equals
    # the request rendered alone with its pipeline, whatever shared its batch.
    import time
    ctx, items, _ = setup_ctx(30, max_tris=12000)
    own = [None,
           PipelineDesc.make(Shader.PHONG, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.BACK),
           PipelineDesc.make(Shader.GOURAUD, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE),
           PipelineDesc.make(Shader.TEXTURED, Xform.MVP, depth_mode=DepthMode.GREATER_WRITE, cull=CullMode.BACK,
                             band_policy=BandPolicy.SEQ)]
    clears = [1.0, 1.0, 1.0, 0.0]  # GREATER_WRITE starts from depth 0
    try:
        picks = [i % len(own) for i in range(len(items))]
        ref = []
        for it, k in zip(items, picks):
            c, d = ctx.color_target(it.width, it.height), ctx.depth_target(it.width, it.height)
            ctx.render(own[k] or DESC, it.uniforms, it.mesh, c, d, texture=it.texture, clear_color=0,
                       clear_depth=clears[k])
            ref.append((c.download(), d.download()))
            c.close(), d.close()
        srv = Server(ctx, DESC, budget_bytes=0, max_batch=12, slots=2)
        tickets = [srv.submit(it.mesh, it.uniforms, it.width, it.height, texture=it.texture, desc=own[k],
                              clear_color=0, clear_depth=clears[k]) for it, k in zip(items, picks)]
        got, t0 = {}, time.time()
        while len(got) < len(tickets):
            for cpl in srv.poll():
                assert cpl.status == 0
                got[cpl.ticket] = (cpl.pixel.download(), cpl.depth.download())
                srv.release(cpl.ticket)
            assert time.time() - t0 < 60, "requests still pending"
        assert srv.stats()["batches"] >= 3
        for i, t in enumerate(tickets):
            assert_same(got[t][0], got[t][1], ref[i][0], ref[i][1], f"request {i} ({own[picks[i]] is not None})")
        bad = PipelineDesc.make(Shader.PHONG, Xform.MVP)
        with pytest.raises(Exception):  # no clears: pooled targets would leak earlier results
            srv.submit(items[0].mesh, items[0].uniforms, 8, 8, desc=bad, clear_color=None, clear_depth=None)
        srv.close()
    finally:
        ctx.close()
