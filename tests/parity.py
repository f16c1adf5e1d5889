"""Shared helpers of the parity tests: run one pipeline call through the HIP
path (C ABI) and through the CPU oracle on identical inputs."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def run_pair(ctx, desc, uniforms, mesh, w, h, *, color_init=0, depth_init=1.0, pixel=True, depth=True,
             dw=None, dh=None, texture=None, shadow=None, gmesh=None, fused_clear=True):
    """Render with both paths.  Targets start as color_init / depth_init (the GPU
    side gets them as fused load-op clears when fused_clear, else uploaded).
    Returns dict with oracle/gpu colour and depth arrays, status codes, frags."""
    dw = w if dw is None else dw
    dh = h if dh is None else dh
    oc = np.full((h, w), color_init, dtype=np.uint32) if pixel else None
    od = np.full((dh, dw), depth_init, dtype=np.float32) if depth else None
    if isinstance(color_init, np.ndarray) and pixel:
        oc = np.ascontiguousarray(color_init, dtype=np.uint32).copy()
    if isinstance(depth_init, np.ndarray) and depth:
        od = np.ascontiguousarray(depth_init, dtype=np.float32).copy()
    o_rc, frags = O.render(desc, uniforms, mesh, oc, od, texture=texture, shadow=shadow)

    from euc_amd._lib import EuchipError
    gm = gmesh if gmesh is not None else ctx.mesh(mesh)
    gp = ctx.color_target(w, h) if pixel else None
    gd = ctx.depth_target(dw, dh) if depth else None
    gt = ctx.texture(texture) if texture is not None else None
    gs = ctx.texture(shadow) if shadow is not None else None
    scalar_c = not isinstance(color_init, np.ndarray)
    scalar_d = not isinstance(depth_init, np.ndarray)
    cc = cd = None
    if gp is not None:
        if fused_clear and scalar_c:
            cc = int(color_init)
        else:
            gp.upload(np.full((h, w), color_init, np.uint32) if scalar_c else color_init)
    if gd is not None:
        if fused_clear and scalar_d:
            cd = float(depth_init)
        else:
            gd.upload(np.full((dh, dw), depth_init, np.float32) if scalar_d else depth_init)
    g_rc = 0
    try:
        ctx.render(desc, uniforms, gm, gp, gd, texture=gt, shadow=gs, clear_color=cc, clear_depth=cd)
    except EuchipError as e:
        g_rc = e.status
    return {
        "o_rc": o_rc, "g_rc": g_rc, "frags": frags,
        "o_color": oc, "o_depth": od,
        "g_color": gp.download() if gp is not None else None,
        "g_depth": gd.download() if gd is not None else None,
    }


def assert_bit_exact(r, what=""):
    assert r["o_rc"] == r["g_rc"], f"{what}: status oracle={r['o_rc']} gpu={r['g_rc']}"
    if r["o_color"] is not None:
        bad = r["o_color"] != r["g_color"]
        n = int(bad.sum())
        if n:
            ys, xs = np.nonzero(bad)
            raise AssertionError(f"{what}: {n} colour mismatches, first at (x={xs[0]}, y={ys[0]}): "
                                 f"oracle={r['o_color'][ys[0], xs[0]]:#010x} gpu={r['g_color'][ys[0], xs[0]]:#010x}")
    if r["o_depth"] is not None:
        ob = r["o_depth"].view(np.uint32)
        gb = r["g_depth"].view(np.uint32)
        bad = ob != gb
        n = int(bad.sum())
        if n:
            ys, xs = np.nonzero(bad)
            raise AssertionError(f"{what}: {n} depth mismatches, first at (x={xs[0]}, y={ys[0]}): "
                                 f"oracle={r['o_depth'][ys[0], xs[0]]!r} gpu={r['g_depth'][ys[0], xs[0]]!r}")


def channel_mismatch(r):
    """(max per-channel |difference|, fraction of pixels differing) of the colour targets."""
    a = r["o_color"].view(np.uint8).astype(np.int16)
    b = r["g_color"].view(np.uint8).astype(np.int16)
    d = np.abs(a - b)
    return int(d.max()) if d.size else 0, float((r["o_color"] != r["g_color"]).mean())


def assert_same(g_color, g_depth, r_color, r_depth, what=""):
    """Bit-exact comparison of colour (uint32) and depth (f32 bits) arrays."""
    assert_bit_exact({"o_rc": 0, "g_rc": 0, "o_color": r_color, "g_color": g_color,
                      "o_depth": r_depth, "g_depth": g_depth}, what)
