"""Benchmark of the triangle path on one MI355X.

`--workload torus` (default) measures one 1080p frame per step; `--workload
icons` measures one batch of 4096 icon renders per step (see bench_icons).

Default workload: four 65,536-triangle tori (256 x 128 segments, seeded model
matrices) rendered with per-fragment Phong shading (2 directional lights),
LESS_WRITE depth, back-face culling, into a 1920x1080 RGBA8 colour target and
an f32 depth target, both cleared inside the frame.  One step = one frame.

Reported `value` = shaded fragments per second: the number of emit_fragment
calls euc 0.6.0 makes for this frame (coverage, z-clip and depth test passed,
in submission order), taken from the oracle's run of the same frame in the
cpu_baseline leg, times frames per second of the HIP path.  Inputs are
resident in HBM before the timed region starts.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def measured_traffic(kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by the copy-probe factor), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        k = t["kernels"][kernel]
        return {"bytes": k["fetch_bytes"] + k["write_bytes"], "fetch_bytes": k["fetch_bytes"],
                "write_bytes": k["write_bytes"], "source": "profiles/traffic.json"}
    except (OSError, KeyError, ValueError):
        return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--shader", default="phong", choices=["phong", "gouraud"])
    ap.add_argument("--workload", default="torus", choices=["torus", "plain", "icons", "serving", "heightfield"])
    ap.add_argument("--rate", type=float, default=40000.0, help="Poisson arrival rate of the serving workload (req/s)")
    ap.add_argument("--requests", type=int, default=20000, help="requests of the serving workload")
    ap.add_argument("--budget-gib", type=float, default=16.0, help="device-memory budget of the serving workload")
    ap.add_argument("--icons", type=int, default=4096, help="requests per batch of the icons workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound of the CPU baseline sample")
    ap.add_argument("--check", action="store_true", help="also compare one GPU frame with the oracle")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB read + 256 MiB write elementwise kernel three times "
                         "(the reference point that scales the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def copy_probe():
    import torch
    n = 64 << 20  # 64 Mi f32 = 256 MiB each way
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.add(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "dispatches": 3}))


def algorithmic_bytes(scene) -> dict:
    m = scene.mesh
    mesh_bytes = m.pos.nbytes + m.nrm.nbytes + (m.idx.nbytes if m.idx is not None else 0)
    target_bytes = scene.width * scene.height * 8  # RGBA8 + f32 depth, each written once
    return {"frame": mesh_bytes + target_bytes, "raster": target_bytes, "mesh": mesh_bytes}


def cpu_baseline(scene, desc, seconds: float):
    """The oracle's render_par (row bands, G = 20000*2^L/W, T = min(H/G, cores) threads),
    timed on this host on whole frames until `seconds` have passed (at least 2 frames)."""
    from oracle import oracle as O
    cores = os.cpu_count() or 1
    G = desc.group_rows(scene.width)
    threads = max(1, min(scene.height // max(G, 1), cores))
    desc.threads = threads
    times, frags = [], 0
    t_end = time.perf_counter() + seconds
    first = None
    while True:
        color = np.zeros((scene.height, scene.width), np.uint32)   # clear inside the frame
        depth = np.full((scene.height, scene.width), 1.0, np.float32)
        t0 = time.perf_counter()
        rc, frags = O.render(desc, scene.uniforms, scene.mesh, color, depth)
        times.append(time.perf_counter() - t0)
        assert rc == 0
        if first is None:
            first = (color, depth)
        if len(times) >= 2 and time.perf_counter() > t_end:
            break
    med = float(np.median(times))
    return {"frame_s": med, "frames": len(times), "threads": threads, "frags": frags, "first": first}


ICON_SEED = 0xC0FFEE00
ICON_CLEAR = 0xFF202020


def icon_cpu_baseline(scenes, desc, seconds: float):
    """The oracle's render over the icon requests, one request per host thread
    (render_seq: each request is one target), until `seconds` have passed."""
    import concurrent.futures as cf
    import threading
    from oracle import oracle as O
    cores = min(16, os.cpu_count() or 1)
    t_end = time.perf_counter() + seconds
    done = [0]
    lock = threading.Lock()

    def worker(k):
        i = k
        n = 0
        while time.perf_counter() < t_end:
            sc = scenes[i % len(scenes)]
            pix = np.full((sc.height, sc.width), ICON_CLEAR, np.uint32)
            dep = np.full((sc.height, sc.width), 1.0, np.float32)
            rc, _ = O.render(desc, sc.uniforms, sc.mesh, pix, dep)
            assert rc == 0
            n += 1
            i += cores
        with lock:
            done[0] += n

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(cores) as ex:
        list(ex.map(worker, range(cores)))
    el = time.perf_counter() - t0
    return {"value": done[0] / el, "unit": "icons/s", "cores": cores, "kind": "port",
            "sample": f"{done[0]} icon requests of this batch (in order, {cores} at a time, one host thread "
                      f"each) in {el:.1f} s, oracle render_seq"}


def bench_icons(args):
    """Batched icons: `--icons` requests (seeded spheres, tori, boxes of 200-20k
    log-uniform triangles; square viewports 32..256 px; Phong, LESS_WRITE, back-face
    culling, render_seq bands, colour and depth cleared in the frame).  One step =
    one whole batch (a hipGraph replay: one setup and one raster launch for all
    requests) from launch to completion; the step times give p50/p99 latency.
    Compared with the plain path: the same requests as one render() each."""
    from euc_amd import scenes as S
    from euc_amd.pipeline import BandPolicy, CullMode, DepthMode, PipelineDesc, Shader, Xform
    from euc_amd.renderer import Context, Request

    n = args.icons
    desc = PipelineDesc.make(Shader.PHONG, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.BACK,
                             band_policy=BandPolicy.SEQ)
    scenes = [S.icon_request(ICON_SEED + i) for i in range(n)]
    ctx = Context(0)
    meshes = [ctx.mesh(sc.mesh) for sc in scenes]
    tg = [(ctx.color_target(sc.width, sc.height), ctx.depth_target(sc.width, sc.height)) for sc in scenes]
    px = sum(sc.width * sc.height for sc in scenes)
    tris = sum(sc.mesh.n_triangles for sc in scenes)

    # plain path: one render() per request, all enqueued, one wait
    def plain_pass():
        for sc, m, (c, d) in zip(scenes, meshes, tg):
            ctx.render(desc, sc.uniforms, m, c, d, clear_color=ICON_CLEAR, clear_depth=1.0, sync=False)
        ctx.synchronize()

    plain_pass()
    plain_pass()
    sample = list(range(0, n, max(1, n // 64)))
    ref = {i: (tg[i][0].download(), tg[i][1].download()) for i in sample}
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        plain_pass()
    plain_s = (time.perf_counter() - t0) / reps
    # plain per-request latency: render() alone, waited for
    lat_plain = []
    for i in sample:
        sc, m, (c, d) = scenes[i], meshes[i], tg[i]
        t = time.perf_counter()
        ctx.render(desc, sc.uniforms, m, c, d, clear_color=ICON_CLEAR, clear_depth=1.0, sync=True)
        lat_plain.append(time.perf_counter() - t)

    for c, d in tg:  # the batch must write every target itself
        c.clear(0x5A5A5A5A)
        d.clear(0.5)
    b = ctx.batch().set(desc, [Request(m, c, d, sc.uniforms) for sc, m, (c, d) in zip(scenes, meshes, tg)],
                        clear_color=ICON_CLEAR, clear_depth=1.0)
    tile_w, tune_ms = b.autotune(reps=3)
    for _ in range(args.warmup):
        b.run(graph=True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        b.launch(graph=True)
        b.wait()
        lat.append(time.perf_counter() - t)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    parity = all(np.array_equal(tg[i][0].download(), ref[i][0]) and
                 np.array_equal(tg[i][1].download().view(np.uint32), ref[i][1].view(np.uint32)) for i in sample)

    # per-kernel times of the same batch, eager launches with events on the batch's stream
    ctx.reset_stats()
    ctx.set_profiling(True)
    for _ in range(min(args.steps, 20)):
        b.run(graph=False)
    ctx.set_profiling(False)
    stats = ctx.kernel_stats()
    kernels = {k: {"launches": v[0], "avg_us": v[1] * 1e3 / max(v[0], 1)} for k, v in stats.items()}
    raster_us = kernels.get("batch_raster", {}).get("avg_us", float("nan"))
    target_bytes = px * 8
    achieved = target_bytes / (raster_us * 1e-6) / 1e9 if raster_us == raster_us else None

    cpu = None
    if not args.no_cpu_baseline:
        cpu = icon_cpu_baseline(scenes, desc, args.cpu_seconds)
    icons_per_s = n / (ms_per_step * 1e-3)
    lat_ms = np.array(lat) * 1e3
    line = {
        "metric": "batched icons/s (4096 Phong icon requests per batch, p50/p99 batch latency)",
        "value": icons_per_s, "unit": "icons/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n} seeded icon requests (sphere/torus/box, 200-20k tris log-uniform, "
                               f"32-256 px square RGBA8+f32, phong, LESS_WRITE, cull back, render_seq bands, "
                               f"clears in frame), one hipGraph replay per batch",
                   "requests": n, "triangles": tris, "pixels": px, "tile_w": tile_w},
        "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
        "autotune_ms": tune_ms,
        "plain": {"icons_per_s": n / plain_s, "ms_per_batch": plain_s * 1e3,
                  "request_p50_ms": float(np.percentile(np.array(lat_plain) * 1e3, 50)),
                  "request_p99_ms": float(np.percentile(np.array(lat_plain) * 1e3, 99))},
        "speedup_vs_plain": icons_per_s / (n / plain_s),
        "parity_with_plain_path": parity,
        "kernels_us": kernels,
        "roofline": {"bound": "hbm", "kernel": "batch_raster", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBPS) if achieved else None, "traffic": None,
                     "algorithmic_bytes_per_launch": target_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    b.close()
    ctx.close()


def bench_plain(args):
    """The reference's own CPU-runnable case: 1024 seeded clip-space triangles, one
    flat RGBA8 colour per triangle, into 640x480 RGBA8 + f32, LESS_WRITE, back-face
    cull, render_par bands; one step = one frame, both targets cleared inside it."""
    from euc_amd import scenes as S
    from euc_amd.pipeline import (AaMode, BandPolicy, CoordinateMode, CullMode, DepthMode, PipelineDesc, PixelMode,
                                  Shader, Xform)
    from euc_amd.renderer import Context

    def make_desc():
        return PipelineDesc.make(Shader.FLAT, Xform.CLIP, PixelMode.WRITE, DepthMode.LESS_WRITE, CoordinateMode.VULKAN,
                                 CullMode.BACK, AaMode.NONE, BandPolicy.PAR)

    sc = S.plain_scene()
    desc = make_desc()
    W, H = sc.width, sc.height
    ctx = Context(0)
    mesh = ctx.mesh(sc.mesh)
    color, depth = ctx.color_target(W, H), ctx.depth_target(W, H)

    def frame():
        ctx.render(desc, sc.uniforms, mesh, color, depth, clear_color=0, clear_depth=1.0, sync=False)

    for _ in range(args.warmup):
        frame()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    ctx.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    ctx.reset_stats()
    ctx.set_profiling(True)
    for _ in range(args.steps):
        frame()
    ctx.synchronize()
    ctx.set_profiling(False)
    stats = ctx.kernel_stats()
    kernels = {k: {"launches": n, "avg_us": ms * 1e3 / max(n, 1)} for k, (n, ms) in stats.items()}
    raster_us = kernels.get("raster", {}).get("avg_us", float("nan"))
    out_color, out_depth = color.download(), depth.download()
    target_bytes = W * H * 8
    cpu, frags = None, None
    if not args.no_cpu_baseline:
        cb = cpu_baseline(sc, make_desc(), args.cpu_seconds)
        frags = cb["frags"]
        oc, od = cb["first"]
        parity = bool(np.array_equal(oc, out_color) and np.array_equal(od.view(np.uint32), out_depth.view(np.uint32)))
        cpu = {"value": frags / cb["frame_s"], "unit": "shaded frags/s", "cores": cb["threads"], "kind": "port",
               "sample": f"{cb['frames']} whole 640x480 frames of this workload (median frame "
                         f"{cb['frame_s'] * 1e3:.2f} ms), oracle render_par with {cb['threads']} threads",
               "frames_per_s": 1.0 / cb["frame_s"], "parity_with_gpu_frame": parity}
    frames_per_s = 1e3 / ms_per_step
    achieved = target_bytes / (raster_us * 1e-6) / 1e9 if raster_us == raster_us else None
    line = {
        "metric": "shaded frags/s (640x480 flat, 1024 random clip-space triangles)",
        "value": frags * frames_per_s if frags else None, "unit": "frags/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "1024 seeded clip-space triangles, flat RGBA8 per triangle, 640x480 RGBA8+f32, "
                               "LESS_WRITE, cull back, render_par bands, clears in frame", "width": W, "height": H,
                   "triangles": int(sc.mesh.n_triangles)},
        "frames_per_s": frames_per_s, "frags_per_frame": frags, "kernels_us": kernels,
        "roofline": {"bound": "hbm", "kernel": "raster", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBPS) if achieved else None, "traffic": None,
                     "algorithmic_bytes_per_launch": target_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()


def bench_heightfield(args):
    """Two-pass shadow mapping + texture: a depth-only pass of the 199,712-triangle
    value-noise heightfield into a 2048^2 f32 shadow map, then the 1080p main pass
    (bilinear 1024^2 RGBA8 texture, shadow lookup, Msaa { level: 1 }) sampling it.
    One step = both passes; both targets are cleared inside the frame."""
    from euc_amd import scenes as S
    from euc_amd.pipeline import AaMode, CullMode, DepthMode, PipelineDesc, PixelMode, Shader, Xform
    from euc_amd.renderer import Context

    sc = S.heightfield_scene()
    SS = sc.extra["shadow_size"]
    sd = PipelineDesc.make(Shader.DEPTH_ONLY, Xform.MVP, pixel_mode=PixelMode.PASS,
                           depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE)
    md = PipelineDesc.make(Shader.TEX_SHADOW, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, aa_mode=AaMode.Msaa(1))
    ctx = Context(0)
    mesh = ctx.mesh(sc.mesh)
    tex = ctx.texture(sc.texture)
    shadow = ctx.depth_target(SS, SS)
    color = ctx.color_target(sc.width, sc.height)
    depth = ctx.depth_target(sc.width, sc.height)

    def frame():
        ctx.render(sd, sc.extra["shadow_uniforms"], mesh, None, shadow, clear_depth=1.0, sync=False)
        ctx.render(md, sc.uniforms, mesh, color, depth, texture=tex, shadow=shadow, clear_color=0, clear_depth=1.0,
                   sync=False)

    for _ in range(args.warmup):
        frame()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    ctx.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    ctx.reset_stats()
    ctx.set_profiling(True)
    for _ in range(args.steps):
        frame()
    ctx.synchronize()
    ctx.set_profiling(False)
    stats = ctx.kernel_stats()
    kernels = {k: {"launches": n, "avg_us": ms * 1e3 / max(n, 1)} for k, (n, ms) in stats.items()}
    g_color, g_depth, g_shadow = color.download(), depth.download(), shadow.download()

    # algorithmic bytes per frame: shadow pass (positions, indices, 2048^2 f32 written),
    # main pass (positions + normals + uv, indices, texture and shadow map read once,
    # targets written once)
    m = sc.mesh
    nv, ni = m.pos.shape[0], m.idx.size
    frame_bytes = (nv * 12 + ni * 4 + SS * SS * 4) + (nv * 32 + ni * 4 + sc.texture.nbytes + SS * SS * 4 +
                                                      sc.width * sc.height * 8)
    cpu = None
    frags = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        cores = os.cpu_count() or 1
        times, first = [], None
        t_end = time.perf_counter() + args.cpu_seconds
        while True:
            sh = np.full((SS, SS), 1.0, np.float32)
            col = np.zeros((sc.height, sc.width), np.uint32)
            dep = np.full((sc.height, sc.width), 1.0, np.float32)
            sd.threads = md.threads = cores
            t = time.perf_counter()
            rc1, f1 = O.render(sd, sc.extra["shadow_uniforms"], sc.mesh, None, sh)
            rc2, f2 = O.render(md, sc.uniforms, sc.mesh, col, dep, texture=sc.texture, shadow=sh)
            times.append(time.perf_counter() - t)
            assert rc1 == 0 and rc2 == 0
            if first is None:
                first = (col, dep, sh)
            frags = f1 + f2
            if len(times) >= 2 and time.perf_counter() > t_end:
                break
        fs = float(np.median(times))
        T1 = max(1, min(SS // max(sd.group_rows(SS), 1), cores))
        T2 = max(1, min(sc.height // max(md.group_rows(sc.width), 1), cores))
        parity = bool(np.array_equal(first[0], g_color) and np.array_equal(first[1].view(np.uint32), g_depth.view(np.uint32))
                      and np.array_equal(first[2].view(np.uint32), g_shadow.view(np.uint32)))
        cpu = {"value": frags / fs, "unit": "shaded frags/s", "cores": max(T1, T2), "kind": "port",
               "sample": f"{len(times)} whole two-pass frames (median {fs * 1e3:.1f} ms), oracle render_par "
                         f"({T1} threads for the shadow pass, {T2} for the main pass)",
               "frames_per_s": 1.0 / fs, "parity_with_gpu_frame": parity}
    fps = 1e3 / ms_per_step
    line = {
        "metric": "shaded frags/s (two-pass shadow + texture + MSAA heightfield, 1080p)",
        "value": frags * fps if frags else None, "unit": "frags/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "199,712-triangle value-noise heightfield: depth-only pass into a 2048^2 f32 shadow "
                               "map, then 1920x1080 RGBA8+f32 main pass, 1024^2 bilinear texture, shadow lookup, "
                               "Msaa{level:1}, clears in frame", "triangles": int(m.n_triangles)},
        "frames_per_s": fps, "frame_hbm_gbps": frame_bytes / (ms_per_step * 1e-3) / 1e9,
        "frame_algorithmic_bytes": frame_bytes, "frags_per_frame": frags,
        "kernels_us": kernels, "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()


SERVE_SEED = 0x5EED0000 + 5


def bench_serving(args):
    """Serving loop: `--requests` textured renders arriving as a Poisson stream of
    `--rate` req/s; resident meshes of 1k-500k triangles and 256^2-2048^2 RGBA8
    textures; square targets 64..1024 px (log-uniform), allocated per request
    from a pool; requests admitted into 2 batch slots of up to 64 under a
    `--budget-gib` device budget.  Reported: frames/s and p50/p99 latency
    (completion seen by the host minus arrival), next to the plain path (one
    render() per request as it arrives) and a saturation run (all requests at
    once) of both.  The timed region is the whole stream; steps = requests."""
    from euc_amd import scenes as S
    from euc_amd.pipeline import BandPolicy, CullMode, DepthMode, PipelineDesc, Shader, Xform
    from euc_amd.renderer import Context
    from euc_amd.serving import ServeItem, Server, poisson_arrivals, run_poisson

    desc = PipelineDesc.make(Shader.TEXTURED, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.BACK,
                             band_policy=BandPolicy.SEQ)
    meshes, texs = S.serving_assets(SERVE_SEED)
    ctx = Context(0)
    gm = [ctx.mesh(m) for m in meshes]
    gt = [ctx.texture(t) for t in texs]
    resident = ctx.bytes_in_use()
    n = args.requests
    items, hosts = [], []
    for i in range(n):
        mi, ti, u, side = S.serving_request(SERVE_SEED * 7919 + i, len(gm), len(gt))
        items.append(ServeItem(gm[mi], u, side, side, gt[ti]))
        hosts.append((mi, ti, u, side))
    budget = int(args.budget_gib * (1 << 30))

    srv = Server(ctx, desc, budget_bytes=budget, max_batch=64, slots=2, clear_color=0, clear_depth=1.0)

    def served(rate, seed, keep=None):
        return run_poisson(srv, items, rate_hz=rate, seed=seed, keep=keep)

    def plain(rate, seed):
        """One render() per request as it arrives (targets from a per-size pool)."""
        arrivals = poisson_arrivals(n, rate, seed)
        pool = {}
        lat = np.empty(n)
        t0 = time.perf_counter()
        for i, it in enumerate(items):
            while time.perf_counter() - t0 < arrivals[i]:
                pass
            tg = pool.get(it.width)
            if tg is None:
                tg = pool[it.width] = (ctx.color_target(it.width, it.height), ctx.depth_target(it.width, it.height))
            ctx.render(desc, it.uniforms, it.mesh, tg[0], tg[1], texture=it.texture, clear_color=0, clear_depth=1.0)
            lat[i] = time.perf_counter() - t0 - arrivals[i]
        el = time.perf_counter() - t0 - arrivals[0]
        for c, d in pool.values():
            c.close()
            d.close()
        return lat, el

    # warm-up: a short stream through both paths sizes the scratch and the pools
    served(args.rate, 99)
    plain(args.rate * 4, 98)
    sample = list(range(0, n, max(1, n // 48)))
    res = served(args.rate, 1, keep=lambda i: i in sample)
    ok = all(s == 0 for s in res.statuses)
    frames_per_s = n / res.elapsed_s
    lat_ms = res.latency_s * 1e3
    sat = served(1e9, 2)
    srv.close()
    ctx.set_budget(0)
    sp_lat, sp_el = plain(1e9, 2)
    p_lat, p_el = plain(args.rate, 1)
    # results in the mix == each request rendered alone
    parity = ok
    for i in sample:
        it = items[i]
        c = ctx.color_target(it.width, it.height)
        d = ctx.depth_target(it.width, it.height)
        ctx.render(desc, it.uniforms, it.mesh, c, d, texture=it.texture, clear_color=0, clear_depth=1.0)
        parity = parity and np.array_equal(c.download(), res.outputs[i][0]) and \
            np.array_equal(d.download().view(np.uint32), res.outputs[i][1].view(np.uint32))
        c.close()
        d.close()

    cpu = None
    if not args.no_cpu_baseline:
        import concurrent.futures as cf
        import threading
        from oracle import oracle as O
        cores = min(16, os.cpu_count() or 1)
        t_end = time.perf_counter() + args.cpu_seconds
        cnt = [0]
        lock = threading.Lock()

        def worker(k):
            i, m = k, 0
            while time.perf_counter() < t_end:
                mi, ti, u, side = hosts[i % n]
                pix = np.zeros((side, side), np.uint32)
                dep = np.ones((side, side), np.float32)
                rc, _ = O.render(desc, u, meshes[mi], pix, dep, texture=texs[ti])
                assert rc == 0
                m += 1
                i += cores
            with lock:
                cnt[0] += m

        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(cores) as ex:
            list(ex.map(worker, range(cores)))
        el = time.perf_counter() - t0
        cpu = {"value": cnt[0] / el, "unit": "frames/s", "cores": cores, "kind": "port",
               "sample": f"{cnt[0]} requests of this stream (in order, one host thread each, {cores} threads) "
                         f"in {el:.1f} s, oracle render_seq"}

    st = res.stats
    line = {
        "metric": "served frames/s at a fixed Poisson rate, p50/p99 latency",
        "value": frames_per_s, "unit": "frames/s",
        "n_gpus": 1, "steps": n, "warmup": 1, "ms_per_step": res.elapsed_s * 1e3 / n,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n} textured render requests, Poisson {args.rate:.0f} req/s, targets 64-1024 px "
                               f"square RGBA8+f32, {len(meshes)} resident meshes 1k-500k tris, {len(texs)} textures "
                               f"256-2048 px, budget {args.budget_gib:g} GiB, 2 slots x 64 requests",
                   "rate_hz": args.rate, "requests": n, "budget_bytes": budget, "resident_bytes": resident},
        "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
        "server": {k: st[k] for k in ("batches", "rejected", "peak_bytes")},
        "plain": {"frames_per_s": n / p_el, "p50_ms": float(np.percentile(p_lat * 1e3, 50)),
                  "p99_ms": float(np.percentile(p_lat * 1e3, 99))},
        "saturation": {"served_frames_per_s": n / sat.elapsed_s,
                       "served_p50_ms": float(np.percentile(sat.latency_s * 1e3, 50)),
                       "served_p99_ms": float(np.percentile(sat.latency_s * 1e3, 99)),
                       "plain_frames_per_s": n / sp_el, "batches": sat.stats["batches"]},
        "results_equal_alone": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this renderer runs on one GPU; --gpus {args.gpus} is not supported "
              f"(several GPUs would each own a set of screen-tile rows)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.workload == "icons":
        bench_icons(args)
        return
    if args.workload == "serving":
        bench_serving(args)
        return
    if args.workload == "heightfield":
        bench_heightfield(args)
        return
    if args.workload == "plain":
        bench_plain(args)
        return

    from euc_amd import scenes as S
    from euc_amd.pipeline import DepthMode, PipelineDesc, Shader, Xform
    from euc_amd.renderer import Context

    scene = S.torus_scene()
    shader = Shader.PHONG if args.shader == "phong" else Shader.GOURAUD
    desc = PipelineDesc.make(shader, Xform.MVP, depth_mode=DepthMode.LESS_WRITE)
    W, H = scene.width, scene.height
    ab = algorithmic_bytes(scene)

    ctx = Context(0)
    mesh = ctx.mesh(scene.mesh)
    color = ctx.color_target(W, H)
    depth = ctx.depth_target(W, H)

    def frame():
        ctx.render(desc, scene.uniforms, mesh, color, depth, clear_color=0, clear_depth=1.0, sync=False)

    for _ in range(args.warmup):
        frame()
    ctx.synchronize()

    # timed region: exactly K frames, synchronised on both sides
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps

    # the same K frames again with per-kernel HIP events on the context's stream
    ctx.reset_stats()
    ctx.set_profiling(True)
    for _ in range(args.steps):
        frame()
    ctx.synchronize()
    ctx.set_profiling(False)
    stats = ctx.kernel_stats()
    kernels = {k: {"launches": n, "avg_us": ms * 1e3 / max(n, 1)} for k, (n, ms) in stats.items()}
    raster_us = kernels.get("raster", {}).get("avg_us", float("nan"))

    out_color = color.download()
    out_depth = depth.download()

    cpu = None
    frags = None
    if not args.no_cpu_baseline:
        cb = cpu_baseline(scene, PipelineDesc.make(shader, Xform.MVP, depth_mode=DepthMode.LESS_WRITE),
                          args.cpu_seconds)
        frags = cb["frags"]
        oc, od = cb["first"]
        parity = bool(np.array_equal(oc, out_color) and np.array_equal(od.view(np.uint32), out_depth.view(np.uint32)))
        cpu = {"value": frags / cb["frame_s"], "unit": "shaded frags/s", "cores": cb["threads"],
               "kind": "port",
               "sample": f"{cb['frames']} whole 1920x1080 frames of this workload (median frame "
                         f"{cb['frame_s'] * 1e3:.1f} ms), oracle render_par with {cb['threads']} threads",
               "frames_per_s": 1.0 / cb["frame_s"], "parity_with_gpu_frame": parity}

    frames_per_s = 1e3 / ms_per_step
    value = frags * frames_per_s if frags else None
    achieved = ab["raster"] / (raster_us * 1e-6) / 1e9 if raster_us == raster_us else None
    traffic = measured_traffic("k_raster<PHONG>" if args.shader == "phong" else "k_raster<GOURAUD>")
    line = {
        "metric": "shaded frags/s (1080p Phong, 4x65536-tri torus)",
        "value": value,
        "unit": "frags/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"torus 256x128 x4 instances, 1920x1080 RGBA8+f32, {args.shader}, "
                               f"LESS_WRITE, cull back, clears in frame", "width": W, "height": H,
                   "triangles": scene.mesh.n_triangles},
        "frames_per_s": frames_per_s,
        "frame_hbm_gbps": ab["frame"] / (ms_per_step * 1e-3) / 1e9,
        "frags_per_frame": frags,
        "kernels_us": kernels,
        "roofline": {"bound": "hbm", "kernel": "raster", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBPS) if achieved else None,
                     "traffic": traffic["bytes"] if traffic else None, "traffic_detail": traffic,
                     "algorithmic_bytes_per_launch": ab["raster"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
