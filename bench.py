"""Benchmark of the triangle path on one MI355X.

Default (`--workload all`): one JSON line whose headline is the 1080p Phong
torus frame -- four 65,536-triangle tori (256 x 128 segments, seeded model
matrices), per-fragment Phong shading (2 directional lights), LESS_WRITE depth,
back-face culling, 1920x1080 RGBA8 colour and f32 depth targets both cleared
inside the frame; one step = one frame.  `value` = shaded fragments per second:
the number of emit_fragment calls euc 0.6.0 makes for this frame (coverage,
z-clip and depth test passed, in submission order), taken from the oracle's run
of the same frame in the cpu_baseline leg, times frames per second of the HIP
path.  Inputs are resident in HBM before the timed region starts.

The same line carries the other workloads under "workloads": the torus frame
with Gouraud shading, the 1024-triangle 640x480 flat frame, the two-pass
shadow + texture heightfield frame, the 4096-request batched icons (icons/s,
p50/p99 batch latency, speed-up over one render() per icon) and the serving
loop (frames/s, p50/p99 latency) -- each with its kernel times, roofline and a
short CPU sample of the oracle.  `--workload X` runs one of them alone as the
headline.

Warm-up steps are blocking renders, so bins and overflow storage have grown to
what the frame needs before the timed region; the timed steps are enqueued
back to back and synchronised once at each end.
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
WORKLOADS = ["torus", "gouraud", "plain", "heightfield", "icons", "serving"]


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per dispatch of `kernel` in `workload` from the committed rocprofv3
    PMC summary (FETCH_SIZE and WRITE_SIZE, each scaled by the copy-probe factor),
    or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        k = t["workloads"][workload][kernel]
        return {"bytes": k["fetch_bytes"] + k["write_bytes"], "fetch_bytes": k["fetch_bytes"],
                "write_bytes": k["write_bytes"], "source": "profiles/traffic.json"}
    except (OSError, KeyError, ValueError):
        return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="all", choices=["all"] + WORKLOADS)
    ap.add_argument("--rate", type=float, default=40000.0, help="Poisson arrival rate of the serving workload (req/s)")
    ap.add_argument("--requests", type=int, default=20000, help="requests of the serving workload")
    ap.add_argument("--budget-gib", type=float, default=16.0, help="device-memory budget of the serving workload")
    ap.add_argument("--icons", type=int, default=4096, help="requests per batch of the icons workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--icon-tile-width", type=int, choices=(8, 16, 32), default=None,
                    help="icons: this tile width instead of the autotuner's pick (counter runs, so that each "
                         "counter pass measures the same kernel)")
    ap.add_argument("--no-plain-path", action="store_true",
                    help="icons: leave out the comparison with one render() per request (runs under a counter "
                         "profiler, which cannot follow thousands of queued launches)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound of the headline's CPU sample")
    ap.add_argument("--cpu-seconds-extra", type=float, default=4.0, help="bound of each other workload's CPU sample")
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
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(10):
        torch.add(a, 1.0, out=b)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / 10
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "dispatches": 13,
                      "copy_gbps": 2 * n * 4 / (ms * 1e-3) / 1e9}))


def median_timed(fn, seconds: float, max_samples: int = 32, warm_s: float = 1.0):
    """benches/teapot.rs:228-234's protocol: run for `warm_s` untimed, then take
    up to `max_samples` samples (fewer if `seconds` runs out first; at least 2)
    and return (median seconds, samples, first result)."""
    t_end = time.perf_counter() + warm_s
    first = fn()
    while time.perf_counter() < t_end:
        fn()
    times = []
    t_end = time.perf_counter() + seconds
    while len(times) < max_samples:
        t = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t)
        if len(times) >= 2 and time.perf_counter() > t_end:
            break
    return float(np.median(times)), len(times), first


def kernel_table(ctx):
    stats = ctx.kernel_stats()
    return {k: {"launches": n, "avg_us": ms * 1e3 / max(n, 1)} for k, (n, ms) in stats.items()}


def timed_frames(ctx, frame, args):
    """Blocking warm-up frames, then exactly `steps` frames back to back between two
    synchronisations; then the same frames again with HIP events around each
    kernel on the context's stream.  Returns (ms per step, kernel table)."""
    for _ in range(max(args.warmup, 2)):
        frame(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame(False)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    ctx.reset_stats()
    ctx.set_profiling(True)
    for _ in range(args.steps):
        frame(False)
    ctx.synchronize()
    ctx.set_profiling(False)
    return ms, kernel_table(ctx)


def roofline(kernels, name, algorithmic_bytes, workload, traffic_kernel):
    us = kernels.get(name, {}).get("avg_us")
    achieved = algorithmic_bytes / (us * 1e-6) / 1e9 if us else None
    tr = measured_traffic(workload, traffic_kernel)
    return {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBPS) if achieved else None, "traffic": tr["bytes"] if tr else None,
            "traffic_detail": tr, "algorithmic_bytes_per_launch": algorithmic_bytes}


def line_base(metric, value, unit, args, ms_per_step, config, steps=None, warmup=None):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1,
            "steps": args.steps if steps is None else steps, "warmup": args.warmup if warmup is None else warmup,
            "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic", "config": config}


# --------------------------------------------------------------------------- #
# one-frame workloads
# --------------------------------------------------------------------------- #

def frame_cpu_baseline(sc, desc, seconds, texture=None, label=""):
    """The oracle's render_par (row bands, G = 20000*2^L/W, T = min(H/G, cores)
    threads) of the whole frame, both targets cleared inside it."""
    from oracle import oracle as O
    cores = os.cpu_count() or 1
    threads = max(1, min(sc.height // max(desc.group_rows(sc.width), 1), cores))
    desc.threads = threads
    out = {}

    def run():
        color = np.full((sc.height, sc.width), sc.clear_color, np.uint32)
        depth = np.full((sc.height, sc.width), 1.0, np.float32)
        rc, frags = O.render(desc, sc.uniforms, sc.mesh, color, depth, texture=texture)
        assert rc == 0
        out["frags"] = frags
        return color, depth

    med, n, first = median_timed(run, seconds)
    return {"frame_s": med, "samples": n, "threads": threads, "frags": out["frags"], "first": first}


def run_frame_workload(args, ctx, key, sc, desc, metric, config, cpu_seconds, traffic_key):
    """One render() of `sc` per step; the headline value is the oracle's fragment
    count of the frame times frames/s."""
    mesh = ctx.mesh(sc.mesh)
    W, H = sc.width, sc.height
    color, depth = ctx.color_target(W, H), ctx.depth_target(W, H)

    def frame(sync):
        ctx.render(desc, sc.uniforms, mesh, color, depth, clear_color=sc.clear_color, clear_depth=1.0, sync=sync)

    ms, kernels = timed_frames(ctx, frame, args)
    out_color, out_depth = color.download(), depth.download()
    target_bytes = W * H * 8
    m = sc.mesh
    mesh_bytes = m.pos.nbytes + (m.nrm.nbytes if m.nrm is not None else 0) + \
        (m.idx.nbytes if m.idx is not None else 0) + (m.rgba8.nbytes if m.rgba8 is not None else 0)
    cpu, frags = None, None
    if not args.no_cpu_baseline:
        from euc_amd.pipeline import PipelineDesc
        cb = frame_cpu_baseline(sc, PipelineDesc(**{k: getattr(desc, k) for k in desc.__dataclass_fields__}),
                                cpu_seconds)
        frags = cb["frags"]
        oc, od = cb["first"]
        parity = bool(np.array_equal(oc, out_color) and np.array_equal(od.view(np.uint32), out_depth.view(np.uint32)))
        cpu = {"value": frags / cb["frame_s"], "unit": "shaded frags/s", "cores": cb["threads"], "kind": "port",
               "sample": f"median of {cb['samples']} whole {W}x{H} frames of this workload after 1 s of warm-up "
                         f"({cb['frame_s'] * 1e3:.2f} ms per frame), oracle render_par with {cb['threads']} "
                         f"threads",
               "frames_per_s": 1.0 / cb["frame_s"], "parity_with_gpu_frame": parity}
        if not parity:
            print(f"bench.py: {key}: the GPU frame differs from the oracle's", file=sys.stderr)
    fps = 1e3 / ms
    line = line_base(metric, frags * fps if frags else None, "frags/s", args, ms, config)
    line.update({
        "frames_per_s": fps, "frame_hbm_gbps": (mesh_bytes + target_bytes) / (ms * 1e-3) / 1e9,
        "frags_per_frame": frags, "kernels_us": kernels,
        "roofline": roofline(kernels, "raster", target_bytes, traffic_key, "k_raster"),
        "cpu_baseline": cpu,
    })
    return line


def bench_torus(args, ctx, shader_name, cpu_seconds):
    from euc_amd import scenes as S
    from euc_amd.pipeline import DepthMode, PipelineDesc, Shader, Xform
    sc = S.torus_scene()
    shader = Shader.PHONG if shader_name == "phong" else Shader.GOURAUD
    desc = PipelineDesc.make(shader, Xform.MVP, depth_mode=DepthMode.LESS_WRITE)
    cfg = {"workload": f"torus 256x128 x4 instances, 1920x1080 RGBA8+f32, {shader_name}, LESS_WRITE, cull back, "
                       f"clears in frame", "width": sc.width, "height": sc.height,
           "triangles": int(sc.mesh.n_triangles)}
    key = "torus_phong" if shader_name == "phong" else "torus_gouraud"
    return run_frame_workload(args, ctx, key, sc, desc,
                              f"shaded frags/s (1080p {shader_name.capitalize()}, 4x65536-tri torus)", cfg,
                              cpu_seconds, key)


def bench_plain(args, ctx, cpu_seconds):
    """The reference's own CPU-runnable case: 1024 seeded clip-space triangles, one
    flat RGBA8 colour per triangle, into 640x480 RGBA8 + f32, LESS_WRITE, back-face
    cull, render_par bands; one step = one frame, both targets cleared inside it."""
    from euc_amd import scenes as S
    from euc_amd.pipeline import (AaMode, BandPolicy, CoordinateMode, CullMode, DepthMode, PipelineDesc, PixelMode,
                                  Shader, Xform)
    sc = S.plain_scene()
    desc = PipelineDesc.make(Shader.FLAT, Xform.CLIP, PixelMode.WRITE, DepthMode.LESS_WRITE, CoordinateMode.VULKAN,
                             CullMode.BACK, AaMode.NONE, BandPolicy.PAR)
    cfg = {"workload": "1024 seeded clip-space triangles, flat RGBA8 per triangle, 640x480 RGBA8+f32, LESS_WRITE, "
                       "cull back, render_par bands, clears in frame", "width": sc.width, "height": sc.height,
           "triangles": int(sc.mesh.n_triangles)}
    return run_frame_workload(args, ctx, "plain", sc, desc,
                              "shaded frags/s (640x480 flat, 1024 random clip-space triangles)", cfg, cpu_seconds,
                              "plain")


def bench_heightfield(args, ctx, cpu_seconds):
    """Two-pass shadow mapping + texture: a depth-only pass of the 199,712-triangle
    value-noise heightfield into a 2048^2 f32 shadow map, then the 1080p main pass
    (bilinear 1024^2 RGBA8 texture, shadow lookup, Msaa { level: 1 }) sampling it.
    One step = both passes; both targets are cleared inside the frame."""
    from euc_amd import scenes as S
    from euc_amd.pipeline import AaMode, CullMode, DepthMode, PipelineDesc, PixelMode, Shader, Xform
    from oracle import oracle as O

    sc = S.heightfield_scene()
    SS = sc.extra["shadow_size"]

    def descs():
        sd = PipelineDesc.make(Shader.DEPTH_ONLY, Xform.MVP, pixel_mode=PixelMode.PASS,
                               depth_mode=DepthMode.LESS_WRITE, cull=CullMode.NONE)
        md = PipelineDesc.make(Shader.TEX_SHADOW, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, aa_mode=AaMode.Msaa(1))
        return sd, md

    sd, md = descs()
    mesh = ctx.mesh(sc.mesh)
    tex = ctx.texture(sc.texture)
    shadow = ctx.depth_target(SS, SS)
    color = ctx.color_target(sc.width, sc.height)
    depth = ctx.depth_target(sc.width, sc.height)

    def frame(sync):
        ctx.render(sd, sc.extra["shadow_uniforms"], mesh, None, shadow, clear_depth=1.0, sync=sync)
        ctx.render(md, sc.uniforms, mesh, color, depth, texture=tex, shadow=shadow, clear_color=0, clear_depth=1.0,
                   sync=sync)

    ms, kernels = timed_frames(ctx, frame, args)
    g_color, g_depth, g_shadow = color.download(), depth.download(), shadow.download()
    # algorithmic bytes per frame: shadow pass (positions, indices, 2048^2 f32 written),
    # main pass (positions + normals + uv, indices, texture and shadow map read once,
    # targets written once)
    m = sc.mesh
    nv, ni = m.pos.shape[0], m.idx.size
    frame_bytes = (nv * 12 + ni * 4 + SS * SS * 4) + (nv * 32 + ni * 4 + sc.texture.nbytes + SS * SS * 4 +
                                                      sc.width * sc.height * 8)
    cpu, frags = None, None
    if not args.no_cpu_baseline:
        cores = os.cpu_count() or 1
        osd, omd = descs()
        osd.threads = omd.threads = cores
        out = {}

        def run():
            sh = np.full((SS, SS), 1.0, np.float32)
            col = np.zeros((sc.height, sc.width), np.uint32)
            dep = np.full((sc.height, sc.width), 1.0, np.float32)
            rc1, f1 = O.render(osd, sc.extra["shadow_uniforms"], sc.mesh, None, sh)
            rc2, f2 = O.render(omd, sc.uniforms, sc.mesh, col, dep, texture=sc.texture, shadow=sh)
            assert rc1 == 0 and rc2 == 0
            out["frags"] = f1 + f2
            return col, dep, sh

        fs, n, first = median_timed(run, cpu_seconds, warm_s=0.0)
        frags = out["frags"]
        T1 = max(1, min(SS // max(osd.group_rows(SS), 1), cores))
        T2 = max(1, min(sc.height // max(omd.group_rows(sc.width), 1), cores))
        parity = bool(np.array_equal(first[0], g_color) and np.array_equal(first[1].view(np.uint32), g_depth.view(np.uint32))
                      and np.array_equal(first[2].view(np.uint32), g_shadow.view(np.uint32)))
        cpu = {"value": frags / fs, "unit": "shaded frags/s", "cores": max(T1, T2), "kind": "port",
               "sample": f"median of {n} whole two-pass frames ({fs * 1e3:.1f} ms per frame), oracle render_par "
                         f"({T1} threads for the shadow pass, {T2} for the main pass)",
               "frames_per_s": 1.0 / fs, "parity_with_gpu_frame": parity}
    fps = 1e3 / ms
    line = line_base("shaded frags/s (two-pass shadow + texture + MSAA heightfield, 1080p)",
                     frags * fps if frags else None, "frags/s", args, ms,
                     {"workload": "199,712-triangle value-noise heightfield: depth-only pass into a 2048^2 f32 "
                                  "shadow map, then 1920x1080 RGBA8+f32 main pass, 1024^2 bilinear texture, shadow "
                                  "lookup, Msaa{level:1}, clears in frame", "triangles": int(m.n_triangles)})
    rl = roofline(kernels, "raster", sc.width * sc.height * 8 + SS * SS * 4, "heightfield", "k_raster")
    rl["note"] = "both raster launches of the frame: the 2048^2 f32 shadow map and the 1080p RGBA8 + f32 targets"
    if kernels.get("raster"):
        rl["achieved"] = (sc.width * sc.height * 8 + SS * SS * 4) / (2 * kernels["raster"]["avg_us"] * 1e-6) / 1e9
        rl["frac"] = rl["achieved"] / HBM_PEAK_GBPS
    line.update({"frames_per_s": fps, "frame_hbm_gbps": frame_bytes / (ms * 1e-3) / 1e9,
                 "frame_algorithmic_bytes": frame_bytes, "frags_per_frame": frags, "kernels_us": kernels,
                 "roofline": rl, "cpu_baseline": cpu})
    return line


# --------------------------------------------------------------------------- #
# batched icons
# --------------------------------------------------------------------------- #

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


def bench_icons(args, ctx, cpu_seconds):
    """Batched icons: `--icons` requests (seeded spheres, tori, boxes of 200-20k
    log-uniform triangles; square viewports 32..256 px; Phong, LESS_WRITE, back-face
    culling, render_seq bands, colour and depth cleared in the frame).  One step =
    one whole batch (a hipGraph replay: one setup and one raster launch for all
    requests) from launch to completion; the step times give p50/p99 latency.
    Compared with the plain path: the same requests as one render() each."""
    from euc_amd import scenes as S
    from euc_amd.pipeline import BandPolicy, CullMode, DepthMode, PipelineDesc, Shader, Xform
    from euc_amd.renderer import Request

    n = args.icons
    desc = PipelineDesc.make(Shader.PHONG, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.BACK,
                             band_policy=BandPolicy.SEQ)
    scenes = [S.icon_request(ICON_SEED + i) for i in range(n)]
    meshes = [ctx.mesh(sc.mesh) for sc in scenes]
    tg = [(ctx.color_target(sc.width, sc.height), ctx.depth_target(sc.width, sc.height)) for sc in scenes]
    px = sum(sc.width * sc.height for sc in scenes)
    tris = sum(sc.mesh.n_triangles for sc in scenes)

    # plain path: one render() per request, all enqueued, one wait
    def plain_pass():
        for sc, m, (c, d) in zip(scenes, meshes, tg):
            ctx.render(desc, sc.uniforms, m, c, d, clear_color=ICON_CLEAR, clear_depth=1.0, sync=False)
        ctx.synchronize()

    sample = list(range(0, n, max(1, n // 64)))
    ref, plain_s, lat_plain = None, None, []
    if not args.no_plain_path:
        for sc, m, (c, d) in zip(scenes, meshes, tg):  # blocking once: every request's bins sized
            ctx.render(desc, sc.uniforms, m, c, d, clear_color=ICON_CLEAR, clear_depth=1.0, sync=True)
        plain_pass()
        ref = {i: (tg[i][0].download(), tg[i][1].download()) for i in sample}
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            plain_pass()
        plain_s = (time.perf_counter() - t0) / reps
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
    if args.icon_tile_width:
        b.set_tile_width(args.icon_tile_width)
        tile_w, tune_ms = args.icon_tile_width, None
    else:
        tile_w, tune_ms = b.autotune(reps=3)
    for _ in range(max(args.warmup, 2)):
        b.run(graph=True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        b.launch(graph=True)
        b.wait()
        lat.append(time.perf_counter() - t)
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    parity = None if ref is None else all(
        np.array_equal(tg[i][0].download(), ref[i][0]) and
        np.array_equal(tg[i][1].download().view(np.uint32), ref[i][1].view(np.uint32)) for i in sample)
    ctx.reset_stats()
    ctx.set_profiling(True)
    for _ in range(min(args.steps, 20)):
        b.run(graph=False)
    ctx.set_profiling(False)
    kernels = kernel_table(ctx)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = icon_cpu_baseline(scenes, desc, cpu_seconds)
    icons_per_s = n / (ms_per_step * 1e-3)
    lat_ms = np.array(lat) * 1e3
    line = line_base("batched icons/s (4096 Phong icon requests per batch, p50/p99 batch latency)", icons_per_s,
                     "icons/s", args, ms_per_step,
                     {"workload": f"{n} seeded icon requests (sphere/torus/box, 200-20k tris log-uniform, "
                                  f"32-256 px square RGBA8+f32, phong, LESS_WRITE, cull back, render_seq bands, "
                                  f"clears in frame), one hipGraph replay per batch",
                      "requests": n, "triangles": tris, "pixels": px, "tile_w": tile_w})
    line.update({
        "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
        "autotune_ms": tune_ms,
        "plain": None if plain_s is None else {
            "icons_per_s": n / plain_s, "ms_per_batch": plain_s * 1e3,
            "request_p50_ms": float(np.percentile(np.array(lat_plain) * 1e3, 50)),
            "request_p99_ms": float(np.percentile(np.array(lat_plain) * 1e3, 99))},
        "speedup_vs_plain": None if plain_s is None else icons_per_s / (n / plain_s),
        "parity_with_plain_path": parity,
        "kernels_us": kernels,
        "roofline": roofline(kernels, "batch_raster", px * 8, "icons", "k_raster_batch"),
        "cpu_baseline": cpu,
    })
    b.close()
    for c, d in tg:
        c.close()
        d.close()
    for m in meshes:
        m.close()
    return line


# --------------------------------------------------------------------------- #
# serving
# --------------------------------------------------------------------------- #

SERVE_SEED = 0x5EED0000 + 5


def bench_serving(args, ctx, cpu_seconds):
    """Serving loop: `--requests` textured renders arriving as a Poisson stream of
    `--rate` req/s; resident meshes of 1k-500k triangles and 256^2-2048^2 RGBA8
    textures; square targets 64..1024 px (log-uniform), allocated per request
    from a pool; requests admitted into 2 batch slots of up to 64 under a
    `--budget-gib` device budget.  Reported: frames/s and p50/p99 latency
    (completion seen by the server minus arrival), next to the plain path (one
    render() per request as it arrives) and a saturation run (all requests at
    once) of both.  The timed region is the whole stream; steps = requests."""
    from euc_amd import scenes as S
    from euc_amd.pipeline import BandPolicy, CullMode, DepthMode, PipelineDesc, Shader, Xform
    from euc_amd.serving import ServeItem, Server, poisson_arrivals, run_poisson

    desc = PipelineDesc.make(Shader.TEXTURED, Xform.MVP, depth_mode=DepthMode.LESS_WRITE, cull=CullMode.BACK,
                             band_policy=BandPolicy.SEQ)
    meshes, texs = S.serving_assets(SERVE_SEED)
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
            tgt = pool.get(it.width)
            if tgt is None:
                tgt = pool[it.width] = (ctx.color_target(it.width, it.height), ctx.depth_target(it.width, it.height))
            ctx.render(desc, it.uniforms, it.mesh, tgt[0], tgt[1], texture=it.texture, clear_color=0, clear_depth=1.0)
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
    parity = ok  # results in the mix == each request rendered alone
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
        t_end = time.perf_counter() + cpu_seconds
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
    line = line_base("served frames/s at a fixed Poisson rate, p50/p99 latency", frames_per_s, "frames/s", args,
                     res.elapsed_s * 1e3 / n,
                     {"workload": f"{n} textured render requests, Poisson {args.rate:.0f} req/s, targets 64-1024 px "
                                  f"square RGBA8+f32, {len(meshes)} resident meshes 1k-500k tris, {len(texs)} "
                                  f"textures 256-2048 px, budget {args.budget_gib:g} GiB, 2 slots x 64 requests",
                      "rate_hz": args.rate, "requests": n, "budget_bytes": budget, "resident_bytes": resident},
                     steps=n, warmup=1)
    line.update({
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
    })
    for m in gm:
        m.close()
    for t in gt:
        t.close()
    return line


# --------------------------------------------------------------------------- #

SUMMARY_KEYS = ("metric", "value", "unit", "ms_per_step", "frames_per_s", "frags_per_frame", "p50_ms", "p99_ms",
                "speedup_vs_plain", "plain", "saturation", "results_equal_alone", "parity_with_plain_path",
                "kernels_us", "roofline", "cpu_baseline", "config")


def run_one(name, args, headline):
    from euc_amd.renderer import Context
    cs = args.cpu_seconds if headline else args.cpu_seconds_extra
    ctx = Context(0)
    try:
        if name == "torus":
            return bench_torus(args, ctx, "phong", cs)
        if name == "gouraud":
            return bench_torus(args, ctx, "gouraud", cs)
        if name == "plain":
            return bench_plain(args, ctx, cs)
        if name == "heightfield":
            return bench_heightfield(args, ctx, cs)
        if name == "icons":
            return bench_icons(args, ctx, cs)
        if name == "serving":
            return bench_serving(args, ctx, cs)
        raise ValueError(name)
    finally:
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
    if args.workload != "all":
        print(json.dumps(run_one(args.workload, args, True)))
        return
    line = run_one("torus", args, True)
    print(f"bench.py: torus_phong {line['ms_per_step'] * 1e3:.1f} us per frame", file=sys.stderr, flush=True)
    extra = {}
    for name, key in [("gouraud", "torus_gouraud"), ("plain", "plain_c1"), ("heightfield", "heightfield_c3"),
                      ("icons", "icons_c4"), ("serving", "serving_c5")]:
        t = time.perf_counter()
        r = run_one(name, args, False)
        extra[key] = {k: r[k] for k in SUMMARY_KEYS if k in r}
        print(f"bench.py: {key} done in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    line["workloads"] = extra
    print(json.dumps(line))


if __name__ == "__main__":
    main()
