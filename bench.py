"""Benchmark of the triangle path on one MI355X.

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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--shader", default="phong", choices=["phong", "gouraud"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound of the CPU baseline sample")
    ap.add_argument("--check", action="store_true", help="also compare one GPU frame with the oracle")
    return ap.parse_args()


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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the path would shard screen tiles across GPUs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

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
                     "traffic": None, "algorithmic_bytes_per_launch": ab["raster"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
