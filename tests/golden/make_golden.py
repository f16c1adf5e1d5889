"""Generate the golden fixtures of tests/golden/ with the independent pure-Python
restatement (tests/ref_tiny.py) -- not with the C oracle, so the fixtures pin
the oracle against a second transcription of the reference.

Each fixture holds its inputs (clip-space vertex stream, per-vertex RGBA8 or
Rgba<f32>, the pipeline configuration, the initial depth) and the expected
outputs (colour, depth, emit_fragment count).  Run from the repository root:

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests import ref_tiny as R  # noqa: E402

CASES = [
    # name, seed, W, H, shader, depth_test, depth_write, cull, y_up, z_clip, msaa, bands, n_tris
    ("flat_less_seq", 101, 48, 40, "flat", 1, True, 1, False, (0.0, 1.0), 0, "seq", 24),
    ("flat_none_front_yup", 102, 40, 36, "flat", 0, False, 2, True, (-1.0, 1.0), 0, "seq", 20),
    ("color_greater_msaa1", 103, 44, 32, "color", 3, True, 0, False, (0.0, 1.0), 1, "seq", 16),
    ("color_equal_noclip", 104, 36, 36, "color", 2, False, 0, False, None, 0, "seq", 18),
    ("color_less_msaa2_wide", 105, 64, 24, "color", 1, True, 1, True, (0.0, 1.0), 2, "seq", 14),
    ("flat_lesspass_grid", 106, 32, 32, "flat", 1, False, 0, False, (0.0, 1.0), 0, "seq", 30),
]


def make_inputs(seed, n_tris, grid=False):
    rng = np.random.default_rng(seed)
    n = n_tris * 3
    pos = np.empty((n, 4), np.float32)
    pos[:, :2] = rng.uniform(-1.3, 1.3, (n, 2))
    pos[:, 2] = rng.uniform(-0.2, 1.2, n)
    pos[:, 3] = rng.uniform(0.7, 1.6, n)
    if grid:
        pos[:, :2] = np.round(pos[:, :2] * 8) / 8
        pos[:, 3] = 1.0
    rgba8 = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    col = rng.uniform(-0.1, 1.1, (n, 4)).astype(np.float32)
    return pos, rgba8, col


def main():
    for (name, seed, W, H, shader, dt, dw, cull, y_up, zc, msaa, bands, nt) in CASES:
        pos, rgba8, col = make_inputs(seed, nt, grid=name.endswith("grid"))
        rng = np.random.default_rng(seed + 1)
        d0 = rng.uniform(0.0, 1.0, (H, W)).astype(np.float32) if dt in (1, 2, 3) else np.ones((H, W), np.float32)
        if dt == 2:
            # depth pre-pass: the EQUAL pass then re-finds exactly the winning z values
            pre = {"shader": shader, "depth_test": 1, "depth_write": True, "pixel_write": False, "cull": cull,
                   "y_up": y_up, "z_clip": zc, "msaa_level": 0, "bands": bands}
            d0 = np.ones((H, W), np.float32)
            R.render(pre, [list(p) for p in pos], [[np.float32(0)] * 4] * len(pos), None, d0)
        cfg = {"shader": shader, "depth_test": dt, "depth_write": dw, "pixel_write": True, "cull": cull,
               "y_up": y_up, "z_clip": zc, "msaa_level": msaa, "bands": bands}
        vd = [int(c) for c in rgba8] if shader == "flat" else [list(map(np.float32, c)) for c in col]
        pix = np.full((H, W), 0x0A0B0C0D, np.uint32)
        dep = d0.copy()
        frags = R.render(cfg, [list(p) for p in pos], vd, pix, dep)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), pos=pos, rgba8=rgba8, col=col, depth0=d0,
                            pixel=pix, depth=dep, frags=np.int64(frags), cfg=json.dumps(cfg))
        print(f"{name}: {W}x{H}, {nt} triangles, {frags} fragments")


if __name__ == "__main__":
    main()
