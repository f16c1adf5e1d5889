"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the independent Python restatement) against the C oracle (CPU) and the
HIP path (GPU)."""
import glob
import json
import os

import numpy as np
import pytest

from euc_amd import scenes as S
from euc_amd.pipeline import BandPolicy, PipelineDesc, Shader, Xform
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    cfg = json.loads(str(z["cfg"]))
    flat = cfg["shader"] == "flat"
    mesh = S.Mesh(pos=z["pos"], count=z["pos"].shape[0], rgba8=z["rgba8"] if flat else None,
                  col=None if flat else z["col"])
    zc = cfg["z_clip"]
    desc = PipelineDesc(
        shader=Shader.FLAT if flat else Shader.COLOR, xform=Xform.CLIP, pixel_write=cfg["pixel_write"],
        depth_test=cfg["depth_test"], depth_write=cfg["depth_write"], y_up=cfg["y_up"], z_clip=zc is not None,
        z_min=zc[0] if zc else 0.0, z_max=zc[1] if zc else 0.0, cull=cfg["cull"], aa_msaa=cfg["msaa_level"] > 0,
        aa_level=cfg["msaa_level"], band_policy=BandPolicy.SEQ if cfg["bands"] == "seq" else BandPolicy.PAR)
    return z, desc, mesh


def test_fixtures_exist():
    assert len(FIXTURES) >= 6


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_reproduces_golden(path):
    z, desc, mesh = load(path)
    H, W = z["pixel"].shape
    pix = np.full((H, W), 0x0A0B0C0D, np.uint32)
    dep = z["depth0"].copy()
    rc, frags = O.render(desc, S.uniforms(), mesh, pix, dep)
    assert rc == 0 and frags == int(z["frags"])
    assert np.array_equal(pix, z["pixel"])
    assert np.array_equal(dep.view(np.uint32), z["depth"].view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_hip_reproduces_golden(gpu_ctx, path):
    z, desc, mesh = load(path)
    H, W = z["pixel"].shape
    gm = gpu_ctx.mesh(mesh)
    gp = gpu_ctx.color_target(W, H).upload(np.full((H, W), 0x0A0B0C0D, np.uint32))
    gd = gpu_ctx.depth_target(W, H).upload(z["depth0"])
    gpu_ctx.render(desc, S.uniforms(), gm, gp, gd)
    assert np.array_equal(gp.download(), z["pixel"])
    assert np.array_equal(gd.download().view(np.uint32), z["depth"].view(np.uint32))
