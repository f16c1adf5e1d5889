"""The on-disk input loaders (euc_amd/io.py) -- CPU only.

PNG: the reference's example images and PNGs of every 8-bit colour type written
by Pillow decode to the same RGBA8 texels as Pillow's own decoder (an
independent implementation of the format).  OBJ: the teapot's faces come out in
file order with correctly rounded f32 coordinates; polygons split into fans and
negative indices count from the end, as the format defines.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

from euc_amd import io

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")


def pil_rgba(path):
    from PIL import Image
    a = np.asarray(Image.open(path).convert("RGBA"), dtype=np.uint8)
    return np.ascontiguousarray(a).view(np.uint32).reshape(a.shape[:2])


@pytest.mark.parametrize("name", ["rust.png", "checkerboard.png"])
def test_reference_pngs_decode_like_pillow(name):
    path = os.path.join(DATA, name)
    a = io.load_png(path)
    assert np.array_equal(a, pil_rgba(path))


@pytest.mark.parametrize("mode", ["L", "LA", "RGB", "RGBA", "P"])
def test_png_colour_types(tmp_path, mode):
    from PIL import Image
    rng = np.random.default_rng(len(mode))
    h, w = 37, 53
    base = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    base[:, :w // 2] = base[:, :1]  # smooth and noisy halves: the encoder picks several filters
    img = Image.fromarray(base, "RGBA")
    if mode == "P":
        img = img.convert("RGB").quantize(colors=40)
    else:
        img = img.convert(mode)
    path = str(tmp_path / f"t_{mode}.png")
    img.save(path, optimize=False)
    assert np.array_equal(io.load_png(path), pil_rgba(path))


def test_teapot_obj():
    m = io.load_obj(os.path.join(DATA, "teapot.obj"))
    assert m.count == 3 * 2256 and m.pos.shape == (6768, 3) and m.nrm.shape == (6768, 3) and m.uv is None
    # the first face is "f 1//1 3//2 2//3": v 1 = (0.605903, 1.343409, 0)
    assert m.pos[0].tolist() == [np.float32(0.605903), np.float32(1.343409), 0.0]
    assert float(m.pos[1][0]) == float(np.float32(0.584584))


def test_f32_parse_is_correctly_rounded():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        s = f"{rng.uniform(-10, 10):.{int(rng.integers(1, 12))}f}"
        f = io._f32(s)
        exact = Fraction(s)
        err = abs(Fraction(float(f)) - exact)
        for g in (np.nextafter(f, np.float32(np.inf)), np.nextafter(f, np.float32(-np.inf))):
            assert err <= abs(Fraction(float(g)) - exact)
    # a halfway case rounds to the even mantissa: 1 + 2^-24 lies between 1 and 1 + 2^-23
    assert io._f32(str(Fraction(1) + Fraction(1, 2 ** 24))) == np.float32(1.0)
    assert io._f32("16777217") == np.float32(16777216.0)


def test_obj_fans_and_negative_indices(tmp_path):
    path = tmp_path / "q.obj"
    path.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 1\nf 1/1 2/2 3/1 4/2\nf -1/-1 -2/-2 -4/-1\n")
    m = io.load_obj(str(path))
    assert m.count == 9 and m.nrm is None
    assert m.pos.tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 0, 0], [1, 1, 0], [0, 1, 0],
                              [0, 1, 0], [1, 1, 0], [0, 0, 0]]
    assert m.uv.tolist() == [[0, 0], [1, 1], [0, 0], [0, 0], [0, 0], [1, 1], [1, 1], [0, 0], [1, 1]]
