"""On-disk inputs of the reference's examples: Wavefront OBJ meshes and PNG textures.

`load_obj` plays the role of `wavefront::Obj::vertices()` feeding
`Pipeline::render` in examples/teapot.rs:157 and benches/teapot.rs:150-151: the
triangles of every face in file order, each corner carrying its position and
(when the face names them) its texture coordinate and normal.  Faces with more
than three corners are split into a fan (0, i, i + 1).  Coordinates are parsed
to the nearest f32 (correctly rounded, as Rust's `str::parse::<f32>`).

`load_png` plays `image::open(..).to_rgba8()` followed by
`Buffer2d::from_texture` (src/buffer.rs:32-58): an 8-bit PNG decoded to RGBA8
texels in row-major order (x + w*y), packed little-endian into uint32 as the
device textures hold them.  Supported: bit depth 8, colour types grey, RGB,
palette (with tRNS), grey + alpha and RGBA, no interlacing; anything else raises.
"""
from __future__ import annotations

import struct
import zlib
from fractions import Fraction

import numpy as np

from .scenes import Mesh


def _f32(text: str) -> np.float32:
    """The f32 nearest to the decimal `text` (ties to even)."""
    exact = Fraction(text)
    f = np.float32(float(exact))
    if not np.isfinite(f):
        return f
    best, best_err = f, abs(Fraction(float(f)) - exact)
    for g in (np.nextafter(f, np.float32(np.inf)), np.nextafter(f, np.float32(-np.inf))):
        err = abs(Fraction(float(g)) - exact)
        if err < best_err or (err == best_err and (int(g.view(np.uint32)) & 1) == 0):
            best, best_err = g, err
    return np.float32(best)


def load_obj(path: str) -> Mesh:
    """The corners of every triangle of the OBJ file at `path`, in file order,
    as a non-indexed Mesh (pos: object-space xyz, nrm and uv when present)."""
    v, vt, vn = [], [], []
    corners = []  # (v, vt or None, vn or None), 0-based
    with open(path, "r") as f:
        for line in f:
            parts = line.split()
            if not parts or parts[0].startswith("#"):
                continue
            tag = parts[0]
            if tag == "v":
                v.append([_f32(x) for x in parts[1:4]])
            elif tag == "vt":
                vt.append([_f32(x) for x in parts[1:3]])
            elif tag == "vn":
                vn.append([_f32(x) for x in parts[1:4]])
            elif tag == "f":
                face = []
                for c in parts[1:]:
                    f3 = (c.split("/") + ["", ""])[:3]

                    def ix(s, n):
                        if not s:
                            return None
                        i = int(s)
                        return i - 1 if i > 0 else n + i

                    face.append((ix(f3[0], len(v)), ix(f3[1], len(vt)), ix(f3[2], len(vn))))
                for i in range(1, len(face) - 1):
                    corners += [face[0], face[i], face[i + 1]]
    if not corners:
        raise ValueError(f"{path}: no faces")
    pos = np.array([v[c[0]] for c in corners], np.float32)
    nrm = np.array([vn[c[2]] for c in corners], np.float32) if all(c[2] is not None for c in corners) else None
    uv = np.array([vt[c[1]] for c in corners], np.float32) if all(c[1] is not None for c in corners) else None
    return Mesh(pos=pos, nrm=nrm, uv=uv, count=len(corners), n_instances=1,
                models=np.eye(4, dtype=np.float32).reshape(1, 16))


def _unfilter(raw: bytes, w: int, h: int, bpp: int) -> np.ndarray:
    """PNG scanline filters (None, Sub, Up, Average, Paeth) undone."""
    stride = w * bpp
    out = np.zeros((h, stride), np.uint8)
    prev = [0] * stride
    pos = 0
    for y in range(h):
        ft = raw[pos]
        line = raw[pos + 1:pos + 1 + stride]
        pos += 1 + stride
        if ft == 0:
            cur = list(line)
        elif ft == 2:
            cur = [(a + b) & 0xFF for a, b in zip(line, prev)]
        elif ft in (1, 3, 4):
            cur = [0] * stride
            for x in range(stride):
                a = cur[x - bpp] if x >= bpp else 0
                if ft == 1:
                    p = a
                elif ft == 3:
                    p = (a + prev[x]) >> 1
                else:
                    b = prev[x]
                    c = prev[x - bpp] if x >= bpp else 0
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                cur[x] = (line[x] + p) & 0xFF
        else:
            raise ValueError(f"bad PNG filter type {ft}")
        out[y] = cur
        prev = cur
    return out


def load_png(path: str) -> np.ndarray:
    """RGBA8 texels (h, w) uint32 of the 8-bit PNG at `path`."""
    data = open(path, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, plte, trns, hdr = 8, [], None, None, None
    while pos < len(data):
        n, tag = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if tag == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif tag == b"PLTE":
            plte = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif tag == b"tRNS":
            trns = body
        elif tag == b"IDAT":
            idat.append(body)
        elif tag == b"IEND":
            break
    w, h, depth, ctype, _, _, interlace = hdr
    if depth != 8 or interlace != 0:
        raise ValueError(f"{path}: only 8-bit, non-interlaced PNGs are supported (depth {depth}, interlace {interlace})")
    bpp = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    px = _unfilter(zlib.decompress(b"".join(idat)), w, h, bpp).reshape(h, w, bpp)
    rgba = np.empty((h, w, 4), np.uint8)
    if ctype == 6:
        rgba[:] = px
    elif ctype == 2:
        rgba[..., :3], rgba[..., 3] = px, 255
    elif ctype == 0:
        rgba[..., :3], rgba[..., 3] = px[..., :1], 255
    elif ctype == 4:
        rgba[..., :3], rgba[..., 3] = px[..., :1], px[..., 1]
    else:
        alpha = np.full(len(plte), 255, np.uint8)
        if trns:
            alpha[:len(trns)] = np.frombuffer(trns, np.uint8)
        rgba[..., :3], rgba[..., 3] = plte[px[..., 0]], alpha[px[..., 0]]
    return np.ascontiguousarray(rgba).view(np.uint32).reshape(h, w)
