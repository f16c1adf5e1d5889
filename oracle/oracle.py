"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path (euc_amd/)
never imports this module.

The oracle restates zesterer/euc 0.6.0's triangle path (src/pipeline.rs,
src/rasterizer/triangles.rs, src/math.rs, src/sampler/*.rs) in C; see
oracle/euc_oracle.c for the per-function citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FLAT, COLOR, GOURAUD, PHONG, DEPTH_ONLY, TEXTURED, PHONG_SHADOW, TEX_SHADOW = range(8)
DEPTH_NONE, DEPTH_LESS, DEPTH_EQUAL, DEPTH_GREATER = range(4)
CULL_NONE, CULL_BACK, CULL_FRONT = range(3)
BANDS_PAR, BANDS_SEQ = range(2)
XFORM_CLIP, XFORM_MVP = range(2)
OK, SIZE_MISMATCH, BAND_ZERO, INVALID = range(4)


class OrcDesc(C.Structure):
    _fields_ = [
        ("shader", C.c_int32), ("xform", C.c_int32), ("pixel_write", C.c_int32),
        ("depth_test", C.c_int32), ("depth_write", C.c_int32), ("y_up", C.c_int32),
        ("z_clip", C.c_int32), ("z_min", C.c_float), ("z_max", C.c_float),
        ("cull", C.c_int32), ("aa_msaa", C.c_int32), ("aa_level", C.c_int32),
        ("band_policy", C.c_int32), ("threads", C.c_int32), ("group_rows", C.c_int32),
        ("pixel_format", C.c_int32), ("sampler", C.c_int32), ("geometry", C.c_int32), ("primitives", C.c_int32),
    ]


class OrcStream(C.Structure):
    _fields_ = [
        ("pos", C.c_void_p), ("pos_comps", C.c_int32), ("nrm", C.c_void_p),
        ("col", C.c_void_p), ("uv", C.c_void_p), ("rgba8", C.c_void_p),
        ("idx", C.c_void_p), ("count", C.c_int64), ("n_vertices", C.c_int64),
        ("n_instances", C.c_int32), ("models", C.c_void_p),
    ]


class OrcTexture(C.Structure):
    _fields_ = [("data", C.c_void_p), ("w", C.c_int32), ("h", C.c_int32), ("is_f32", C.c_int32)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_render_count.restype = C.c_int
        L.orc_render_count.argtypes = [
            C.POINTER(OrcDesc), C.c_void_p, C.POINTER(OrcStream), C.POINTER(OrcTexture),
            C.POINTER(OrcTexture), C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
            C.c_int32, C.POINTER(C.c_int64)]
        L.orc_setup.restype = C.c_int
        L.orc_setup.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                C.c_int32, C.c_float, C.c_float, C.c_int32, C.c_void_p]
        L.orc_walk.restype = C.c_float
        L.orc_walk.argtypes = [C.c_float, C.c_float, C.c_int64]
        L.orc_sample.restype = None
        L.orc_sample.argtypes = [C.POINTER(OrcTexture), C.c_float, C.c_float, C.c_int32, C.c_void_p]
        L.orc_vertex.restype = C.c_int
        L.orc_vertex.argtypes = [C.POINTER(OrcDesc), C.c_void_p, C.POINTER(OrcStream), C.c_int32, C.c_int64,
                                 C.c_void_p, C.c_void_p]
        L.orc_fragment.restype = C.c_int
        L.orc_fragment.argtypes = [C.POINTER(OrcDesc), C.c_void_p, C.POINTER(OrcTexture), C.POINTER(OrcTexture),
                                   C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]
        L.orc_sample_with.restype = None
        L.orc_sample_with.argtypes = [C.POINTER(OrcTexture), C.c_int32, C.c_float, C.c_float, C.c_void_p]
        L.orc_rem_euclid.restype = C.c_float
        L.orc_rem_euclid.argtypes = [C.c_float, C.c_float]
        L.orc_clip_line.restype = C.c_int64
        L.orc_clip_line.argtypes = [C.c_int64] * 8 + [C.c_void_p, C.c_int64]
        L.orc_lighting2.restype = None
        L.orc_lighting2.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def make_desc(p) -> OrcDesc:
    """p: euc_amd.pipeline.PipelineDesc (duck-typed)."""
    return OrcDesc(
        shader=p.shader, xform=p.xform, pixel_write=int(p.pixel_write), depth_test=p.depth_test,
        depth_write=int(p.depth_write), y_up=int(p.y_up), z_clip=int(p.z_clip),
        z_min=float(p.z_min), z_max=float(p.z_max), cull=p.cull, aa_msaa=int(p.aa_msaa),
        aa_level=int(p.aa_level), band_policy=p.band_policy, threads=int(p.threads),
        group_rows=int(getattr(p, "band_rows_override", 0)), pixel_format=int(getattr(p, "pixel_format", 0)),
        sampler=int(getattr(p, "sampler", 0)), geometry=int(getattr(p, "geometry", 0)),
        primitives=int(getattr(p, "primitives", 0)))


def _texture(t):
    if t is None:
        return None, None
    arr = np.ascontiguousarray(t)
    if arr.dtype == np.float32:
        tex = OrcTexture(arr.ctypes.data, arr.shape[1], arr.shape[0], 1)
    else:
        arr = arr.view(np.uint32).reshape(arr.shape[0], -1) if arr.dtype == np.uint8 else arr
        tex = OrcTexture(arr.ctypes.data, arr.shape[1], arr.shape[0], 0)
    return tex, arr


def _stream(mesh, arr):
    pos = arr(mesh.pos, np.float32)
    return OrcStream(
        pos=_ptr(pos), pos_comps=pos.shape[1], nrm=_ptr(arr(mesh.nrm, np.float32)),
        col=_ptr(arr(mesh.col, np.float32)), uv=_ptr(arr(mesh.uv, np.float32)),
        rgba8=_ptr(arr(mesh.rgba8, np.uint32)), idx=_ptr(arr(mesh.idx, np.uint32)),
        count=mesh.count, n_vertices=pos.shape[0], n_instances=mesh.n_instances,
        models=_ptr(arr(mesh.models, np.float32)))


def render(desc, uniforms, mesh, pixel=None, depth=None, texture=None, shadow=None):
    """Run the oracle in place on numpy targets.

    desc: PipelineDesc; uniforms: float32[64]; mesh: euc_amd.scenes.Mesh;
    pixel: uint32[h, w] (RGBA8 packed) or None (Empty); depth: float32[h, w] or None.
    Returns (status, shaded_fragment_count).
    """
    L = lib()
    d = make_desc(desc)
    u = np.ascontiguousarray(uniforms, dtype=np.float32)
    keep = []

    def arr(a, dt):
        if a is None:
            return None
        b = np.ascontiguousarray(a, dtype=dt)
        keep.append(b)
        return b

    s = _stream(mesh, arr)
    tex, ta = _texture(texture)
    sh, sa = _texture(shadow)
    if pixel is not None:
        assert pixel.dtype == np.uint32 and pixel.flags.c_contiguous
    if depth is not None:
        assert depth.dtype == np.float32 and depth.flags.c_contiguous
    frags = C.c_int64(0)
    rc = L.orc_render_count(
        C.byref(d), u.ctypes.data, C.byref(s), C.byref(tex) if tex else None,
        C.byref(sh) if sh else None,
        _ptr(pixel), 0 if pixel is None else pixel.shape[1], 0 if pixel is None else pixel.shape[0],
        _ptr(depth), 0 if depth is None else depth.shape[1], 0 if depth is None else depth.shape[0],
        C.byref(frags))
    del keep, ta, sa
    return rc, frags.value


SETUP_FIELDS = 37


def setup(hom, w, h, by0=0, by1=None, y_up=False, z_clip=True, z_min=0.0, z_max=1.0, cull=CULL_BACK):
    """Triangle setup of triangles.rs:55-171 for one band; returns (kept, dict)."""
    L = lib()
    hom = np.ascontiguousarray(hom, dtype=np.float32).reshape(12)
    out = np.zeros(SETUP_FIELDS, dtype=np.float32)
    ok = L.orc_setup(hom.ctypes.data, w, h, by0, h if by1 is None else by1, int(y_up), int(z_clip),
                     z_min, z_max, cull, out.ctypes.data)
    f = {
        "M": out[0:9].reshape(3, 3), "o": out[9:12], "dx": out[12:15], "dy": out[15:18],
        "bmin": out[18:20].astype(np.int64), "bmax": out[20:22].astype(np.int64),
        "verts_by_y": out[22:28].reshape(3, 2), "hom_z": out[28:31], "allin": bool(out[31]),
        "reversed": bool(out[32]), "culled": bool(out[33]), "winding": out[34],
    }
    return bool(ok), f


def walk(a, d, n):
    return np.float32(lib().orc_walk(float(a), float(d), int(n)))


def sample(texture, u, v, clamped=True):
    tex, arr = _texture(texture)
    out = np.zeros(4, dtype=np.float32)
    lib().orc_sample(C.byref(tex), float(u), float(v), int(clamped), out.ctypes.data)
    return out[:1] if tex.is_f32 else out


def sample_with(texture, kind, u, v):
    tex, arr = _texture(texture)
    out = np.zeros(4, dtype=np.float32)
    lib().orc_sample_with(C.byref(tex), int(kind), float(u), float(v), out.ctypes.data)
    return out[:1] if tex.is_f32 else out


def rem_euclid(x, m):
    return np.float32(lib().orc_rem_euclid(float(x), float(m)))


def lighting2(uniforms, wpos, nrm):
    u = np.ascontiguousarray(uniforms, dtype=np.float32)
    a = np.ascontiguousarray(wpos, dtype=np.float32)
    b = np.ascontiguousarray(nrm, dtype=np.float32)
    out = np.zeros(4, dtype=np.float32)
    lib().orc_lighting2(u.ctypes.data, a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out


def vertex(desc, uniforms, mesh, inst, vi):
    """Pipeline::vertex of the fixed shader: (clip position float32[4], VertexData float32[n])."""
    keep = []

    def arr(a, dt):
        if a is None:
            return None
        b = np.ascontiguousarray(a, dtype=dt)
        keep.append(b)
        return b

    d = make_desc(desc)
    u = np.ascontiguousarray(uniforms, dtype=np.float32)
    s = _stream(mesh, arr)
    pos = np.zeros(4, np.float32)
    vd = np.zeros(12, np.float32)
    n = lib().orc_vertex(C.byref(d), u.ctypes.data, C.byref(s), int(inst), int(vi), pos.ctypes.data, vd.ctypes.data)
    del keep
    return pos, vd[:n]


def fragment(desc, uniforms, vd, texture=None, shadow=None, old_pixel=0):
    """Pipeline::fragment then Pipeline::blend of the fixed shader: (Rgba<f32> float32[4], blended u32)."""
    d = make_desc(desc)
    u = np.ascontiguousarray(uniforms, dtype=np.float32)
    v = np.zeros(12, np.float32)
    v[:len(vd)] = np.asarray(vd, np.float32)
    tex, ta = _texture(texture)
    sh, sa = _texture(shadow)
    out = np.zeros(4, np.float32)
    blended = C.c_uint32(0)
    rc = lib().orc_fragment(C.byref(d), u.ctypes.data, C.byref(tex) if tex else None, C.byref(sh) if sh else None,
                            v.ctypes.data, out.ctypes.data, int(old_pixel), C.byref(blended))
    assert rc == 0
    del ta, sa
    return out, blended.value


def clip_line(p1, p2, rect_min, rect_max):
    """Pixels [(x, y), ...] of the clipped Bresenham line of the lines rasterizer."""
    out = np.zeros(2 * 8192, np.int64)
    n = lib().orc_clip_line(int(p1[0]), int(p1[1]), int(p2[0]), int(p2[1]), int(rect_min[0]), int(rect_min[1]),
                            int(rect_max[0]), int(rect_max[1]), out.ctypes.data, 8192)
    assert n <= 8192
    return [tuple(map(int, out[2 * i:2 * i + 2])) for i in range(n)]
