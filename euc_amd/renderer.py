"""Python host API over the C ABI (libeuchip.so).

`Pipeline` mirrors euc's trait (src/pipeline.rs:171-301): subclasses override
`pixel_mode`, `depth_mode`, `coordinate_mode`, `aa_mode`, `rasterizer_config`
and name a fixed device shader plus its uniforms; `render(vertices, pixel,
depth)` has the reference's meaning, with `Buffer2d` device targets and
`None` standing for `Empty`.  Errors that the reference raises as panics come
back as `EuchipError` with the reference's message.
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Optional

import numpy as np

from . import _lib
from ._lib import EuchipError, MeshDescC, PipelineDescC, KernelStatC, RequestC, FORMAT_F32, FORMAT_RGBA8
from .pipeline import (AaMode, BandPolicy, CoordinateMode, CullMode, DepthMode, PipelineDesc, PixelMode,
                       Shader, Xform)
from .scenes import Mesh, N_UNIFORMS


def _check(ctx_handle, status: int):
    if status != 0:
        L = _lib.lib()
        msg = L.euchip_last_error(ctx_handle).decode() if ctx_handle else ""
        raise EuchipError(status, msg)


class Buffer2d:
    """Device-resident Buffer2d<T> (src/buffer.rs:19-22): RGBA8 packed in uint32, or f32."""

    def __init__(self, ctx: "Context", fmt: int, w: int, h: int, device_ptr: Optional[int] = None):
        L = _lib.lib()
        self.ctx, self.fmt, self.w, self.h = ctx, fmt, int(w), int(h)
        h_ = C.c_void_p()
        if device_ptr is None:
            _check(ctx.handle, L.euchip_buffer_create(ctx.handle, fmt, self.w, self.h, C.byref(h_)))
        else:
            _check(ctx.handle, L.euchip_buffer_wrap(ctx.handle, fmt, self.w, self.h, C.c_void_p(device_ptr),
                                                    C.byref(h_)))
        self.handle = h_.value
        ctx._children.add(self)

    @property
    def dtype(self):
        return np.float32 if self.fmt == FORMAT_F32 else np.uint32

    def size(self):
        return [self.w, self.h]

    def upload(self, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.size == self.w * self.h
        _check(self.ctx.handle, _lib.lib().euchip_buffer_upload(self.handle, a.ctypes.data, a.nbytes))
        return self

    def download(self) -> np.ndarray:
        a = np.empty((self.h, self.w), dtype=self.dtype)
        _check(self.ctx.handle, _lib.lib().euchip_buffer_download(self.handle, a.ctypes.data, a.nbytes))
        return a

    def clear(self, value):
        """Target::clear (src/buffer.rs:213-218)."""
        v = np.array([value], dtype=self.dtype)
        _check(self.ctx.handle, _lib.lib().euchip_buffer_clear(self.handle, v.ctypes.data))
        return self

    def device_ptr(self) -> int:
        return _lib.lib().euchip_buffer_device_ptr(self.handle)

    def close(self):
        if self.handle and self.ctx.handle:
            _lib.lib().euchip_buffer_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuMesh:
    """Device-resident vertex stream (TriangleList over an optional index buffer)."""

    def __init__(self, ctx: "Context", mesh: Mesh):
        self.ctx = ctx
        self.mesh = mesh
        keep = []

        def p(a, dt):
            if a is None:
                return None
            b = np.ascontiguousarray(a, dtype=dt)
            keep.append(b)
            return b.ctypes.data

        d = MeshDescC(
            pos=p(mesh.pos, np.float32), pos_comps=int(mesh.pos.shape[1]), nrm=p(mesh.nrm, np.float32),
            col=p(mesh.col, np.float32), uv=p(mesh.uv, np.float32), rgba8=p(mesh.rgba8, np.uint32),
            n_vertices=mesh.n_vertices, idx=p(mesh.idx, np.uint32), count=int(mesh.count),
            n_instances=int(mesh.n_instances), models=p(mesh.models, np.float32))
        h = C.c_void_p()
        _check(ctx.handle, _lib.lib().euchip_mesh_create(ctx.handle, C.byref(d), C.byref(h)))
        self.handle = h.value
        ctx._children.add(self)

    def set_models(self, models: np.ndarray):
        m = np.ascontiguousarray(models, dtype=np.float32)
        _check(self.ctx.handle, _lib.lib().euchip_mesh_set_models(self.handle, m.ctypes.data, m.shape[0]))

    def close(self):
        if self.handle and self.ctx.handle:
            _lib.lib().euchip_mesh_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def desc_c(desc: PipelineDesc, texture=None, shadow=None, clear_color=None, clear_depth=None) -> PipelineDescC:
    """The C pipeline descriptor of `desc` plus textures and in-frame clears."""
    return PipelineDescC(
        shader=int(desc.shader), xform=int(desc.xform), pixel_write=int(desc.pixel_write),
        depth_test=int(desc.depth_test), depth_write=int(desc.depth_write), y_up=int(desc.y_up),
        z_clip=int(desc.z_clip), z_min=float(desc.z_min), z_max=float(desc.z_max), cull=int(desc.cull),
        aa_msaa=int(desc.aa_msaa), aa_level=max(0, int(desc.aa_level)), band_policy=int(desc.band_policy),
        threads=int(desc.threads),
        clear_color_enable=int(clear_color is not None),
        clear_color=0 if clear_color is None else int(clear_color) & 0xFFFFFFFF,
        clear_depth_enable=int(clear_depth is not None),
        clear_depth=0.0 if clear_depth is None else float(clear_depth),
        texture=texture.handle if texture else None, shadow=shadow.handle if shadow else None,
        pixel_format=int(desc.pixel_format), sampler=int(desc.sampler), geometry=int(desc.geometry),
        primitives=int(desc.primitives))


class Request:
    """One render() of a batch: a mesh into its own targets with its own uniforms."""

    def __init__(self, mesh: "GpuMesh", pixel: Optional["Buffer2d"], depth: Optional["Buffer2d"],
                 uniforms: np.ndarray, texture: Optional["Buffer2d"] = None, shadow: Optional["Buffer2d"] = None):
        self.mesh, self.pixel, self.depth = mesh, pixel, depth
        self.uniforms = np.ascontiguousarray(uniforms, dtype=np.float32)
        assert self.uniforms.size == N_UNIFORMS
        self.texture, self.shadow = texture, shadow


def _requests_c(reqs):
    arr = (RequestC * max(len(reqs), 1))()
    for i, r in enumerate(reqs):
        arr[i] = RequestC(mesh=r.mesh.handle, pixel=r.pixel.handle if r.pixel else None,
                          depth=r.depth.handle if r.depth else None, uniforms=r.uniforms.ctypes.data,
                          texture=r.texture.handle if r.texture else None,
                          shadow=r.shadow.handle if r.shadow else None)
    return arr


def estimate_bytes(desc: PipelineDesc, reqs, texture=None, shadow=None) -> int:
    """Device bytes a batch of `reqs` adds as scratch (the admission estimate)."""
    d = desc_c(desc, texture, shadow)
    return int(_lib.lib().euchip_batch_estimate_bytes(C.byref(d), _requests_c(reqs), len(reqs)))


class Batch:
    """Many render() calls of one pipeline in one setup + one raster launch, on
    the batch's own stream; each request's targets end up as euchip_render of
    that request alone would leave them."""

    def __init__(self, ctx: "Context"):
        self.ctx = ctx
        h = C.c_void_p()
        _check(ctx.handle, _lib.lib().euchip_batch_create(ctx.handle, C.byref(h)))
        self.handle = h.value
        self.requests = []
        ctx._children.add(self)

    def set(self, desc, reqs, texture=None, shadow=None, clear_color=None, clear_depth=None):
        """Replace the requests.  `desc` is one PipelineDesc for all of them, or a
        list of one per request (requests of equal descs still share a launch;
        each distinct desc runs on a stream of its own)."""
        arr = _requests_c(reqs)
        self.requests = list(reqs)  # keeps uniforms and targets alive
        if isinstance(desc, PipelineDesc):
            d = desc_c(desc, texture, shadow, clear_color, clear_depth)
            _check(self.ctx.handle, _lib.lib().euchip_batch_set(self.handle, C.byref(d), arr, len(reqs)))
            return self
        descs = list(desc)
        if len(descs) != len(reqs):
            raise ValueError(f"{len(descs)} descs for {len(reqs)} requests")
        ds = (PipelineDescC * max(len(descs), 1))(*[desc_c(x, texture, shadow, clear_color, clear_depth) for x in descs])
        _check(self.ctx.handle, _lib.lib().euchip_batch_set_multi(self.handle, ds, arr, len(reqs)))
        return self

    def set_tile_width(self, w: int):
        _check(self.ctx.handle, _lib.lib().euchip_batch_set_tile_width(self.handle, int(w)))

    def autotune(self, reps: int = 3):
        """Time tile widths 8, 16, 32; keep the fastest.  Returns (width, {width: ms})."""
        best = C.c_uint32()
        times = (C.c_float * 3)()
        _check(self.ctx.handle, _lib.lib().euchip_batch_autotune(self.handle, int(reps), C.byref(best), times))
        return int(best.value), {8: times[0], 16: times[1], 32: times[2]}

    def launch(self, graph: bool = False):
        _check(self.ctx.handle, _lib.lib().euchip_batch_launch(self.handle, int(graph)))

    def wait(self):
        _check(self.ctx.handle, _lib.lib().euchip_batch_wait(self.handle))

    def done(self) -> bool:
        s = _lib.lib().euchip_batch_query(self.handle)
        if s == 8:
            return False
        _check(self.ctx.handle, s)
        return True

    def run(self, graph: bool = False):
        self.launch(graph)
        self.wait()

    def stream(self) -> int:
        return _lib.lib().euchip_batch_stream(self.handle)

    def scratch_bytes(self) -> int:
        return int(_lib.lib().euchip_batch_scratch_bytes(self.handle))

    def close(self):
        if self.handle and self.ctx.handle:
            _lib.lib().euchip_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One device context (one HIP stream)."""

    def __init__(self, device: int = 0):
        L = _lib.lib()
        h = C.c_void_p()
        _check(None, L.euchip_ctx_create(device, C.byref(h)))
        self.handle = h.value
        self._children = weakref.WeakSet()

    def buffer(self, fmt: str, w: int, h: int) -> Buffer2d:
        return Buffer2d(self, FORMAT_RGBA8 if fmt == "rgba8" else FORMAT_F32, w, h)

    def color_target(self, w, h) -> Buffer2d:
        return Buffer2d(self, FORMAT_RGBA8, w, h)

    def depth_target(self, w, h) -> Buffer2d:
        return Buffer2d(self, FORMAT_F32, w, h)

    def texture(self, texels: np.ndarray) -> Buffer2d:
        a = np.ascontiguousarray(texels)
        fmt = FORMAT_F32 if a.dtype == np.float32 else FORMAT_RGBA8
        return Buffer2d(self, fmt, a.shape[1], a.shape[0]).upload(a)

    def mesh(self, mesh: Mesh) -> GpuMesh:
        return GpuMesh(self, mesh)

    def render(self, desc: PipelineDesc, uniforms: np.ndarray, mesh: GpuMesh, pixel: Optional[Buffer2d] = None,
               depth: Optional[Buffer2d] = None, texture: Optional[Buffer2d] = None,
               shadow: Optional[Buffer2d] = None, clear_color: Optional[int] = None,
               clear_depth: Optional[float] = None, sync: bool = True):
        u = np.ascontiguousarray(uniforms, dtype=np.float32)
        assert u.size == N_UNIFORMS
        d = desc_c(desc, texture, shadow, clear_color, clear_depth)
        L = _lib.lib()
        f = L.euchip_render if sync else L.euchip_render_async
        _check(self.handle, f(self.handle, C.byref(d), u.ctypes.data, mesh.handle,
                              pixel.handle if pixel else None, depth.handle if depth else None))

    def synchronize(self):
        _check(self.handle, _lib.lib().euchip_synchronize(self.handle))

    def bytes_in_use(self) -> int:
        """Device bytes held through this context (targets, meshes, scratch)."""
        return int(_lib.lib().euchip_ctx_bytes_in_use(self.handle))

    def set_budget(self, nbytes: int):
        """Device-memory budget (0 = none): allocations past it fail with OVER_BUDGET."""
        _check(self.handle, _lib.lib().euchip_ctx_set_budget(self.handle, int(nbytes)))

    def batch(self) -> "Batch":
        return Batch(self)

    def stream(self) -> int:
        return _lib.lib().euchip_ctx_stream(self.handle)

    def set_profiling(self, on: bool):
        _check(self.handle, _lib.lib().euchip_set_profiling(self.handle, int(on)))

    def kernel_stats(self) -> dict:
        L = _lib.lib()
        n = L.euchip_kernel_stats(self.handle, None, 0)
        arr = (KernelStatC * max(n, 1))()
        L.euchip_kernel_stats(self.handle, arr, n)
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(n)}

    def reset_stats(self):
        _lib.lib().euchip_reset_stats(self.handle)

    def close(self):
        """Destroy every buffer and mesh of this context, then the context."""
        if self.handle:
            kids = list(self._children)
            for c in kids:  # batches and servers first: they use the buffers and meshes
                if not isinstance(c, (Buffer2d, GpuMesh)):
                    c.close()
            for c in kids:
                c.close()
            _lib.lib().euchip_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pipeline:
    """euc's `Pipeline` trait with a fixed device shader (src/pipeline.rs:171-301)."""

    shader: int = Shader.FLAT
    xform: int = Xform.MVP
    band_policy: int = BandPolicy.PAR
    threads: int = 8

    def pixel_mode(self) -> PixelMode:
        return PixelMode.WRITE

    def depth_mode(self) -> DepthMode:
        return DepthMode.NONE

    def coordinate_mode(self) -> CoordinateMode:
        return CoordinateMode.VULKAN

    def aa_mode(self) -> AaMode:
        return AaMode.NONE

    def rasterizer_config(self) -> CullMode:
        return CullMode.BACK

    def uniforms(self) -> np.ndarray:
        return np.zeros(N_UNIFORMS, dtype=np.float32)

    def texture(self) -> Optional[Buffer2d]:
        return None

    def shadow(self) -> Optional[Buffer2d]:
        return None

    def desc(self) -> PipelineDesc:
        return PipelineDesc.make(self.shader, self.xform, self.pixel_mode(), self.depth_mode(),
                                 self.coordinate_mode(), self.rasterizer_config(), self.aa_mode(),
                                 self.band_policy, self.threads)

    def render(self, vertices: GpuMesh, pixel: Optional[Buffer2d], depth: Optional[Buffer2d]):
        """Pipeline::render (src/pipeline.rs:248-300); `None` is `Empty`."""
        vertices.ctx.render(self.desc(), self.uniforms(), vertices, pixel, depth, self.texture(), self.shadow())
