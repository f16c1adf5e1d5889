"""Serving loop over the C ABI's euchip_server_* entry points.

The native server (euc_amd/csrc/server.cpp) keeps the queue, admits requests
into batch slots by device-memory budget and overlaps the slots' batches on
the device.  This module drives it: `Server` wraps the handle, `run_poisson`
submits a request list at seeded Poisson arrival times and polls until every
request has completed, recording each request's latency (completion seen by
the host minus arrival).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _lib
from ._lib import CompletionC, EuchipError, ServeRequestC, ServerStatsC, FORMAT_F32, FORMAT_RGBA8
from .pipeline import PipelineDesc
from .renderer import Context, GpuMesh, Buffer2d, _check, desc_c
from .scenes import N_UNIFORMS


class BorrowedBuffer:
    """A target the server owns, readable until its ticket is released."""

    def __init__(self, ctx: Context, handle: int, fmt: int, w: int, h: int):
        self.ctx, self.handle, self.fmt, self.w, self.h = ctx, handle, fmt, w, h

    def download(self) -> np.ndarray:
        dt = np.float32 if self.fmt == FORMAT_F32 else np.uint32
        a = np.empty((self.h, self.w), dtype=dt)
        _check(self.ctx.handle, _lib.lib().euchip_buffer_download(self.handle, a.ctypes.data, a.nbytes))
        return a


@dataclass
class Completion:
    ticket: int
    status: int
    pixel: Optional[BorrowedBuffer]
    depth: Optional[BorrowedBuffer]


class Server:
    """Continuous-batching render server on one context."""

    def __init__(self, ctx: Context, desc: PipelineDesc, budget_bytes: int = 0, max_batch: int = 64,
                 slots: int = 2, texture: Optional[Buffer2d] = None, shadow: Optional[Buffer2d] = None,
                 clear_color: Optional[int] = 0, clear_depth: Optional[float] = 1.0):
        self.ctx = ctx
        d = desc_c(desc, texture, shadow, clear_color, clear_depth)
        h = C.c_void_p()
        _check(ctx.handle, _lib.lib().euchip_server_create(ctx.handle, C.byref(d), int(budget_bytes), int(max_batch),
                                                           int(slots), C.byref(h)))
        self.handle = h.value
        self.sizes = {}
        self._buf = (CompletionC * 512)()
        ctx._children.add(self)

    def submit(self, mesh: GpuMesh, uniforms: np.ndarray, width: int, height: int,
               texture: Optional[Buffer2d] = None, shadow: Optional[Buffer2d] = None) -> int:
        u = np.ascontiguousarray(uniforms, dtype=np.float32)
        assert u.size == N_UNIFORMS
        r = ServeRequestC(mesh=mesh.handle, texture=texture.handle if texture else None,
                          shadow=shadow.handle if shadow else None, uniforms=u.ctypes.data,
                          width=int(width), height=int(height))
        t = C.c_uint64()
        _check(self.ctx.handle, _lib.lib().euchip_server_submit(self.handle, C.byref(r), C.byref(t)))
        self.sizes[t.value] = (int(width), int(height))
        return t.value

    def poll(self) -> list:
        n = C.c_uint32()
        out = []
        while True:
            _check(self.ctx.handle, _lib.lib().euchip_server_poll(self.handle, self._buf, len(self._buf), C.byref(n)))
            for i in range(n.value):
                c = self._buf[i]
                w, h = self.sizes.pop(c.ticket)
                px = BorrowedBuffer(self.ctx, c.pixel, FORMAT_RGBA8, w, h) if c.pixel else None
                dp = BorrowedBuffer(self.ctx, c.depth, FORMAT_F32, w, h) if c.depth else None
                out.append(Completion(c.ticket, c.status, px, dp))
            if n.value < len(self._buf):
                return out

    def release(self, ticket: int):
        _check(self.ctx.handle, _lib.lib().euchip_server_release(self.handle, int(ticket)))

    def stats(self) -> dict:
        s = ServerStatsC()
        _check(self.ctx.handle, _lib.lib().euchip_server_get_stats(self.handle, C.byref(s)))
        return {k: int(getattr(s, k)) for k, _ in ServerStatsC._fields_}

    def close(self):
        if self.handle and self.ctx.handle:
            _lib.lib().euchip_server_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ServeItem:
    mesh: GpuMesh
    uniforms: np.ndarray
    width: int
    height: int
    texture: Optional[Buffer2d] = None


@dataclass
class ServeResult:
    latency_s: np.ndarray          # per request, in submission order
    elapsed_s: float               # first arrival to last completion
    statuses: list
    stats: dict
    outputs: dict = field(default_factory=dict)   # request index -> (colour, depth) when kept


def poisson_arrivals(n: int, rate_hz: float, seed: int) -> np.ndarray:
    """Arrival times (s) of a Poisson process of the given rate, seeded."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_hz, n))


def run_poisson(server: Server, items: list, rate_hz: float, seed: int = 1, keep: Optional[Callable] = None,
                timeout_s: float = 120.0) -> ServeResult:
    """Submit `items` at Poisson arrival times of rate `rate_hz` and poll until all
    complete.  keep(i) -> bool selects requests whose colour and depth are
    downloaded into the result; every target is released after completion."""
    arrivals = poisson_arrivals(len(items), rate_hz, seed)
    n = len(items)
    ticket_of = {}
    done_at = np.full(n, np.nan)
    statuses = [None] * n
    outputs = {}
    i_next = 0
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        while i_next < n and arrivals[i_next] <= now:
            it = items[i_next]
            ticket_of[server.submit(it.mesh, it.uniforms, it.width, it.height, it.texture)] = i_next
            i_next += 1
        for c in server.poll():
            i = ticket_of.pop(c.ticket)
            done_at[i] = time.perf_counter() - t0
            statuses[i] = c.status
            if c.status == 0 and keep is not None and keep(i):
                outputs[i] = (c.pixel.download(), c.depth.download())
            if c.status == 0:
                server.release(c.ticket)
        if i_next == n and not ticket_of:
            break
        if time.perf_counter() - t0 > timeout_s:
            raise TimeoutError(f"serving loop: {len(ticket_of)} requests still pending after {timeout_s} s")
        if i_next < n and not ticket_of:
            wait = arrivals[i_next] - (time.perf_counter() - t0)
            if wait > 0.0005:
                time.sleep(wait - 0.0002)
    lat = done_at - arrivals
    elapsed = float(np.nanmax(done_at) - arrivals[0]) if n else 0.0
    return ServeResult(lat, elapsed, statuses, server.stats(), outputs)
