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
               texture: Optional[Buffer2d] = None, shadow: Optional[Buffer2d] = None,
               desc: Optional[PipelineDesc] = None, clear_color: Optional[int] = 0,
               clear_depth: Optional[float] = 1.0) -> int:
        """Queue one request; `desc` (with its clears) renders it with a pipeline
        of its own instead of the server's."""
        u = np.ascontiguousarray(uniforms, dtype=np.float32)
        assert u.size == N_UNIFORMS
        r = ServeRequestC(mesh=mesh.handle, texture=texture.handle if texture else None,
                          shadow=shadow.handle if shadow else None, uniforms=u.ctypes.data,
                          width=int(width), height=int(height))
        t = C.c_uint64()
        if desc is None:
            _check(self.ctx.handle, _lib.lib().euchip_server_submit(self.handle, C.byref(r), C.byref(t)))
        else:
            d = desc_c(desc, None, None, clear_color, clear_depth)
            _check(self.ctx.handle, _lib.lib().euchip_server_submit_desc(self.handle, C.byref(r), C.byref(d),
                                                                          C.byref(t)))
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
    complete.  Requests are submitted in bulk as they arrive (one native call per
    loop turn) and completion times are the server's own (euchip_clock_ns), so
    the loop's cost does not grow with the request count.  keep(i) -> bool selects
    requests whose colour and depth are downloaded into the result; every target
    is released after completion."""
    L = _lib.lib()
    n = len(items)
    reqs = (ServeRequestC * max(n, 1))()
    hold = []
    for i, it in enumerate(items):
        u = np.ascontiguousarray(it.uniforms, dtype=np.float32)
        assert u.size == N_UNIFORMS
        hold.append(u)
        reqs[i] = ServeRequestC(mesh=it.mesh.handle, texture=it.texture.handle if it.texture else None, shadow=None,
                                uniforms=u.ctypes.data, width=int(it.width), height=int(it.height))
    tickets = (C.c_uint64 * max(n, 1))()
    comp = (CompletionC * 1024)()
    rel = (C.c_uint64 * 1024)()
    got = C.c_uint32()
    rsize = C.sizeof(ServeRequestC)
    arrivals = poisson_arrivals(n, rate_hz, seed)
    arr_ns = (arrivals * 1e9).astype(np.int64)
    done_ns = np.zeros(n, np.int64)
    statuses = [None] * n
    outputs = {}
    base = None
    i_next, remaining = 0, n
    t0 = L.euchip_clock_ns()
    while remaining:
        now = L.euchip_clock_ns() - t0
        j = int(np.searchsorted(arr_ns, now, side="right"))
        if j > i_next:
            _check(server.ctx.handle, L.euchip_server_submit_n(server.handle, C.byref(reqs, i_next * rsize), j - i_next,
                                                               C.byref(tickets, i_next * 8)))
            if base is None:
                base = tickets[0]
            i_next = j
        _check(server.ctx.handle, L.euchip_server_poll(server.handle, comp, len(comp), C.byref(got)))
        nrel = 0
        for c in comp[:got.value]:
            i = c.ticket - base
            server.sizes.pop(c.ticket, None)
            done_ns[i] = c.done_ns - t0
            statuses[i] = c.status
            remaining -= 1
            if c.status == 0:
                if keep is not None and keep(i):
                    w, h = items[i].width, items[i].height
                    outputs[i] = (BorrowedBuffer(server.ctx, c.pixel, FORMAT_RGBA8, w, h).download(),
                                  BorrowedBuffer(server.ctx, c.depth, FORMAT_F32, w, h).download())
                rel[nrel] = c.ticket
                nrel += 1
        if nrel:
            _check(server.ctx.handle, L.euchip_server_release_n(server.handle, rel, nrel))
        if L.euchip_clock_ns() - t0 > timeout_s * 1e9:
            raise TimeoutError(f"serving loop: {remaining} requests still pending after {timeout_s} s")
        if i_next < n and remaining == n - i_next:  # nothing in flight: wait for the next arrival
            wait = (arr_ns[i_next] - (L.euchip_clock_ns() - t0)) * 1e-9
            if wait > 0.0005:
                time.sleep(wait - 0.0002)
    lat = (done_ns - arr_ns) * 1e-9
    elapsed = float((done_ns.max() - arr_ns[0]) * 1e-9) if n else 0.0
    return ServeResult(lat, elapsed, statuses, server.stats(), outputs)
