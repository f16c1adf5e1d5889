"""ctypes binding of libeuchip.so (the C ABI of include/euchip.h).

The library is the only compute path: importing this module never falls back
to anything else, and `lib()` raises if the built library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeuchip.so")

STATUS = {
    0: "OK", 1: "SIZE_MISMATCH", 2: "BAND_ZERO", 3: "INVALID_ARG", 4: "OOM",
    5: "OVER_BUDGET", 6: "HIP_ERROR", 7: "NO_DEVICE", 8: "NOT_READY",
}
FORMAT_RGBA8, FORMAT_F32 = 0, 1


class EuchipError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status
        self.name = STATUS.get(status, str(status))


class PipelineDescC(C.Structure):
    _fields_ = [
        ("shader", C.c_uint32), ("xform", C.c_uint32), ("pixel_write", C.c_uint32),
        ("depth_test", C.c_uint32), ("depth_write", C.c_uint32), ("y_up", C.c_uint32),
        ("z_clip", C.c_uint32), ("z_min", C.c_float), ("z_max", C.c_float),
        ("cull", C.c_uint32), ("aa_msaa", C.c_uint32), ("aa_level", C.c_uint32),
        ("band_policy", C.c_uint32), ("threads", C.c_uint32),
        ("clear_color_enable", C.c_uint32), ("clear_color", C.c_uint32),
        ("clear_depth_enable", C.c_uint32), ("clear_depth", C.c_float),
        ("texture", C.c_void_p), ("shadow", C.c_void_p),
        ("pixel_format", C.c_uint32), ("sampler", C.c_uint32), ("geometry", C.c_uint32), ("primitives", C.c_uint32),
    ]


class MeshDescC(C.Structure):
    _fields_ = [
        ("pos", C.c_void_p), ("pos_comps", C.c_uint32), ("nrm", C.c_void_p), ("col", C.c_void_p),
        ("uv", C.c_void_p), ("rgba8", C.c_void_p), ("n_vertices", C.c_uint64), ("idx", C.c_void_p),
        ("count", C.c_uint64), ("n_instances", C.c_uint32), ("models", C.c_void_p),
    ]


class RequestC(C.Structure):
    _fields_ = [
        ("mesh", C.c_void_p), ("pixel", C.c_void_p), ("depth", C.c_void_p), ("uniforms", C.c_void_p),
        ("texture", C.c_void_p), ("shadow", C.c_void_p),
    ]


class ServeRequestC(C.Structure):
    _fields_ = [
        ("mesh", C.c_void_p), ("texture", C.c_void_p), ("shadow", C.c_void_p), ("uniforms", C.c_void_p),
        ("width", C.c_uint32), ("height", C.c_uint32),
    ]


class CompletionC(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("status", C.c_int), ("pixel", C.c_void_p), ("depth", C.c_void_p),
                ("done_ns", C.c_uint64)]


class ServerStatsC(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
        ("batches", C.c_uint64), ("queued", C.c_uint32), ("in_flight", C.c_uint32),
        ("bytes_in_use", C.c_size_t), ("peak_bytes", C.c_size_t), ("budget", C.c_size_t),
    ]


class KernelStatC(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double)]


# name -> (restype, argtypes); every entry point include/euchip.h declares.
SIGNATURES = {
    "euchip_status_string": (C.c_char_p, [C.c_int]),
    "euchip_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "euchip_ctx_destroy": (C.c_int, [C.c_void_p]),
    "euchip_last_error": (C.c_char_p, [C.c_void_p]),
    "euchip_ctx_stream": (C.c_void_p, [C.c_void_p]),
    "euchip_synchronize": (C.c_int, [C.c_void_p]),
    "euchip_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "euchip_kernel_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "euchip_reset_stats": (None, [C.c_void_p]),
    "euchip_buffer_create": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "euchip_buffer_wrap": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p,
                                     C.POINTER(C.c_void_p)]),
    "euchip_buffer_destroy": (C.c_int, [C.c_void_p]),
    "euchip_buffer_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "euchip_buffer_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "euchip_buffer_clear": (C.c_int, [C.c_void_p, C.c_void_p]),
    "euchip_buffer_device_ptr": (C.c_void_p, [C.c_void_p]),
    "euchip_mesh_create": (C.c_int, [C.c_void_p, C.POINTER(MeshDescC), C.POINTER(C.c_void_p)]),
    "euchip_mesh_destroy": (C.c_int, [C.c_void_p]),
    "euchip_mesh_set_models": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "euchip_render": (C.c_int, [C.c_void_p, C.POINTER(PipelineDescC), C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p]),
    "euchip_render_async": (C.c_int, [C.c_void_p, C.POINTER(PipelineDescC), C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    "euchip_render_noalloc": (C.c_int, [C.c_void_p, C.POINTER(PipelineDescC), C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p]),
    "euchip_ctx_bytes_in_use": (C.c_size_t, [C.c_void_p]),
    "euchip_ctx_set_budget": (C.c_int, [C.c_void_p, C.c_size_t]),
    "euchip_batch_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "euchip_batch_destroy": (C.c_int, [C.c_void_p]),
    "euchip_batch_set": (C.c_int, [C.c_void_p, C.POINTER(PipelineDescC), C.POINTER(RequestC), C.c_uint32]),
    "euchip_batch_set_multi": (C.c_int, [C.c_void_p, C.POINTER(PipelineDescC), C.POINTER(RequestC), C.c_uint32]),
    "euchip_batch_set_tile_width": (C.c_int, [C.c_void_p, C.c_uint32]),
    "euchip_batch_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]),
    "euchip_batch_launch": (C.c_int, [C.c_void_p, C.c_int]),
    "euchip_batch_wait": (C.c_int, [C.c_void_p]),
    "euchip_batch_query": (C.c_int, [C.c_void_p]),
    "euchip_batch_stream": (C.c_void_p, [C.c_void_p]),
    "euchip_batch_scratch_bytes": (C.c_size_t, [C.c_void_p]),
    "euchip_batch_estimate_bytes": (C.c_size_t, [C.POINTER(PipelineDescC), C.POINTER(RequestC), C.c_uint32]),
    "euchip_server_create": (C.c_int, [C.c_void_p, C.POINTER(PipelineDescC), C.c_size_t, C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_void_p)]),
    "euchip_server_destroy": (C.c_int, [C.c_void_p]),
    "euchip_server_submit": (C.c_int, [C.c_void_p, C.POINTER(ServeRequestC), C.POINTER(C.c_uint64)]),
    "euchip_server_submit_desc": (C.c_int, [C.c_void_p, C.POINTER(ServeRequestC), C.POINTER(PipelineDescC),
                                            C.POINTER(C.c_uint64)]),
    "euchip_server_submit_n": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "euchip_server_poll": (C.c_int, [C.c_void_p, C.POINTER(CompletionC), C.c_uint32, C.POINTER(C.c_uint32)]),
    "euchip_server_release": (C.c_int, [C.c_void_p, C.c_uint64]),
    "euchip_server_release_n": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "euchip_clock_ns": (C.c_uint64, []),
    "euchip_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStatsC)]),
}

_lib = None


def lib():
    """Load libeuchip.so; raises FileNotFoundError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -m euc_amd.build` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
