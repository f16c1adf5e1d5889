"""euc_amd: euc 0.6.0's triangle-rasterisation path (Pipeline::render) on MI355X.

The compute path is the HIP library euc_amd/libeuchip.so behind the C ABI of
include/euchip.h; `euc_amd.renderer` is the Python host API over it.
"""
from .pipeline import (AaMode, BandPolicy, CoordinateMode, CullMode, DepthMode, Ordering, PipelineDesc,
                       PixelMode, Shader, Xform)

__all__ = ["AaMode", "BandPolicy", "CoordinateMode", "CullMode", "DepthMode", "Ordering", "PipelineDesc",
           "PixelMode", "Shader", "Xform"]
