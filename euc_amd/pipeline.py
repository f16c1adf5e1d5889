"""Host-side mirror of euc's `Pipeline` configuration surface.

Names, meanings and defaults follow the reference:
  DepthMode      src/pipeline.rs:11-53   (NONE, LESS_WRITE, GREATER_WRITE, LESS_PASS, GREATER_PASS)
  PixelMode      src/pipeline.rs:55-73   (WRITE default, PASS)
  CoordinateMode src/pipeline.rs:93-163  (OPENGL, VULKAN default, METAL, DIRECTX, without_z_clip)
  AaMode         src/pipeline.rs:102-113 (None, Msaa { level })
  CullMode       src/rasterizer/mod.rs:9-18 (None, Back default, Front)
Band policy PAR is euc's default `par` feature (render_par, pipeline.rs:303-366);
SEQ is render_seq (pipeline.rs:368-394).

Rust shader closures cannot cross an FFI boundary, so a pipeline names one of
the fixed device shaders (`Shader`) and passes its uniforms as a 64-float block.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, replace
from typing import Optional


class Shader(enum.IntEnum):
    FLAT = 0          # provoking RGBA8 colour (examples with a flat colour)
    COLOR = 1         # interpolated Rgba<f32> vertex colour (examples/triangle.rs)
    GOURAUD = 2       # per-vertex lighting, 2 directional lights
    PHONG = 3         # per-fragment lighting, 2 directional lights
    DEPTH_ONLY = 4    # Unit fragment (examples/teapot.rs TeapotShadow)
    TEXTURED = 5      # bilinear RGBA8 texture (examples/texture_mapping.rs)
    PHONG_SHADOW = 6  # examples/teapot.rs Teapot: point light + shadow map
    TEX_SHADOW = 7    # texture x directional light x shadow map


class Ordering(enum.IntEnum):
    NONE = 0   # DepthMode.test == None
    LESS = 1
    EQUAL = 2
    GREATER = 3


class CullMode(enum.IntEnum):
    NONE = 0
    BACK = 1
    FRONT = 2


class BandPolicy(enum.IntEnum):
    PAR = 0   # euc default features: row bands of G = 20000 * 2^L / W rows
    SEQ = 1   # one band covering the whole target


class Xform(enum.IntEnum):
    CLIP = 0  # vertex positions are homogeneous clip coordinates
    MVP = 1   # clip = view_proj * (model[instance] * (p, 1))


class PixelFormat(enum.IntEnum):
    """How the Rgba<f32> shaders' blend packs a pixel into the u32 texel."""
    RGBA8 = 0  # u32::from_le_bytes([r, g, b, a])
    BGRA8 = 1  # u32::from_le_bytes([b, g, r, a]) (examples/teapot.rs:135-140, a window framebuffer)


class Sampler(enum.IntEnum):
    """Texture sampler of TEXTURED / TEX_SHADOW (src/sampler/*.rs): Linear or
    Nearest, bare or wrapped in Clamped / Tiled / Mirrored."""
    DEFAULT = 0          # TEXTURED: Linear; TEX_SHADOW: Clamped<Linear>
    LINEAR = 1
    LINEAR_CLAMPED = 2
    LINEAR_TILED = 3
    LINEAR_MIRRORED = 4
    NEAREST = 5
    NEAREST_CLAMPED = 6
    NEAREST_TILED = 7
    NEAREST_MIRRORED = 8


class Geometry(enum.IntEnum):
    """Fixed Pipeline::geometry stages (src/pipeline.rs:221-229) of a triangle [a, b, c]."""
    IDENTITY = 0    # output([a, b, c]) -- the default
    REVERSED = 1    # output([a, c, b]): the other winding
    BOTH_SIDES = 2  # output([a, b, c]); output([a, c, b])


class Primitives(enum.IntEnum):
    """PrimitiveKind of the vertex stream (src/primitives.rs)."""
    TRIANGLE_LIST = 0
    LINE_LIST = 1
    LINE_TRIANGLE_LIST = 2


@dataclass(frozen=True)
class DepthMode:
    test: Ordering = Ordering.NONE
    write: bool = False

    def uses_depth(self) -> bool:
        return self.test != Ordering.NONE or self.write


DepthMode.NONE = DepthMode(Ordering.NONE, False)
DepthMode.LESS_WRITE = DepthMode(Ordering.LESS, True)
DepthMode.GREATER_WRITE = DepthMode(Ordering.GREATER, True)
DepthMode.LESS_PASS = DepthMode(Ordering.LESS, False)
DepthMode.GREATER_PASS = DepthMode(Ordering.GREATER, False)


@dataclass(frozen=True)
class PixelMode:
    write: bool = True


PixelMode.WRITE = PixelMode(True)
PixelMode.PASS = PixelMode(False)


@dataclass(frozen=True)
class CoordinateMode:
    handedness: str = "left"          # carried for parity; euc never reads it
    y_up: bool = False                # YAxisDirection::Up
    z_clip_range: Optional[tuple] = (0.0, 1.0)

    def without_z_clip(self) -> "CoordinateMode":
        return replace(self, z_clip_range=None)


CoordinateMode.OPENGL = CoordinateMode("right", True, (-1.0, 1.0))
CoordinateMode.VULKAN = CoordinateMode("left", False, (0.0, 1.0))
CoordinateMode.METAL = CoordinateMode("right", False, (0.0, 1.0))
CoordinateMode.DIRECTX = CoordinateMode("left", True, (0.0, 1.0))


@dataclass(frozen=True)
class AaMode:
    msaa: bool = False
    level: int = 0

    @staticmethod
    def Msaa(level: int) -> "AaMode":
        return AaMode(True, level)


AaMode.NONE = AaMode(False, 0)


@dataclass
class PipelineDesc:
    """Flat description of one `Pipeline` impl, as both the HIP path and the
    oracle consume it."""
    shader: int = Shader.FLAT
    xform: int = Xform.MVP
    pixel_write: bool = True
    depth_test: int = Ordering.NONE
    depth_write: bool = False
    y_up: bool = False
    z_clip: bool = True
    z_min: float = 0.0
    z_max: float = 1.0
    cull: int = CullMode.BACK
    aa_msaa: bool = False
    aa_level: int = 0
    band_policy: int = BandPolicy.PAR
    threads: int = 8
    pixel_format: int = PixelFormat.RGBA8
    sampler: int = Sampler.DEFAULT
    geometry: int = Geometry.IDENTITY
    primitives: int = Primitives.TRIANGLE_LIST

    @staticmethod
    def make(shader=Shader.FLAT, xform=Xform.MVP, pixel_mode=PixelMode.WRITE, depth_mode=DepthMode.NONE,
             coordinate_mode=CoordinateMode.VULKAN, cull=CullMode.BACK, aa_mode=AaMode.NONE,
             band_policy=BandPolicy.PAR, threads=8, pixel_format=PixelFormat.RGBA8, sampler=Sampler.DEFAULT,
             geometry=Geometry.IDENTITY, primitives=Primitives.TRIANGLE_LIST) -> "PipelineDesc":
        zc = coordinate_mode.z_clip_range
        return PipelineDesc(
            shader=int(shader), xform=int(xform), pixel_write=pixel_mode.write,
            depth_test=int(depth_mode.test), depth_write=depth_mode.write,
            y_up=coordinate_mode.y_up, z_clip=zc is not None,
            z_min=0.0 if zc is None else float(zc[0]), z_max=0.0 if zc is None else float(zc[1]),
            cull=int(cull), aa_msaa=aa_mode.msaa, aa_level=int(aa_mode.level),
            band_policy=int(band_policy), threads=threads, pixel_format=int(pixel_format), sampler=int(sampler),
            geometry=int(geometry), primitives=int(primitives))

    def msaa_level(self) -> int:
        """pipeline.rs:291-294: level.max(0).min(6), 0 for AaMode::None."""
        return max(0, min(6, self.aa_level)) if self.aa_msaa else 0

    def group_rows(self, width: int) -> int:
        """render_par's band height G = 20000 * 2^L / max(W, 1) (pipeline.rs:328-329)."""
        return 20000 * (1 << self.msaa_level()) // max(width, 1)
