"""Seeded synthetic scenes for the five measured workloads.

Every input is produced on the host in float32 and handed bit-identically to
both the HIP path and the CPU oracle.  The RNG is SplitMix64 with floats drawn
as (u >> 40) * 2**-24, so the random parts of a scene are exact functions of
the seed on any machine.  Geometry built with sin/cos (tori, spheres, camera
matrices) goes through numpy's float64 libm and is then rounded to float32;
tests that need machine-independent inputs use only the RNG.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive SplitMix64 outputs for the given seed (uint64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def rand_f32(seed: int, n: int) -> np.ndarray:
    """n uniform floats in [0, 1) with 24 random bits each (exact in f32)."""
    u = splitmix64(seed, n)
    return ((u >> np.uint64(40)).astype(np.float64) * 2.0**-24).astype(np.float32)


@dataclass
class Mesh:
    """A vertex stream: element e -> (instance e // count, vertex idx[e % count])."""
    pos: np.ndarray                      # (n, 3) object space or (n, 4) clip space, f32
    count: int                           # stream elements per instance
    nrm: Optional[np.ndarray] = None     # (n, 3) f32
    col: Optional[np.ndarray] = None     # (n, 4) f32 Rgba
    uv: Optional[np.ndarray] = None      # (n, 2) f32
    rgba8: Optional[np.ndarray] = None   # (n,) u32 packed RGBA8
    idx: Optional[np.ndarray] = None     # (count,) u32 or None
    n_instances: int = 1
    models: Optional[np.ndarray] = None  # (n_instances, 16) f32 row-major

    @property
    def n_vertices(self) -> int:
        return int(self.pos.shape[0])

    @property
    def n_triangles(self) -> int:
        return (self.count * self.n_instances) // 3

    def input_bytes(self) -> int:
        b = self.pos.nbytes
        for a in (self.nrm, self.col, self.uv, self.rgba8, self.idx):
            if a is not None:
                b += a.nbytes
        return b


# --------------------------------------------------------------------------- #
# Uniform block layout (64 floats), shared with include/euchip.h
# --------------------------------------------------------------------------- #
U_VP, U_L0, U_L1, U_CAM, U_SURF, U_K, U_LPOS, U_LVP = 0, 16, 20, 24, 28, 32, 36, 40
N_UNIFORMS = 64


def uniforms(view_proj=None, lights=None, cam=(0.0, 0.0, 0.0), surf=(1.0, 0.8, 0.7, 1.0),
             ambient=0.1, k_diffuse=0.5, k_specular=3.0, shadow_bias=0.0001,
             light_pos=(0.0, 0.0, 0.0), light_vp=None) -> np.ndarray:
    u = np.zeros(N_UNIFORMS, dtype=np.float32)
    u[U_VP:U_VP + 16] = np.eye(4, dtype=np.float32).reshape(16) if view_proj is None else \
        np.asarray(view_proj, dtype=np.float32).reshape(16)
    if lights is None:
        lights = [((0.0, 0.0, -1.0), 1.0), ((0.0, 1.0, 0.0), 0.0)]
    for i, (d, inten) in enumerate(lights[:2]):
        d = np.asarray(d, dtype=np.float64)
        d = d / np.linalg.norm(d)
        base = U_L0 if i == 0 else U_L1
        u[base:base + 3] = d.astype(np.float32)
        u[base + 3] = inten
    u[U_CAM:U_CAM + 3] = cam
    u[U_SURF:U_SURF + 4] = surf
    u[U_K:U_K + 4] = (ambient, k_diffuse, k_specular, shadow_bias)
    u[U_LPOS:U_LPOS + 3] = light_pos
    u[U_LVP:U_LVP + 16] = np.eye(4, dtype=np.float32).reshape(16) if light_vp is None else \
        np.asarray(light_vp, dtype=np.float32).reshape(16)
    return u


# --------------------------------------------------------------------------- #
# Matrices (row-major, column vectors), built in float64 and rounded once
# --------------------------------------------------------------------------- #

def perspective_fov_lh_zo(fov_y, width, height, near, far) -> np.ndarray:
    """Left-handed perspective with z in [0, 1] (the matrix the examples use)."""
    h = np.cos(0.5 * fov_y) / np.sin(0.5 * fov_y)
    w = h * height / width
    m = np.zeros((4, 4))
    m[0, 0] = w
    m[1, 1] = h
    m[2, 2] = far / (far - near)
    m[2, 3] = -(far * near) / (far - near)
    m[3, 2] = 1.0
    return m


def perspective_fov_rh_no(fov_y, width, height, near, far) -> np.ndarray:
    """Right-handed perspective with z in [-1, 1] (examples/texture_mapping.rs:123)."""
    h = np.cos(0.5 * fov_y) / np.sin(0.5 * fov_y)
    w = h * height / width
    m = np.zeros((4, 4))
    m[0, 0] = w
    m[1, 1] = h
    m[2, 2] = -(far + near) / (far - near)
    m[2, 3] = -(2.0 * far * near) / (far - near)
    m[3, 2] = -1.0
    return m


def look_at_lh(eye, target, up) -> np.ndarray:
    eye, target, up = (np.asarray(a, dtype=np.float64) for a in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(up, f)
    s /= np.linalg.norm(s)
    u = np.cross(f, s)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, -f @ eye
    return m


def rotation(ax, ay, az) -> np.ndarray:
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    rx = np.array([[1, 0, 0, 0], [0, cx, -sx, 0], [0, sx, cx, 0], [0, 0, 0, 1]])
    ry = np.array([[cy, 0, sy, 0], [0, 1, 0, 0], [-sy, 0, cy, 0], [0, 0, 0, 1]])
    rz = np.array([[cz, -sz, 0, 0], [sz, cz, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])
    return rz @ ry @ rx


def translation(x, y, z) -> np.ndarray:
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def scaling(s) -> np.ndarray:
    m = np.eye(4)
    m[0, 0] = m[1, 1] = m[2, 2] = s
    return m


def f32(m) -> np.ndarray:
    return np.asarray(m, dtype=np.float64).astype(np.float32)


# --------------------------------------------------------------------------- #
# Meshes
# --------------------------------------------------------------------------- #

def torus(R=1.0, r=0.35, nu=256, nv=128):
    """Wrapped, indexed torus; outward faces wind so that CullMode::Back keeps them
    under a left-handed, y-down (VULKAN) projection."""
    u = np.arange(nu) * (2 * np.pi / nu)
    v = np.arange(nv) * (2 * np.pi / nv)
    U, V = np.meshgrid(u, v, indexing="ij")
    cx, sx = np.cos(U), np.sin(U)
    cv, sv = np.cos(V), np.sin(V)
    pos = np.stack([(R + r * cv) * cx, (R + r * cv) * sx, r * sv], -1).reshape(-1, 3)
    nrm = np.stack([cv * cx, cv * sx, sv], -1).reshape(-1, 3)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    tris = np.stack([a, c, b, a, d, c], -1).reshape(-1, 3)
    return f32(pos), f32(nrm), tris.astype(np.uint32).reshape(-1)


def uv_sphere(nu, nv, radius=1.0):
    """Indexed UV sphere with poles as degenerate quads (same winding as torus())."""
    u = np.arange(nu) * (2 * np.pi / nu)
    v = np.linspace(0, np.pi, nv + 1)
    U, V = np.meshgrid(u, v, indexing="ij")
    nrm = np.stack([np.sin(V) * np.cos(U), np.sin(V) * np.sin(U), np.cos(V)], -1).reshape(-1, 3)
    pos = nrm * radius
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = i * (nv + 1) + j
    b = ((i + 1) % nu) * (nv + 1) + j
    c = ((i + 1) % nu) * (nv + 1) + j + 1
    d = i * (nv + 1) + j + 1
    tris = np.stack([a, b, c, a, c, d], -1).reshape(-1, 3)
    return f32(pos), f32(nrm), tris.astype(np.uint32).reshape(-1)


def box(n):
    """Axis-aligned cube [-1,1]^3, each face an n x n grid (2 n^2 tris per face)."""
    ps, ns, ts = [], [], []
    base = 0
    g = np.linspace(-1.0, 1.0, n + 1)
    A, B = np.meshgrid(g, g, indexing="ij")
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            p = np.zeros((n + 1, n + 1, 3))
            p[..., axis] = sgn
            p[..., (axis + 1) % 3] = A
            p[..., (axis + 2) % 3] = B
            nn = np.zeros_like(p)
            nn[..., axis] = sgn
            i = np.arange(n)[:, None]
            j = np.arange(n)[None, :]
            a = base + i * (n + 1) + j
            b = base + (i + 1) * (n + 1) + j
            c = base + (i + 1) * (n + 1) + j + 1
            d = base + i * (n + 1) + j + 1
            if sgn > 0:
                t = np.stack([a, b, c, a, c, d], -1)
            else:
                t = np.stack([a, c, b, a, d, c], -1)
            ps.append(p.reshape(-1, 3))
            ns.append(nn.reshape(-1, 3))
            ts.append(t.reshape(-1))
            base += (n + 1) * (n + 1)
    return f32(np.concatenate(ps)), f32(np.concatenate(ns)), np.concatenate(ts).astype(np.uint32)


# --------------------------------------------------------------------------- #
# Workloads
# --------------------------------------------------------------------------- #

@dataclass
class Scene:
    name: str
    mesh: Mesh
    uniforms: np.ndarray
    width: int
    height: int
    clear_color: int = 0
    clear_depth: float = 1.0
    texture: Optional[np.ndarray] = None   # (h, w) uint32 RGBA8
    extra: dict = field(default_factory=dict)


def plain_scene(seed=0x5EED0000 + 1, n_tris=1024, width=640, height=480) -> Scene:
    """1024 triangles, clip x,y in [-1,1], z in [0,1], w = 1, flat colour per triangle."""
    r = rand_f32(seed, n_tris * 3 * 3 + n_tris).astype(np.float64)
    xyz = r[: n_tris * 9].reshape(-1, 3)
    pos = np.empty((n_tris * 3, 4), dtype=np.float32)
    pos[:, 0] = (xyz[:, 0] * 2.0 - 1.0).astype(np.float32)
    pos[:, 1] = (xyz[:, 1] * 2.0 - 1.0).astype(np.float32)
    pos[:, 2] = xyz[:, 2].astype(np.float32)
    pos[:, 3] = 1.0
    cbits = splitmix64(seed ^ 0xC0105, n_tris).astype(np.uint64)
    col = ((cbits & np.uint64(0x00FFFFFF)) | np.uint64(0xFF000000)).astype(np.uint32)
    rgba8 = np.repeat(col, 3)
    mesh = Mesh(pos=pos, count=n_tris * 3, rgba8=rgba8)
    return Scene("plain", mesh, uniforms(), width, height, 0, 1.0)


def torus_scene(seed=0x5EED0000 + 2, width=1920, height=1080, nu=256, nv=128, instances=4) -> Scene:
    """Procedural torus (nu x nv segments) x `instances` seeded model matrices."""
    pos, nrm, idx = torus(1.0, 0.35, nu, nv)
    r = rand_f32(seed, 8 * instances).astype(np.float64).reshape(instances, 8)
    models = []
    cols = int(np.ceil(np.sqrt(instances)))
    rows = int(np.ceil(instances / cols))
    for k in range(instances):
        gx, gy = k % cols, k // cols
        tx = (gx - (cols - 1) / 2) * 3.0 + (r[k, 0] - 0.5) * 0.4
        ty = (gy - (rows - 1) / 2) * 2.6 + (r[k, 1] - 0.5) * 0.3
        tz = 3.6 + r[k, 2] * 0.8
        rot = rotation(r[k, 3] * 2 * np.pi, r[k, 4] * 2 * np.pi, r[k, 5] * 2 * np.pi)
        models.append(translation(tx, ty, tz) @ rot)
    models = np.stack([f32(m).reshape(16) for m in models])
    vp = perspective_fov_lh_zo(1.3, width, height, 0.01, 100.0)
    u = uniforms(view_proj=f32(vp), lights=[((-0.4, -0.7, -0.6), 1.0), ((0.8, 0.2, -0.5), 0.6)],
                 cam=(0.0, 0.0, 0.0), surf=(1.0, 0.8, 0.7, 1.0), ambient=0.1, k_diffuse=0.5, k_specular=3.0)
    mesh = Mesh(pos=pos, nrm=nrm, idx=idx, count=int(idx.size), n_instances=instances, models=models)
    return Scene("torus", mesh, u, width, height, 0, 1.0)


def teapot_scene(mesh: Mesh, width: int, height: int, shadow_size: int = 512) -> Scene:
    """The scene of the reference's benchmark (benches/teapot.rs:143-205) for a
    loaded teapot mesh: a 512^2 shadow map from a light at (-8, 5, -5) looking at
    the origin (perspective fov 0.75, near 0.1, far 100), then the Phong +
    shadow-lookup colour pass from a camera 4.5 units away (fov 1.3, near 0.01,
    far 100); model = rotation_x(pi) rotation_x(-0.55) rotation_y(-0.25).
    Uniforms of the colour pass (PHONG_SHADOW) in `uniforms`, of the shadow pass
    (DEPTH_ONLY) in extra["shadow_uniforms"]; the mesh carries the model matrix."""
    light_pos = np.array([-8.0, 5.0, -5.0])
    light_vp = perspective_fov_lh_zo(0.75, shadow_size, shadow_size, 0.1, 100.0) @ \
        look_at_lh(light_pos, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    p = perspective_fov_lh_zo(1.3, width, height, 0.01, 100.0)
    v = translation(0.0, 0.0, 4.5)
    m = rotation(np.pi, 0.0, 0.0) @ rotation(-0.55, 0.0, 0.0) @ rotation(0.0, -0.25, 0.0)
    cam = np.linalg.inv(v) @ np.array([0.0, 0.0, 0.0, 1.0])
    u = uniforms(view_proj=f32(p @ v), cam=tuple(cam[:3]), surf=(1.0, 0.8, 0.7, 1.0), ambient=0.1, k_diffuse=0.5,
                 k_specular=3.0, shadow_bias=0.0001, light_pos=tuple(light_pos), light_vp=f32(light_vp))
    tm = Mesh(pos=mesh.pos, nrm=mesh.nrm, uv=mesh.uv, idx=mesh.idx, count=mesh.count, n_instances=1,
              models=f32(m).reshape(1, 16))
    sc = Scene("teapot", tm, u, width, height, 0, 1.0)
    sc.extra["shadow_size"] = shadow_size
    sc.extra["shadow_uniforms"] = uniforms(view_proj=f32(light_vp))
    return sc


CUBE_CORNERS = [(-1, -1, 1), (-1, 1, 1), (1, 1, 1), (1, -1, 1), (-1, -1, -1), (-1, 1, -1), (1, 1, -1), (1, -1, -1),
                (-1, 1, 1), (-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, -1, 1), (-1, -1, -1), (1, -1, -1), (1, -1, 1),
                (1, -1, 1), (1, -1, -1), (1, 1, -1), (1, 1, 1), (-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1)]
CUBE_UVS = [(0, 1), (0, 0), (1, 0), (1, 1), (0, 0), (0, 1), (1, 1), (1, 0), (0, 0), (0, 1), (1, 1), (1, 0),
            (0, 0), (0, 1), (1, 1), (1, 0), (1, 1), (1, 0), (0, 0), (0, 1), (0, 1), (0, 0), (1, 0), (1, 1)]
CUBE_INDICES = [0, 3, 1, 1, 3, 2, 4, 5, 7, 5, 6, 7, 8, 11, 9, 9, 11, 10, 12, 13, 15, 13, 14, 15,
                16, 17, 19, 17, 18, 19, 20, 23, 21, 21, 23, 22]


def texture_cube_scene(texture: np.ndarray, frame: int = 0, width: int = 800, height: int = 600) -> Scene:
    """examples/texture_mapping.rs: a cube of 24 corners and 12 triangles with
    per-face texture coordinates, mvp = p * v * m at animation frame `frame`, the
    colour target cleared to 180 (TEXTURED, bare Linear sampler, no depth test)."""
    p = perspective_fov_rh_no(1.4, width, height, 0.01, 100.0)
    v = translation(0.0, 0.0, -2.0) @ scaling(0.6) @ rotation(0.6, 0.0, 0.0)
    i = float(frame)
    m = rotation(np.sin(i * 0.004) * 0.4, 0.0, 0.0) @ rotation(0.0, i * 0.0008 * 4.0, 0.0) @ \
        rotation(0.0, 0.0, np.cos(i * 0.006) * 0.4)
    pos = np.array([c + (1.0,) for c in CUBE_CORNERS], np.float32)
    uv = np.array(CUBE_UVS, np.float32)
    mesh = Mesh(pos=pos[:, :3].copy(), uv=uv, idx=np.array(CUBE_INDICES, np.uint32), count=len(CUBE_INDICES),
                n_instances=1, models=np.eye(4, dtype=np.float32).reshape(1, 16))
    u = uniforms(view_proj=f32(p @ v @ m))
    return Scene("texture_cube", mesh, u, width, height, 180, 1.0, texture=texture)


def value_noise(seed, size, cells):
    """Bilinear value noise in [0,1), size x size, from a (cells+1)^2 seeded lattice."""
    lat = rand_f32(seed, (cells + 1) * (cells + 1)).astype(np.float64).reshape(cells + 1, cells + 1)
    t = np.arange(size) * (cells / size)
    i = np.floor(t).astype(int)
    f = t - i
    f = f * f * (3 - 2 * f)
    a = lat[i][:, i]
    b = lat[i + 1][:, i]
    c = lat[i][:, i + 1]
    d = lat[i + 1][:, i + 1]
    fy, fx = f[:, None], f[None, :]
    return (a * (1 - fy) + b * fy) * (1 - fx) + (c * (1 - fy) + d * fy) * fx


def noise_texture(seed, size):
    """size x size RGBA8 (packed u32) texture from two octaves of value noise."""
    tex_n = value_noise(seed + 99, size, 32)
    tex_m = value_noise(seed + 98, size, 7)
    rr = (40 + 150 * tex_m + 60 * tex_n).clip(0, 255).astype(np.uint32)
    gg = (90 + 120 * tex_n).clip(0, 255).astype(np.uint32)
    bb = (30 + 80 * (1 - tex_m)).clip(0, 255).astype(np.uint32)
    return (rr | (gg << 8) | (bb << 16) | (np.uint32(255) << 24)).astype(np.uint32)


def heightfield_scene(seed=0x5EED0000 + 3, width=1920, height=1080, grid=317, tex_size=1024,
                      shadow_size=2048) -> Scene:
    """Seeded value-noise heightfield (grid x grid vertices), textured, lit, shadowed."""
    hgt = value_noise(seed, grid, 12) * 0.6 + value_noise(seed + 17, grid, 40) * 0.08
    g = np.linspace(-4.0, 4.0, grid)
    X, Z = np.meshgrid(g, g, indexing="xy")
    pos = np.stack([X, hgt, Z], -1).reshape(-1, 3)
    dx = np.gradient(hgt, g, axis=1)
    dz = np.gradient(hgt, g, axis=0)
    nrm = np.stack([-dx, np.ones_like(hgt), -dz], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    uv = np.stack(np.meshgrid(np.linspace(0, 1, grid), np.linspace(0, 1, grid), indexing="xy"), -1).reshape(-1, 2)
    i = np.arange(grid - 1)[:, None]
    j = np.arange(grid - 1)[None, :]
    a = i * grid + j
    b = i * grid + j + 1
    c = (i + 1) * grid + j + 1
    d = (i + 1) * grid + j
    tris = np.stack([a, b, c, a, c, d], -1).reshape(-1).astype(np.uint32)
    tex = noise_texture(seed, tex_size)
    cam = np.array([0.0, 3.5, -6.5])
    view = look_at_lh(cam, (0.0, 0.0, 0.5), (0.0, 1.0, 0.0))
    vp = perspective_fov_lh_zo(1.1, width, height, 0.05, 100.0) @ view
    sun = np.array([-0.5, 0.8, -0.3])
    sun /= np.linalg.norm(sun)
    lpos = sun * 12.0
    lvp = perspective_fov_lh_zo(0.75, shadow_size, shadow_size, 0.1, 100.0) @ look_at_lh(lpos, (0, 0, 0), (0, 1, 0))
    u = uniforms(view_proj=f32(vp), lights=[(tuple(sun), 1.0), ((0, 1, 0), 0.0)], cam=tuple(cam),
                 ambient=0.25, k_diffuse=0.9, k_specular=0.0, shadow_bias=0.0001,
                 light_pos=tuple(lpos), light_vp=f32(lvp))
    mesh = Mesh(pos=f32(pos), nrm=f32(nrm), uv=f32(uv), idx=tris, count=int(tris.size), n_instances=1,
                models=np.eye(4, dtype=np.float32).reshape(1, 16))
    sc = Scene("heightfield", mesh, u, width, height, 0, 1.0, texture=tex)
    sc.extra["shadow_size"] = shadow_size
    sc.extra["shadow_uniforms"] = uniforms(view_proj=f32(lvp))
    return sc


def icon_request(seed: int, min_tris=200, max_tris=20000, min_side=32, max_side=256):
    """One batched-icon request: a seeded sphere, torus or box with a log-uniform
    triangle count, a square viewport and a model matrix framing it."""
    r = rand_f32(seed, 8).astype(np.float64)
    kind = int(r[0] * 3)
    n = float(np.exp(np.log(min_tris) + r[1] * (np.log(max_tris) - np.log(min_tris))))
    side = int(min_side + r[2] * (max_side - min_side + 1))
    side = min(side, max_side)
    if kind == 0:
        nu = max(3, int(round(np.sqrt(n / 2.0) * 1.4)))
        nv = max(2, int(round(n / (2.0 * nu))))
        pos, nrm, idx = uv_sphere(nu, nv)
    elif kind == 1:
        nu = max(3, int(round(np.sqrt(n) * 1.4)))
        nv = max(3, int(round(n / (2.0 * nu))))
        pos, nrm, idx = torus(1.0, 0.4, nu, nv)
        pos = f32(pos.astype(np.float64) / 1.4)
    else:
        k = max(1, int(round(np.sqrt(n / 12.0))))
        pos, nrm, idx = box(k)
        pos = f32(pos.astype(np.float64) * 0.6)
    rot = rotation(r[3] * 2 * np.pi, r[4] * 2 * np.pi, r[5] * 2 * np.pi)
    model = f32(translation(0.0, 0.0, 3.2) @ rot).reshape(1, 16)
    vp = perspective_fov_lh_zo(0.9, side, side, 0.1, 20.0)
    u = uniforms(view_proj=f32(vp), lights=[((-0.3, -0.6, -0.7), 1.0), ((0.7, 0.3, -0.4), 0.5)])
    mesh = Mesh(pos=pos, nrm=nrm, idx=idx, count=int(idx.size), n_instances=1, models=model)
    return Scene(f"icon{seed:x}", mesh, u, side, side, 0, 1.0)


# --------------------------------------------------------------------------- #
# Serving workload: resident textured meshes, requests with their own camera
# --------------------------------------------------------------------------- #

def textured_mesh(seed: int, n_tris: int) -> Mesh:
    """A sphere or torus of about n_tris triangles with UVs from its parametrisation."""
    r = rand_f32(seed, 2).astype(np.float64)
    if r[0] < 0.5:
        nu = max(3, int(round(np.sqrt(n_tris))))
        nv = max(2, int(round(n_tris / (2.0 * nu))))
        pos, nrm, idx = uv_sphere(nu, nv)
        U, V = np.meshgrid(np.arange(nu) / nu, np.linspace(0, 1, nv + 1), indexing="ij")
    else:
        nu = max(3, int(round(np.sqrt(n_tris) * 1.4)))
        nv = max(3, int(round(n_tris / (2.0 * nu))))
        pos, nrm, idx = torus(1.0, 0.4, nu, nv)
        pos = f32(pos.astype(np.float64) / 1.4)
        U, V = np.meshgrid(np.arange(nu) / nu, np.arange(nv) / nv, indexing="ij")
    uv = f32(np.stack([U * 4.0, V * 2.0], -1).reshape(-1, 2))
    return Mesh(pos=pos, nrm=nrm, uv=uv, idx=idx, count=int(idx.size), n_instances=1,
                models=np.eye(4, dtype=np.float32).reshape(1, 16))


def serving_assets(seed=0x5EED0000 + 5, n_meshes=12, n_textures=6, min_tris=1000, max_tris=500000,
                   tex_sizes=(256, 512, 1024, 2048)):
    """Resident data of the serving workload: meshes with log-uniform triangle
    counts (the first and last at the bounds) and value-noise textures."""
    r = rand_f32(seed, n_meshes).astype(np.float64)
    r[0], r[-1] = 0.0, 1.0
    counts = np.exp(np.log(min_tris) + r * (np.log(max_tris) - np.log(min_tris))).astype(int)
    meshes = [textured_mesh(seed + 1 + i, int(n)) for i, n in enumerate(counts)]
    textures = [noise_texture(seed + 100 + i, tex_sizes[i % len(tex_sizes)]) for i in range(n_textures)]
    return meshes, textures


def serving_request(seed: int, n_meshes: int, n_textures: int, min_side=64, max_side=1024):
    """(mesh index, texture index, uniforms, side): a seeded camera on a sphere
    of radius 3.2 around the mesh, a square target of log-uniform side."""
    r = rand_f32(seed, 6).astype(np.float64)
    side = int(np.exp(np.log(min_side) + r[0] * (np.log(max_side) - np.log(min_side))))
    side = max(min_side, min(max_side, side))
    th, ph = r[1] * 2 * np.pi, (r[2] - 0.5) * 2.6
    eye = 3.2 * np.array([np.cos(ph) * np.cos(th), np.sin(ph), np.cos(ph) * np.sin(th)])
    vp = perspective_fov_lh_zo(0.9, side, side, 0.1, 20.0) @ look_at_lh(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    return int(r[3] * n_meshes) % n_meshes, int(r[4] * n_textures) % n_textures, uniforms(view_proj=f32(vp)), side
