"""Scene construction for the workloads the library is measured and tested on.

* ``cornell_box``  - Cornell box built in code: 5 walls (10 triangles) in the
  unit cube, white 0.73, red (0.65, 0.05, 0.05) at x=0, green (0.12, 0.45, 0.15)
  at x=1, two boxes of 12 triangles rotated +15 / -18 degrees, a 0.25 x 0.25
  ceiling light of emission 17; all diffuse; camera of
  examples/cornell_box/src/main.rs:500-507 (eye (0.5, 0.5, -1), +z, 45 deg).
* ``procedural``   - N triangles with centres uniform in [-10, 10]^3 and vertex
  offsets of length LogNormal(-2, 0.5) in uniform directions; per-triangle
  materials 60% diffuse (albedo U[0.1, 0.9]), 30% rough GGX metal (roughness
  U[0.05, 1]), 10% dielectric (IOR 1.5, optionally a fraction of them
  dispersive: Cauchy A = 1.5, B = 0.004 um^2); white sky; camera at
  (0, 0, -25) looking at the origin, 45 deg, 16:9.
* ``furnace_box``  - closed cube, all faces one material, camera inside: the
  radiance of every path is known in closed form.

Per-triangle albedo and roughness are expressed through the reference API
(lib.rs:132-162): the albedo is a texel of a recolour atlas addressed by
``tex_pos_recolour``, the roughness a texel of an attributes atlas addressed by
the three texture coordinates (all three equal to the texel centre).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import api


@dataclass
class SceneSpec:
    vertices: np.ndarray                  # (n, 4) float32
    indices: Optional[np.ndarray]         # (3m,) uint32 or None
    material_idx: np.ndarray              # (m,) uint32
    materials: np.ndarray                 # (k, 8) uint32 material words
    diffuse: list = field(default_factory=list)      # (h, w, 4) uint8 images
    emission: list = field(default_factory=list)
    attributes: list = field(default_factory=list)
    recolour: Optional[np.ndarray] = None
    background: Optional[np.ndarray] = None          # (6, n, n, 4) uint8
    proj: np.ndarray = None               # 4 columns
    view: np.ndarray = None
    width: int = 256
    height: int = 256
    opaque: bool = True                   # False: any-hit alpha test

    @property
    def num_triangles(self) -> int:
        return len(self.material_idx)

    def camera(self) -> api.Camera:
        return api.Camera.from_proj_view(self.proj, self.view)

    def texture_set(self) -> api.TextureSet:
        return api.TextureSet(self.diffuse, self.emission, self.attributes, self.recolour,
                              self.background)

    def build(self) -> api.Scene:
        """Uploads the scene to the GPU (BVH build included)."""
        tex = self.texture_set()
        scene = api.Scene(self.vertices, self.indices, self.material_idx, self.materials, tex,
                          opaque=self.opaque)
        scene._textures = tex
        return scene


def _u8(v: float) -> int:
    return int(round(min(max(v, 0.0), 1.0) * 255.0))


def _rgba(r, g, b, a=1.0) -> np.ndarray:
    return api.solid_rgba8(_u8(r), _u8(g), _u8(b), _u8(a))


def _box(center, size, angle_deg):
    """8 vertices + 12 outward-wound triangles of a box rotated about y."""
    cx, cy, cz = center
    sx, sy, sz = size
    a = math.radians(angle_deg)
    c, s = math.cos(a), math.sin(a)
    verts = []
    for k in range(8):
        x = (sx / 2) * (1 if k & 1 else -1)
        y = (sy / 2) * (1 if k & 2 else -1)
        z = (sz / 2) * (1 if k & 4 else -1)
        verts.append((cx + c * x + s * z, cy + y, cz - s * x + c * z))
    quads = [(0, 2, 6, 4), (1, 5, 7, 3), (0, 4, 5, 1), (2, 3, 7, 6), (0, 1, 3, 2), (4, 6, 7, 5)]
    tris = []
    for q in quads:
        tris.append((q[0], q[1], q[2]))
        tris.append((q[0], q[2], q[3]))
    return verts, tris


def _default_materials(n_diffuse_textures: int, with_light: bool = True):
    mats = []
    for i in range(n_diffuse_textures):
        mats.append(api.Material.new([0.5, 0.5], [0.5, 0.5], [0.5, 0.5], [0, 0], i, None, None,
                                     api.EmissionScale(0.0, 0.0), None, api.MaterialType.Diffuse))
    return mats


def cornell_box(width: int = 256, height: int = 256) -> SceneSpec:
    verts: list = []
    tris: list = []
    mat_of: list = []

    def quad(p0, p1, p2, p3, m):
        b = len(verts)
        verts.extend([p0, p1, p2, p3])
        tris.extend([(b, b + 1, b + 2), (b, b + 2, b + 3)])
        mat_of.extend([m, m])

    WHITE, RED, GREEN, LIGHT = 0, 1, 2, 3
    quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0), WHITE)   # floor y=0 (normal +y)
    quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1), WHITE)   # ceiling y=1
    quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1), WHITE)   # back z=1
    quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1), RED)     # x=0
    quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0), GREEN)   # x=1
    lo, hi, ly = 0.375, 0.625, 0.999
    quad((lo, ly, lo), (hi, ly, lo), (hi, ly, hi), (lo, ly, hi), LIGHT)
    for center, size, ang in (((0.33, 0.15, 0.37), (0.3, 0.3, 0.3), -18.0),
                              ((0.66, 0.3, 0.63), (0.3, 0.6, 0.3), 15.0)):
        bv, bt = _box(center, size, ang)
        base = len(verts)
        verts.extend(bv)
        tris.extend([(base + a, base + b, base + c) for a, b, c in bt])
        mat_of.extend([WHITE] * len(bt))
    v = np.zeros((len(verts), 4), np.float32)
    v[:, :3] = np.asarray(verts, np.float32)
    idx = np.asarray(tris, np.uint32).reshape(-1)
    diffuse = [_rgba(0.73, 0.73, 0.73), _rgba(0.65, 0.05, 0.05), _rgba(0.12, 0.45, 0.15)]
    emission = [_rgba(1.0, 1.0, 1.0)]
    mats = _default_materials(3)
    mats.append(api.Material.new([0.5, 0.5], [0.5, 0.5], [0.5, 0.5], [0, 0], 0, 0, None,
                                 api.EmissionScale.from_emission_scale(17.0), None,
                                 api.MaterialType.Diffuse))
    proj = api.perspective(45.0, width / height, 0.1, 200.0)
    view = api.look_to_rh((0.5, 0.5, -1.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0))
    return SceneSpec(v, idx, np.asarray(mat_of, np.uint32), api.materials_array(mats), diffuse,
                     emission, [], None, api.uniform_background(0.0), proj, view, width, height)


def _pack_f16_pair(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    lo = np.asarray(a, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
    hi = np.asarray(b, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
    return lo | (hi << 16)


ATLAS = 1024


CAUCHY_A, CAUCHY_B = 1.5, 0.004  # um^2


def cauchy_range(a: float = CAUCHY_A, b: float = CAUCHY_B) -> tuple:
    """(n(760 nm), n(340 nm)) of n = a + b / lambda^2: the IOR-range endpoints
    of a dispersive dielectric (lib.rs:132-143 ior argument)."""
    return a + b / 0.76 ** 2, a + b / 0.34 ** 2


def procedural_materials(kinds: np.ndarray, dispersive: Optional[np.ndarray] = None) -> np.ndarray:
    """Material words for triangle i: its own texel in the 1024 x 1024 atlases.

    Equivalent to Material.new(uv, uv, uv, [x, y], 0, None, attr, scale 0, ior,
    ty) with uv the texel centre; the equivalence is checked by a test.
    ``dispersive`` marks dielectrics whose IOR range is cauchy_range()."""
    n = len(kinds)
    i = np.arange(n, dtype=np.uint64)
    tx = (i % ATLAS).astype(np.uint32)
    ty = (i // ATLAS).astype(np.uint32)
    u = (tx.astype(np.float64) + 0.5) / ATLAS
    v = (ty.astype(np.float64) + 0.5) / ATLAS
    uv = _pack_f16_pair(u.astype(np.float32), v.astype(np.float32))
    words = np.zeros((n, 8), np.uint32)
    words[:, 0] = uv
    words[:, 1] = uv
    words[:, 2] = uv
    words[:, 3] = tx | (ty << 16)
    words[:, 4] = 0 | (0xFFFF << 16)
    attr = np.where(kinds == 1, 0, 0xFFFF).astype(np.uint32)
    words[:, 5] = attr | (kinds.astype(np.uint32) << 16)
    words[:, 6] = 0
    one = _pack_f16_pair(np.float32(1.0), np.float32(1.0))
    glass = _pack_f16_pair(np.float32(1.5), np.float32(1.5))
    words[:, 7] = np.where(kinds == 2, glass, one)
    if dispersive is not None:
        lo, hi = cauchy_range()
        words[:, 7] = np.where((kinds == 2) & dispersive,
                               _pack_f16_pair(np.float32(lo), np.float32(hi)), words[:, 7])
    return words


def procedural(n_tri: int = 1_000_000, seed: int = 2, width: int = 1920,
               height: int = 1080, dispersive: float = 0.0,
               mix: tuple = (0.6, 0.3, 0.1)) -> SceneSpec:
    """``dispersive``: fraction of the dielectrics given the Cauchy IOR range
    (A = 1.5, B = 0.004 um^2); drawn after every other random number, so the
    geometry and the other materials do not depend on it.  ``mix``: fractions
    of diffuse, GGX metal and dielectric triangles."""
    if n_tri > ATLAS * ATLAS:
        raise ValueError("at most 2^20 triangles fit the material atlases")
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(-10.0, 10.0, size=(n_tri, 1, 3))
    dirs = rng.normal(size=(n_tri, 3, 3))
    dirs /= np.linalg.norm(dirs, axis=2, keepdims=True)
    lens = rng.lognormal(mean=-2.0, sigma=0.5, size=(n_tri, 3, 1))
    tri = (centres + lens * dirs).astype(np.float32)
    v = np.zeros((n_tri * 3, 4), np.float32)
    v[:, :3] = tri.reshape(-1, 3)
    u = rng.uniform(size=n_tri)
    c0, c1 = mix[0], mix[0] + mix[1]
    kinds = np.where(u < c0, 0, np.where(u < c1, 1, 2)).astype(np.uint32)
    albedo = rng.uniform(0.1, 0.9, size=(n_tri, 3))
    rough = rng.uniform(0.05, 1.0, size=n_tri)
    recolour = np.zeros((ATLAS * ATLAS, 4), np.uint8)
    rgb = np.rint(albedo * 255.0).astype(np.uint8)
    rgb[kinds == 2] = 255  # clear glass
    recolour[:n_tri, :3] = rgb
    recolour[:, 3] = 255
    attributes = np.zeros((ATLAS * ATLAS, 4), np.uint8)
    r8 = np.rint(rough * 255.0).astype(np.uint8)
    attributes[:n_tri, 0] = r8
    attributes[:n_tri, 1] = r8
    attributes[:, 3] = 0
    disp = rng.uniform(size=n_tri) < dispersive if dispersive > 0.0 else None
    mats = procedural_materials(kinds, disp)
    proj = api.perspective(45.0, width / height, 0.1, 200.0)
    view = api.look_to_rh((0.0, 0.0, -25.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0))
    return SceneSpec(v, None, np.arange(n_tri, dtype=np.uint32), mats, [_rgba(1, 1, 1)], [],
                     [attributes.reshape(ATLAS, ATLAS, 4)], recolour.reshape(ATLAS, ATLAS, 4),
                     api.uniform_background(1.0), proj, view, width, height)


BATCH_RESOLUTIONS = (128, 256, 512, 1024)


def batch_workload(n_scenes: int = 64, seed: int = 4, tri_min: int = 1_000,
                   tri_max: int = 500_000, resolutions=BATCH_RESOLUTIONS) -> list:
    """Uneven independent scenes for the batched tracer: triangle counts
    log-uniform in [tri_min, tri_max], square resolutions drawn uniformly from
    ``resolutions``, 2/3 diffuse and 1/3 GGX metal triangles (no
    dielectrics), each a procedural scene of its own seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = np.exp(rng.uniform(np.log(tri_min), np.log(tri_max), n_scenes)).astype(np.int64)
    res = rng.choice(np.asarray(resolutions), n_scenes)
    return [procedural(int(c), seed=seed * 1000 + i, width=int(r), height=int(r),
                       mix=(2.0 / 3.0, 1.0 / 3.0, 0.0))
            for i, (c, r) in enumerate(zip(counts, res))]


def furnace_box(albedo: float, emission: float, ty: api.MaterialType = api.MaterialType.Diffuse,
                width: int = 64, height: int = 64) -> SceneSpec:
    """Closed unit cube, every face one material (albedo texel round(albedo*255),
    emission texel white scaled by ``emission``), camera at the centre."""
    verts = []
    tris = []
    for q in [((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)), ((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)),
              ((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)), ((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)),
              ((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0)), ((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))]:
        b = len(verts)
        verts.extend(q)
        tris.extend([(b, b + 1, b + 2), (b, b + 2, b + 3)])
    v = np.zeros((len(verts), 4), np.float32)
    v[:, :3] = np.asarray(verts, np.float32)
    mat = api.Material.new([0.5, 0.5], [0.5, 0.5], [0.5, 0.5], [0, 0], 0, 0, None,
                           api.EmissionScale.from_emission_scale(emission), None, ty)
    proj = api.perspective(60.0, width / height, 0.1, 200.0)
    view = api.look_to_rh((0.5, 0.5, 0.5), (0.3, 0.2, 1.0), (0.0, 1.0, 0.0))
    return SceneSpec(v, np.asarray(tris, np.uint32).reshape(-1), np.zeros(len(tris), np.uint32),
                     api.materials_array([mat]), [_rgba(albedo, albedo, albedo)],
                     [_rgba(1.0, 1.0, 1.0)], [], None, api.uniform_background(0.0), proj, view,
                     width, height)
