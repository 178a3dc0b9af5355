"""The scene of the reference's own renderer test (src/tests/mod.rs:49-160,
319-371, 459-473, 655-790), rebuilt from its data: three faces of the unit
cube (8 triangles, one material each: diffuse with DIAMOND IOR, plain
diffuse, metallic, WATER transparent), its five texture PNGs and its six
64 x 64 background faces, which tests/golden/ref_textures holds as copied
fixtures.  Camera: perspective 60 deg, near 0.1, far 200, looking from
(0.5, 0.8, 0.5) at the origin (the test's first frame); the frame number is
the seed."""
from pathlib import Path

import numpy as np

from lambent_amd import api, scenes

TEX = Path(__file__).resolve().parent / "golden" / "ref_textures"


def png(name: str) -> np.ndarray:
    from PIL import Image
    return np.asarray(Image.open(TEX / name).convert("RGBA"), np.uint8).copy()


def reference_test_scene(width: int = 320, height: int = 180) -> scenes.SceneSpec:
    P = [[1.0, 1.0], [0.0, 1.0], [1.0, 0.0], [0.0, 0.0]]
    M = api.Material.new
    E1 = api.EmissionScale.from_emission_scale(1.0)
    T = api.MaterialType
    ri = api.refractive_indices
    mats = [
        M(P[0], P[1], P[2], [0, 0], 0, 0, 0, E1, ri.DIAMOND, T.Diffuse),        # mod.rs:51
        M(P[3], P[2], P[1], [0, 0], 0, 0, 0, E1, ri.DIAMOND, T.Diffuse),        # mod.rs:63
        M(P[0], P[2], P[1], [0, 0], 1, 1, None, E1, None, T.Diffuse),           # mod.rs:75
        M(P[1], P[3], P[2], [0, 0], 1, 1, None, E1, None, T.Diffuse),           # mod.rs:87
        M(P[2], P[3], P[0], [0, 0], 0, 0, 0, E1, None, T.Metallic),             # mod.rs:99
        M(P[1], P[0], P[3], [0, 0], 0, 0, 0, E1, None, T.Metallic),             # mod.rs:111
        M(P[0], P[1], P[2], [0, 0], 1, None, None, E1, ri.WATER, T.Transparent),  # mod.rs:123
        M(P[1], P[3], P[2], [0, 0], 1, None, None, E1, ri.WATER, T.Transparent),  # mod.rs:135
    ]
    verts = np.array([[x, y, z, 0.0] for x in (0.0, 1.0) for y in (0.0, 1.0) for z in (0.0, 1.0)],
                     np.float32)                                                # mod.rs:150-158
    idx = np.array([0, 1, 2, 3, 2, 1, 0, 4, 1, 1, 5, 4, 0, 2, 4, 6, 4, 2, 2, 3, 6, 7, 6, 3],
                   np.uint32)                                                   # mod.rs:160-162
    background = np.stack([png(f"{s}{a}.png") for a in "xyz" for s in ("plus_", "minus_")])
    proj = api.perspective(60.0, width / height, 0.1, 200.0)
    view = api.look_at_rh((0.5, 0.8, 0.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    return scenes.SceneSpec(verts, idx, np.arange(8, dtype=np.uint32), api.materials_array(mats),
                            [png("diffuse.png"), png("pure_white.png")],
                            [png("emission_partial.png"), png("emission.png")],
                            [png("attributes.png")], None, background, proj, view, width, height)
