"""A cut-out test scene for the any-hit alpha test of non-opaque geometry
(general.wesl:1057-1066): a camera at the origin looking down +z; a 2 x 2
quad at z = 2 whose 4 x 4 diffuse texture has alpha 0 on the odd checker
texels, mapped once over the quad; behind it an opaque white 6 x 6 quad at
z = 4; a uniform grey sky."""
import numpy as np

from lambent_amd import api, scenes

CHECKER = np.zeros((4, 4, 4), np.uint8)
for _y in range(4):
    for _x in range(4):
        CHECKER[_y, _x] = (220, 40, 30, 255) if (_x + _y) % 2 == 0 else (30, 60, 220, 0)


def cutout_scene(front_type=api.MaterialType.Diffuse, opaque=False, width=48, height=48):
    v = np.array([[-1, -1, 2, 0], [1, -1, 2, 0], [1, 1, 2, 0], [-1, 1, 2, 0],
                  [-3, -3, 4, 0], [3, -3, 4, 0], [3, 3, 4, 0], [-3, 3, 4, 0]], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7], np.uint32)
    E0 = api.EmissionScale.from_emission_scale(0.0)
    ior = api.refractive_indices.WATER if front_type == api.MaterialType.Transparent else None
    # u runs with x, v with y: texel row floor(v * 4) is image row y
    f0 = api.Material.new([0, 0], [1, 0], [1, 1], [0, 0], 0, None, None, E0, ior, front_type)
    f1 = api.Material.new([0, 0], [1, 1], [0, 1], [0, 0], 0, None, None, E0, ior, front_type)
    back = api.Material.new([0, 0], [1, 0], [1, 1], [0, 0], 1, None, None, E0, None,
                            api.MaterialType.Diffuse)
    proj = api.perspective(60.0, width / height, 0.1, 100.0)
    view = api.look_to_rh((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0))
    return scenes.SceneSpec(v, idx, np.array([0, 1, 2, 2], np.uint32),
                            api.materials_array([f0, f1, back]),
                            [CHECKER, scenes._rgba(1, 1, 1)], [], [], None,
                            api.uniform_background(0.3), proj, view, width, height,
                            opaque=opaque)


def expected_front(rays: np.ndarray):
    """Per camera ray (n x 8, oracle.create_rays): whether the front quad's
    texel under it is opaque, and how far (in texels) its texture coordinate
    lies from a texel edge (small margins can round either way)."""
    r = rays.astype(np.float64)
    t = 2.0 / r[:, 6]
    px, py = r[:, 0] + t * r[:, 4], r[:, 1] + t * r[:, 5]
    inside = (np.abs(px) < 1.0) & (np.abs(py) < 1.0)
    u, v = (px + 1.0) / 2.0 * 4.0, (py + 1.0) / 2.0 * 4.0
    tx = np.clip(np.floor(u), 0, 3).astype(int)
    ty = np.clip(np.floor(v), 0, 3).astype(int)
    opaque_texel = CHECKER[ty, tx, 3] > 0
    margin = np.minimum(np.abs(u - np.round(u)), np.abs(v - np.round(v)))
    edge = np.minimum(1.0 - np.abs(px), 1.0 - np.abs(py)) * 2.0  # in texels
    return inside, opaque_texel, np.minimum(margin, edge)
