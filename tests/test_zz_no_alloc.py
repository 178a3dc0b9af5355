"""lmb_scene_intersect_device_noalloc (static scratch, no allocation) against
lmb_scene_intersect_device on the same device rays: bit-identical hits on two
calls in a row, the second with fewer rays, on a procedural soup and on a
scene whose BVH is deeper than the 8-entry LDS stack (so the stack spill in
the static scratch is used), with the deep scene's hits also checked against
the oracle's brute force.
"""
import ctypes

import numpy as np
import pytest

import oracle
from lambent_amd import api, scenes

pytestmark = pytest.mark.gpu


class Hip:
    """Device buffers through the HIP runtime the library itself uses."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def check(self, rc, what):
        assert rc == 0, f"{what} failed: hipError {rc}"

    def upload(self, host: np.ndarray) -> int:
        p = ctypes.c_void_p()
        self.check(self.lib.hipMalloc(ctypes.byref(p), host.nbytes), "hipMalloc")
        self.check(self.lib.hipMemcpy(p, host.ctypes.data, host.nbytes, 1), "hipMemcpy H2D")
        return p.value

    def download(self, dev: int, out: np.ndarray) -> np.ndarray:
        self.check(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self.check(self.lib.hipMemcpy(out.ctypes.data, dev, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self, dev: int) -> None:
        self.check(self.lib.hipFree(dev), "hipFree")


HIP = Hip()


def deep_cluster_scene(n=1500, ratio=1.018, soup=20000, seed=3):
    """Triangles at geometric scales ratio^-k along one direction (a BVH of
    more node levels than the LDS stack holds) plus a random soup apart from
    them."""
    k = np.arange(n, dtype=np.float64)
    x = ratio ** -k
    c = np.stack([x, 0.3 * x, -0.2 * x], axis=1)
    s = (0.05 * x)[:, None]
    tri = np.stack([c, c + s * np.array([1.0, 0.0, 0.0]), c + s * np.array([0.0, 1.0, 1.0])], axis=1)
    rng = np.random.default_rng(seed)
    centre = rng.uniform(3.0, 6.0, (soup, 1, 3))
    off = rng.normal(size=(soup, 3, 3))
    off *= (rng.lognormal(-3.0, 0.5, (soup, 3)) / np.linalg.norm(off, axis=2))[..., None]
    tri = np.concatenate([tri, centre + off]).astype(np.float32)
    v = np.zeros((len(tri) * 3, 4), np.float32)
    v[:, :3] = tri.reshape(-1, 3)
    return v, tri


def rays_toward(points, rng, count, spread):
    tgt = points[rng.integers(0, len(points), count)] + rng.normal(scale=spread, size=(count, 3)) * 1e-3
    org = rng.uniform(-3, 3, (count, 3))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((count, 8), np.float32)
    r[:, :3] = org
    r[:, 3] = 1e-4
    r[:, 4:7] = d
    r[:, 7] = 1000.0
    return r


def run(scene, rays, allocate):
    n = len(rays)
    dr = HIP.upload(np.ascontiguousarray(rays, np.float32))
    dh = HIP.upload(np.full((n, 4), 0xAAAAAAAA, np.uint32))
    try:
        scene.intersect_device(dr, n, dh, None, allocate=allocate)
        api.sync()
        return HIP.download(dh, np.empty((n, 4), np.uint32))
    finally:
        HIP.free(dr)
        HIP.free(dh)


def check_pair(scene, rays):
    a = run(scene, rays, True)
    b = run(scene, rays, False)
    np.testing.assert_array_equal(b, a)
    return a


def test_noalloc_matches_allocating_query_on_a_soup():
    spec = scenes.procedural(50_000, seed=9, width=64, height=36)
    scene = spec.build()
    rng = np.random.default_rng(1)
    big = np.zeros((600_000, 8), np.float32)
    big[:, :3] = rng.uniform(-12, 12, (len(big), 3))
    d = rng.normal(size=(len(big), 3))
    big[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    big[:, 3], big[:, 7] = 1e-4, 1000.0
    h1 = check_pair(scene, big)
    assert (h1[:, 0] != 0xFFFFFFFF).mean() > 0.2
    check_pair(scene, big[:7777])        # a second, smaller call


def test_noalloc_spills_on_a_deep_bvh_and_matches_brute_force():
    v, tri = deep_cluster_scene()
    scene = api.Scene(v, None, np.zeros(len(tri), np.uint32), api.materials_array([
        api.Material.new([0.5, 0.5], [0.5, 0.5], [0.5, 0.5], [0, 0], 0, None, None,
                         api.EmissionScale.from_emission_scale(0.0), None, api.MaterialType.Diffuse)]),
        api.TextureSet([api.solid_rgba8(200, 200, 200, 255)], [], [], None, None))
    depth = scene.bvh_stats().max_depth
    print(f"deep scene: {len(tri)} triangles, {depth} node levels")
    assert depth >= 12
    rng = np.random.default_rng(2)
    cluster = tri[:1000].mean(axis=1).astype(np.float64)
    # rays along the cluster's axis pierce its nested boxes and their
    # siblings at every level, so their stacks grow past the LDS part
    axis = np.array([1.0, 0.3, -0.2]) / np.linalg.norm([1.0, 0.3, -0.2])
    along = np.zeros((20_000, 8), np.float32)
    along[:, :3] = 2.5 * axis + rng.normal(scale=1e-3, size=(20_000, 3))
    d = -axis + rng.normal(scale=2e-3, size=(20_000, 3))
    along[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    along[:, 3], along[:, 7] = 1e-4, 1000.0
    rays = np.concatenate([rays_toward(cluster, rng, 60_000, 1e-9),
                           rays_toward(cluster[200:], rng, 20_000, 1e-7), along])
    hits = check_pair(scene, rays)
    assert (hits[:, 0] < 1000).sum() > 100          # rays end in the cluster
    sub = rays[::7]
    ref = oracle.intersect_brute(v, None, sub)
    np.testing.assert_array_equal(hits[::7, 0], ref["prim"])
    hit = ref["prim"] != oracle.MISS
    np.testing.assert_array_equal(hits[::7, 1][hit], ref["t"].view(np.uint32)[hit])
    check_pair(scene, rays[:333])        # a second, smaller call
