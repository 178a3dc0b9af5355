"""GPU parity of the ray-query kernels against the CPU oracle.

Primary-hit triangle ids must be bit-exact and hit distances identical
(required: within 4 ulp); arbitrary rays must return the brute-force
closest hit.  Ties in t resolve to the lower triangle index on both sides.
"""
import numpy as np
import pytest

import oracle
from lambent_amd import api, scenes

pytestmark = pytest.mark.gpu


def _ulp_diff(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    return np.abs(ai - bi)


def _check_hits(got, ref):
    assert np.array_equal(got["prim"], ref["prim"]), (
        f"{np.count_nonzero(got['prim'] != ref['prim'])} of {len(ref)} triangle ids differ")
    hit = ref["prim"] != api.MISS
    assert _ulp_diff(got["t"][hit], ref["t"][hit]).max(initial=0) <= 4
    assert np.array_equal(got["t"][hit].view(np.uint32), ref["t"][hit].view(np.uint32))
    assert np.array_equal(got["u"][hit].view(np.uint32), ref["u"][hit].view(np.uint32))
    assert np.array_equal(got["v"][hit].view(np.uint32), ref["v"][hit].view(np.uint32))


def test_primary_hits_cornell_256():
    spec = scenes.cornell_box(256, 256)
    scene = spec.build()
    cam = spec.camera()
    prim, t = scene.primary_hits(cam, 256, 256)
    ref = oracle.intersect_brute(spec.vertices, spec.indices,
                                 oracle.create_rays(cam.to_bytes(), 256, 256))
    assert np.array_equal(prim.reshape(-1), ref["prim"])
    hit = ref["prim"] != api.MISS
    assert hit.mean() > 0.5
    assert _ulp_diff(t.reshape(-1)[hit], ref["t"][hit]).max() <= 4


def test_primary_hits_procedural_200k():
    spec = scenes.procedural(200_000, seed=7, width=480, height=270)
    scene = spec.build()
    cam = spec.camera()
    prim, t = scene.primary_hits(cam, 480, 270)
    rays = oracle.create_rays(cam.to_bytes(), 480, 270)
    ref = oracle.Bvh(spec.vertices, spec.indices).intersect(rays)
    assert np.array_equal(prim.reshape(-1), ref["prim"])
    hit = ref["prim"] != api.MISS
    assert hit.mean() > 0.2
    assert _ulp_diff(t.reshape(-1)[hit], ref["t"][hit]).max() <= 4


def test_random_rays_inside_procedural():
    spec = scenes.procedural(100_000, seed=3, width=16, height=16)
    scene = spec.build()
    rng = np.random.default_rng(11)
    n = 200_000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(-12, 12, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[::97, rng.integers(0, 3)] = 0.0  # axis-parallel components
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = 1000.0
    got = scene.intersect(rays)
    ref = oracle.Bvh(spec.vertices, spec.indices).intersect(rays)
    _check_hits(got, ref)


def test_shared_edges_and_ties_match_bruteforce():
    # A fan of triangles sharing edges and a duplicated triangle (equal t):
    # the closest hit is the lower index on both sides.
    rng = np.random.default_rng(5)
    ring = 16
    v = [(0.0, 0.0, 0.0, 0.0)]
    for k in range(ring):
        a = 2 * np.pi * k / ring
        v.append((np.cos(a), np.sin(a), 0.0, 0.0))
    idx = []
    for k in range(ring):
        idx += [0, 1 + k, 1 + (k + 1) % ring]
    idx += idx[:3]  # duplicate of triangle 0
    verts = np.asarray(v, np.float32)
    idx = np.asarray(idx, np.uint32)
    mats = api.materials_array([api.Material.new([0, 0], [0, 0], [0, 0], [0, 0], 0, None, None,
                                                 api.EmissionScale(), None,
                                                 api.MaterialType.Diffuse)])
    scene = api.Scene(verts, idx, np.zeros(len(idx) // 3, np.uint32), mats)
    n = 50_000
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:2] = rng.uniform(-1.2, 1.2, size=(n, 2)).astype(np.float32)
    # snap some origins exactly onto the shared edges
    k = rng.integers(0, ring, size=n // 4)
    s = rng.uniform(0, 1, size=n // 4)
    rays[: n // 4, 0] = (s * np.cos(2 * np.pi * k / ring)).astype(np.float32)
    rays[: n // 4, 1] = (s * np.sin(2 * np.pi * k / ring)).astype(np.float32)
    rays[:, 2] = -1.0
    rays[:, 3] = 0.0
    rays[:, 6] = 1.0
    rays[:, 7] = 10.0
    got = scene.intersect(rays)
    ref = oracle.intersect_brute(verts, idx, rays)
    _check_hits(got, ref)
    inside = np.hypot(rays[:, 0], rays[:, 1]) < 0.9
    assert (got["prim"][inside] != api.MISS).all(), "watertightness: a ray fell through an edge"
    assert (got["prim"] != ring).all(), "the duplicate must lose the tie to triangle 0"


def test_empty_scene_misses():
    mats = api.materials_array([api.Material.new([0, 0], [0, 0], [0, 0], [0, 0], 0, None, None,
                                                 api.EmissionScale(), None,
                                                 api.MaterialType.Diffuse)])
    scene = api.Scene(np.zeros((0, 4), np.float32), None, np.zeros(0, np.uint32), mats)
    rays = np.tile(np.array([0, 0, 0, 0, 0, 0, 1, 100], np.float32), (10, 1))
    got = scene.intersect(rays)
    assert (got["prim"] == api.MISS).all()


def test_bvh_stats_sane():
    spec = scenes.procedural(100_000, seed=1, width=16, height=16)
    scene = spec.build()
    st = scene.bvh_stats()
    assert st.num_triangles == 100_000
    assert 1 < st.max_depth < 20
    assert st.num_nodes > 100_000 / 24


@pytest.fixture(scope="module")
def soup_100k():
    spec = scenes.procedural(100_000, seed=5, width=16, height=16)
    return spec, spec.build(), oracle.Bvh(spec.vertices, spec.indices)


@pytest.mark.parametrize("n", [1, 2, 63, 65, 700, 5000])
def test_short_queues_where_idle_lanes_help(soup_100k, n):
    """A queue shorter than the grid's waves gives each wave one ray or a
    few, so nearly every lane of a wave is idle from the start: idle lanes
    traverse the deferred subtrees of the rays still running, their
    triangles go through the carried triangle rounds, and each ray's hit is
    stored once all of its helpers are done.  The hits must be the same
    closest hits as the oracle's."""
    spec, scene, bvh = soup_100k
    rng = np.random.default_rng(100 + n)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(-14, 14, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = 1000.0
    _check_hits(scene.intersect(rays), bvh.intersect(rays))


@pytest.mark.parametrize("n", [1, 33, 1000])
def test_short_queues_on_a_deep_bvh(n):
    """The same with rays along a chain of nested triangles whose BVH is
    deeper than the 8-entry LDS stack: the donated stack entries come from
    the spill area as well."""
    k = np.arange(1500, dtype=np.float64)
    x = 1.018 ** -k
    c = np.stack([x, 0.3 * x, -0.2 * x], axis=1)
    s = (0.05 * x)[:, None]
    tri = np.stack([c, c + s * np.array([1.0, 0.0, 0.0]), c + s * np.array([0.0, 1.0, 1.0])], axis=1)
    v = np.zeros((len(tri) * 3, 4), np.float32)
    v[:, :3] = tri.reshape(-1, 3).astype(np.float32)
    scene = api.Scene(v, None, np.zeros(len(tri), np.uint32), api.materials_array([
        api.Material.new([0.5, 0.5], [0.5, 0.5], [0.5, 0.5], [0, 0], 0, None, None,
                         api.EmissionScale.from_emission_scale(0.0), None, api.MaterialType.Diffuse)]),
        api.TextureSet([api.solid_rgba8(200, 200, 200, 255)], [], [], None, None))
    assert scene.bvh_stats().max_depth > 12
    rng = np.random.default_rng(7 + n)
    # from beyond the chain's large end towards the centroids of its small,
    # deeply nested triangles: every ray crosses the nested boxes level by
    # level and hits at least its target
    axis = np.array([1.0, 0.3, -0.2]) / np.linalg.norm([1.0, 0.3, -0.2])
    target = tri[rng.integers(600, 1500, n)].mean(axis=1)
    org = 2.5 * axis + rng.normal(scale=1e-3, size=(n, 3))
    d = target - org
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = org
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3], rays[:, 7] = 1e-4, 1000.0
    got = scene.intersect(rays)
    ref = oracle.intersect_brute(v, None, rays)
    assert np.array_equal(got["prim"], ref["prim"])
    hit = ref["prim"] != api.MISS
    assert hit.any()
    assert np.array_equal(got["t"][hit].view(np.uint32), ref["t"][hit].view(np.uint32))
