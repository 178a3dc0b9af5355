"""Moving instances: lmb_scene_update_instances, the counterpart of writing
new TlasInstance transforms and rebuilding the TLAS
(examples/cornell_box/src/main.rs:560-563).

After an update the scene's closest hits must equal, bit for bit, the
oracle's over the same triangles transformed ((m0 x + m1 y) + m2 z) + m3 in
float32, and those of a scene made directly with the new transforms; images
must equal that scene's on the same seeds.  The refit keeps the subtree
topology of the creation pose, so only the node boxes differ from a fresh
build, and the boxes never change a closest hit.
"""
import time

import numpy as np
import pytest

import oracle
from lambent_amd import api, scenes

pytestmark = pytest.mark.gpu

IDENT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32)


def _rot_y(deg, t, mirror=False):
    a = np.radians(deg)
    c, s = np.cos(a), np.sin(a)
    m = np.array([[c, 0, s, t[0]], [0, 1, 0, t[1]], [-s, 0, c, t[2]]], np.float32)
    if mirror:
        m[:, 0] = -m[:, 0]
    return m


def _soup(n, seed, size=0.1):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-size, size, (n, 1, 3))
    v = (c + rng.normal(scale=size / 10, size=(n, 3, 3))).reshape(-1, 3)
    out = np.zeros((len(v), 4), np.float32)
    out[:, :3] = v
    return out


class Room:
    """The Cornell room (identity), a box, a quad and several triangle soups
    as instances whose transforms depend on a time t."""

    def __init__(self, n_soups=6, soup_tris=2000):
        self.base = scenes.cornell_box(64, 64)
        box_v = np.array([[x, y, z, 0] for z in (-0.1, 0.1) for y in (0, 0.3)
                          for x in (-0.1, 0.1)], np.float32)
        quads = [(0, 2, 6, 4), (1, 5, 7, 3), (0, 4, 5, 1), (2, 3, 7, 6), (0, 1, 3, 2),
                 (4, 6, 7, 5)]
        box_i = np.array([k for q in quads for k in (q[0], q[1], q[2], q[0], q[2], q[3])],
                         np.uint32)
        quad_v = np.array([[0.1, 0.6, 0.6, 0], [0.3, 0.6, 0.6, 0], [0.3, 0.8, 0.6, 0],
                           [0.1, 0.6, 0.6, 0], [0.3, 0.8, 0.6, 0], [0.1, 0.8, 0.6, 0]],
                          np.float32)
        room_v = self.base.vertices[self.base.indices.reshape(-1)]
        self.nroom = len(room_v) // 3
        self.nmat = len(self.base.materials)
        self.geometries = [(box_v, box_i), (quad_v, None), (room_v, None),
                           (_soup(soup_tris, 1), None)]
        self.mats = [np.arange(12, dtype=np.uint32) % self.nmat,
                     np.full(2, self.nmat - 1, np.uint32), self.base.material_idx,
                     (np.arange(soup_tris) % self.nmat).astype(np.uint32)]
        self.n_soups = n_soups

    def instances(self, t, masks=None, mirror_box=False):
        inst = [(2, IDENT, self.mats[2], 0xFF),
                (0, _rot_y(15 + 40 * t, (0.3, 0.05 * t, 0.35), mirror_box), self.mats[0], 0xFF)]
        for k in range(self.n_soups):
            a = 2 * np.pi * (k / self.n_soups + 0.1 * t)
            pos = (0.5 + 0.25 * np.cos(a), 0.4 + 0.1 * np.sin(a + t), 0.5 + 0.25 * np.sin(a))
            inst.append((3, _rot_y(30 * k + 90 * t, pos), self.mats[3], 0xFF))
        inst.append((1, IDENT, self.mats[1], 0xFF))
        if masks is not None:
            inst = [(g, tf, mi, m) for (g, tf, mi, _), m in zip(inst, masks)]
        return inst

    def scene(self, instances):
        return api.Scene.from_instances(self.geometries, instances, self.base.materials,
                                        self.base.texture_set())

    def flat_vertices(self, instances, created_masks=None):
        """World-space triangles in primitive order; instances hidden now but
        seen at creation keep their primitive ids as degenerate triangles
        far away (never hit)."""
        out = []
        created = created_masks or [m for *_, m in instances]
        for (g, tf, _mi, mask), m0 in zip(instances, created):
            if m0 & 0xFF == 0:
                continue
            v, idx = self.geometries[g]
            p = v[idx] if idx is not None else v
            if mask & 0xFF == 0:
                w = np.full((len(p), 4), 1e6, np.float32)
                w[:, 3] = 0
                out.append(w)
                continue
            x, y, z = p[:, 0], p[:, 1], p[:, 2]
            w = [((tf[r, 0] * x + tf[r, 1] * y) + tf[r, 2] * z) + tf[r, 3] for r in range(3)]
            out.append(np.stack(w + [np.zeros_like(x)], 1).astype(np.float32))
        return np.concatenate(out)

    def oracle_scene(self, instances):
        """The oracle's scene of the same world-space triangles, shading from
        the object-space ones (with each triangle's mirror flag)."""
        obj, mir, mats = [], [], []
        for g, tf, mi, mask in instances:
            if mask & 0xFF == 0:
                continue
            v, idx = self.geometries[g]
            p = v[idx] if idx is not None else v
            obj.append(p[:, :3].reshape(-1, 9))
            mir.append(np.full(len(p) // 3, np.linalg.det(tf[:, :3].astype(np.float64)) < 0,
                               np.uint8))
            mats.append(mi)
        b = self.base
        spec = scenes.SceneSpec(self.flat_vertices(instances), None,
                                np.concatenate(mats).astype(np.uint32), b.materials, b.diffuse,
                                b.emission, b.attributes, b.recolour, b.background, b.proj,
                                b.view, 64, 64)
        return oracle.Scene(spec, object_space=(np.concatenate(obj), np.concatenate(mir)))


def _random_rays(n, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(0.02, 0.98, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = 1000.0
    return rays


def _same_hits(got, ref):
    assert np.array_equal(got["prim"], ref["prim"]), (
        f"{np.count_nonzero(got['prim'] != ref['prim'])} of {len(ref)} triangle ids differ")
    hit = ref["prim"] != api.MISS
    for f in ("t", "u", "v"):
        assert np.array_equal(got[f][hit].view(np.uint32), ref[f][hit].view(np.uint32)), f


def _primary(scene, cam, w=64, h=64):
    p, t = scene.primary_hits(cam, w, h)
    return p.reshape(-1), t.reshape(-1)


def test_update_matches_the_oracle_and_a_fresh_scene():
    room = Room()
    scene = room.scene(room.instances(0.0))
    inst1 = room.instances(1.0)
    scene.update_instances(inst1)
    fresh = room.scene(inst1)
    cam = room.base.camera()
    p_u, t_u = _primary(scene, cam)
    p_f, t_f = _primary(fresh, cam)
    np.testing.assert_array_equal(p_u, p_f)
    np.testing.assert_array_equal(t_u.view(np.uint32), t_f.view(np.uint32))
    rays = oracle.create_rays(cam.to_bytes(), 64, 64)
    ref = oracle.Bvh(room.flat_vertices(inst1), None).intersect(rays)
    np.testing.assert_array_equal(p_u, ref["prim"])
    hit = ref["prim"] != api.MISS
    assert hit.mean() > 0.9
    np.testing.assert_array_equal(t_u[hit].view(np.uint32), ref["t"][hit].view(np.uint32))
    # the room's triangles come first, then the box's 12, then the soups
    assert np.count_nonzero(p_u >= room.nroom + 12) > 50, "soups in view"
    rays = _random_rays(50_000, 3)
    _same_hits(scene.intersect(rays), oracle.Bvh(room.flat_vertices(inst1), None).intersect(rays))
    pipe = api.RayTracer(api.Medium).create_pipeline(
        api.PathTracerOpts(samples=2, bounces=5, guiding=api.Guiding.OFF))
    imgs = []
    for sc in (scene, fresh):
        bufs = api.DataBuffers(64, 64, 1)
        pipe.dispatch(sc, cam, bufs, 5)
        imgs.append(bufs.output(api.Output.COLOR))
    np.testing.assert_array_equal(imgs[0], imgs[1])
    st_u, st_f = scene.bvh_stats(), fresh.bvh_stats()
    assert st_u.num_triangles == st_f.num_triangles
    assert st_u.max_depth == st_f.max_depth
    print(f"after update: {st_u.num_nodes} nodes, SAH {st_u.sah_cost:.1f}; fresh build: "
          f"{st_f.num_nodes} nodes, SAH {st_f.sah_cost:.1f}")


def test_an_animation_stays_exact_every_frame():
    room = Room(n_soups=10, soup_tris=3000)
    scene = room.scene(room.instances(0.0))
    rays = _random_rays(20_000, 9)
    for f in range(1, 13):
        inst = room.instances(0.37 * f)
        scene.update_instances(inst)
        _same_hits(scene.intersect(rays), oracle.Bvh(room.flat_vertices(inst), None).intersect(rays))


def test_masks_hide_and_show_instances():
    room = Room()
    inst0 = room.instances(0.5)
    scene = room.scene(inst0)
    created = [m for *_, m in inst0]
    cam = room.base.camera()
    p0, t0 = _primary(scene, cam)
    rays = oracle.create_rays(cam.to_bytes(), 64, 64)
    # hide the box and two soups
    masks = list(created)
    for k in (1, 2, 5):
        masks[k] = 0
    hidden = room.instances(0.5, masks)
    scene.update_instances(hidden)
    p1, t1 = _primary(scene, cam)
    ref = oracle.Bvh(room.flat_vertices(hidden, created), None).intersect(rays)
    np.testing.assert_array_equal(p1, ref["prim"])
    nroom = room.nroom
    box = (p1 >= nroom) & (p1 < nroom + 12)
    assert not box.any()
    assert np.count_nonzero((p0 >= nroom) & (p0 < nroom + 12)) > 0, "the box is in view"
    # everything hidden: every ray misses
    scene.update_instances(room.instances(0.5, [0] * len(created)))
    p2, _ = _primary(scene, cam)
    assert np.all(p2 == api.MISS)
    assert scene.bvh_stats().max_depth == 1
    # shown again: the hits of the creation pose
    scene.update_instances(inst0)
    p3, t3 = _primary(scene, cam)
    np.testing.assert_array_equal(p3, p0)
    np.testing.assert_array_equal(t3.view(np.uint32), t0.view(np.uint32))


def test_mirroring_update_flips_object_space_facing():
    room = Room(n_soups=2)
    scene = room.scene(room.instances(0.2))
    mirrored = room.instances(0.2, mirror_box=True)
    scene.update_instances(mirrored)
    fresh = room.scene(mirrored)
    cam = room.base.camera()
    out = []
    for sc in (scene, fresh):
        pipe = api.RayTracer(api.debug.FrontFace).create_pipeline(api.PathTracerOpts())
        bufs = api.DataBuffers(64, 64, 1)
        pipe.dispatch(sc, cam, bufs, 0)
        out.append(bufs.output(api.Output.COLOR))
    np.testing.assert_array_equal(out[0], out[1])
    ref = room.oracle_scene(mirrored).trace_debug(cam.to_bytes(), 64, 64, 0)
    np.testing.assert_array_equal(out[0], ref)
    # and back: equal to a scene made unmirrored
    plain = room.instances(0.2)
    scene.update_instances(plain)
    ref = room.scene(plain)
    imgs = []
    for sc in (scene, ref):
        pipe = api.RayTracer(api.debug.FrontFace).create_pipeline(api.PathTracerOpts())
        bufs = api.DataBuffers(64, 64, 1)
        pipe.dispatch(sc, cam, bufs, 0)
        imgs.append(bufs.output(api.Output.COLOR))
    np.testing.assert_array_equal(imgs[0], imgs[1])
    assert not np.array_equal(out[0], imgs[0])


def test_invalid_updates_leave_the_scene_unchanged():
    room = Room(n_soups=2)
    inst = room.instances(0.3)
    inst[3] = (inst[3][0], inst[3][1], inst[3][2], 0)   # a soup hidden at creation
    scene = room.scene(inst)
    cam = room.base.camera()
    p0, t0 = _primary(scene, cam)
    shown = room.instances(0.9)
    moved = list(shown)
    moved[3] = (moved[3][0], moved[3][1], moved[3][2], 0)
    bad = [
        moved[:-1],                                               # count
        [(3 if k == 1 else g, tf, mi, m) for k, (g, tf, mi, m) in enumerate(moved)],  # geometry
        [(g, np.full((3, 4), np.nan, np.float32) if k == 1 else tf, mi, m)
         for k, (g, tf, mi, m) in enumerate(moved)],              # NaN transform
        [(g, tf * np.float32(1e35) if k == 1 else tf, mi, m)
         for k, (g, tf, mi, m) in enumerate(moved)],              # beyond 1e30
        shown,                                                    # shows the hidden soup
    ]
    for b in bad:
        with pytest.raises(api.LambentError):
            scene.update_instances(b)
        p, t = _primary(scene, cam)
        np.testing.assert_array_equal(p, p0)
        np.testing.assert_array_equal(t.view(np.uint32), t0.view(np.uint32))
    flat = room.base.build()
    with pytest.raises(api.LambentError):
        flat.update_instances(moved)
    scene.update_instances(moved)
    p, _ = _primary(scene, cam)
    assert not np.array_equal(p, p0)


def test_many_instances_build_a_deep_top_tree():
    rng = np.random.default_rng(21)
    soup = _soup(40, 2, size=0.05)
    n = 400
    base = scenes.cornell_box(64, 64)
    mats = np.zeros(40, np.uint32)

    def pose(seed):
        r = np.random.default_rng(seed)
        return [(0, _rot_y(r.uniform(0, 360), r.uniform(0.05, 0.95, 3)), mats, 0xFF)
                for _ in range(n)]

    geos = [(soup, None)]
    inst = pose(1)
    scene = api.Scene.from_instances(geos, inst, base.materials, base.texture_set())
    assert scene.bvh_stats().max_depth >= 4
    rays = _random_rays(30_000, int(rng.integers(1 << 30)))

    def flat(instances):
        x, y, z = soup[:, 0], soup[:, 1], soup[:, 2]
        out = []
        for _, tf, _, _ in instances:
            w = [((tf[r, 0] * x + tf[r, 1] * y) + tf[r, 2] * z) + tf[r, 3] for r in range(3)]
            out.append(np.stack(w + [np.zeros_like(x)], 1).astype(np.float32))
        return np.concatenate(out)

    for seed in (2, 3):
        inst = pose(seed)
        scene.update_instances(inst)
        ref = oracle.Bvh(flat(inst), None).intersect(rays)
        assert (ref["prim"] != api.MISS).mean() > 0.03
        _same_hits(scene.intersect(rays), ref)


def test_moving_a_million_triangles_costs_milliseconds():
    """16 instances of a 62,500-triangle soup: an update (top tree on the
    host, placement and refit on the device) against the creation, which
    builds every subtree on the host."""
    soup = _soup(62_500, 4, size=0.12)
    base = scenes.cornell_box(64, 64)
    mats = (np.arange(62_500) % len(base.materials)).astype(np.uint32)

    def pose(t):
        return [(0, _rot_y(20 * k + 45 * t, (0.15 + 0.23 * (k % 4), 0.3 + 0.1 * t,
                                              0.15 + 0.23 * (k // 4))), mats, 0xFF)
                for k in range(16)]

    t0 = time.perf_counter()
    scene = api.Scene.from_instances([(soup, None)], pose(0.0), base.materials,
                                     base.texture_set())
    create_s = time.perf_counter() - t0
    times = []
    for f in range(1, 6):
        t0 = time.perf_counter()
        scene.update_instances(pose(0.1 * f))
        times.append(time.perf_counter() - t0)
    update_s = min(times)
    print(f"1M triangles in 16 instances: created in {create_s * 1e3:.0f} ms, "
          f"updated in {update_s * 1e3:.2f} ms (min of 5)")
    assert update_s * 10 < create_s
    rays = _random_rays(20_000, 5)
    flat = []
    x, y, z = soup[:, 0], soup[:, 1], soup[:, 2]
    for _, tf, _, _ in pose(0.5):
        w = [((tf[r, 0] * x + tf[r, 1] * y) + tf[r, 2] * z) + tf[r, 3] for r in range(3)]
        flat.append(np.stack(w + [np.zeros_like(x)], 1).astype(np.float32))
    _same_hits(scene.intersect(rays), oracle.Bvh(np.concatenate(flat), None).intersect(rays))


def test_updates_reach_batches_and_the_other_tracers():
    """A batch made before an update renders the moved instances (its scene
    table holds the same buffers), equal to per-image traces of the updated
    scene; guided and cached traces of the updated scene stay finite."""
    room = Room(n_soups=3)
    scene = room.scene(room.instances(0.0))
    cam = room.base.camera()
    opts = api.PathTracerOpts(samples=2, bounces=4, guiding=api.Guiding.OFF,
                              restir_temporal=api.RestirTemporalMode.Disabled)
    batch = api.Batch([(scene, cam, 64, 64)], opts, api.Medium)
    batch.trace(3)
    before = batch.output(0)
    scene.update_instances(room.instances(1.3))
    batch.trace(3)
    after = batch.output(0)
    assert not np.array_equal(before, after)
    pipe = api.RayTracer(api.Medium).create_pipeline(opts)
    bufs = api.DataBuffers(64, 64, 1, api.BufferType(0))
    pipe.dispatch(scene, cam, bufs, 3)
    np.testing.assert_array_equal(after, bufs.output(api.Output.COLOR))
    guided = api.RayTracer(api.Medium).create_pipeline(
        api.PathTracerOpts(samples=1, bounces=5, guiding=api.Guiding.MARKOV_CHAIN))
    gb = api.DataBuffers(64, 64, 10_000, api.BufferType.all())
    for f in range(3):
        guided.dispatch(scene, cam, gb, f)
        gb.advance_frame(api.BufferType.all(), api.AdvanceOptions(20))
    assert np.isfinite(gb.output(api.Output.COLOR)).all()
    cache = api.NeuralCache(api.NeuralCacheOpts(train_paths=4096, train_steps=4))
    cached = api.RayTracer(api.Medium).create_pipeline(
        api.PathTracerOpts(samples=2, bounces=5, guiding=api.Guiding.OFF))
    cb = api.DataBuffers(64, 64, 1, api.BufferType(0))
    for f in range(3):
        scene.update_instances(room.instances(1.3 + 0.2 * f))
        cached.dispatch_cached(scene, cam, cb, cache, f)
    img = cb.output(api.Output.COLOR)
    assert np.isfinite(img).all() and img[..., :3].mean() > 0
