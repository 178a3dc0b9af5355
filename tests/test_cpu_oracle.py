"""Pins the CPU oracle itself (independent checks, no GPU).

* rand_u32 / rand_f32 (shared.wesl:266-284): numpy restatement on 2^20
  inputs, injectivity over 2^24 consecutive seeds, chi-square uniformity,
  the fract() edge case;
* create_ray (general.wesl:1158-1173): float64 closed form, the 4-D
  normalize length quirk, the pixel-corner convention (no +0.5);
* brute-force watertight closest hit: float64 Moller-Trumbore on random
  soups, edge watertightness on a triangle fan, t-range semantics;
* oracle BVH: identical to the brute force;
* path tracer: closed furnace identities, background-only image,
  Oren-Nayar at roughness 0, Schlick, GGX VNDF at roughness 0, cosine-lobe
  distribution, the High-quality CIE colour weights against float64.
"""
import numpy as np
import pytest
from scipy import stats

import oracle
from lambent_amd import api, scenes


def pcg_numpy(s):
    s = np.asarray(s, np.uint64)
    m = np.uint64(0xFFFFFFFF)
    state = (s * np.uint64(747796405) + np.uint64(2891336453)) & m
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & m
    return ((word >> np.uint64(22)) ^ word).astype(np.uint32)


def test_rand_u32_matches_numpy():
    rng = np.random.default_rng(0)
    seeds = np.concatenate([np.arange(1 << 16, dtype=np.uint32),
                            rng.integers(0, 2 ** 32, size=1 << 20, dtype=np.uint64).astype(np.uint32),
                            np.array([0xFFFFFFFF, 0x80000000], np.uint32)])
    np.testing.assert_array_equal(oracle.rand_u32(seeds), pcg_numpy(seeds))


def test_rand_u32_is_injective_on_a_block():
    seeds = np.arange(1 << 24, dtype=np.uint32) + np.uint32(123456789)
    out = oracle.rand_u32(seeds)
    assert len(np.unique(out)) == len(out)


def test_rand_f32_uniform_and_fract_edge():
    seeds = np.arange(1 << 20, dtype=np.uint32)
    f = oracle.rand_f32(seeds)
    assert f.min() >= 0.0 and f.max() < 1.0
    counts = np.histogram(f, bins=64, range=(0, 1))[0]
    assert stats.chisquare(counts).pvalue > 1e-4
    # a hash that rounds to 2^32 as f32 gives exactly 0 (fract)
    # (the hash is a bijection, so exactly 128 seeds map into the top range)
    idx = np.zeros(0, np.int64)
    for chunk in range(64):
        base = chunk << 24
        h = oracle.rand_u32(np.arange(base, base + (1 << 24), dtype=np.uint64).astype(np.uint32))
        idx = np.nonzero(h >= np.uint32(0xFFFFFF80))[0] + base
        if len(idx):
            break
    assert len(idx) > 0
    assert (oracle.rand_f32(idx.astype(np.uint32)) == 0.0).all()
    sel = np.arange(1000, dtype=np.uint32)
    np.testing.assert_array_equal(oracle.rand_f32(sel),
                                  (pcg_numpy(sel).astype(np.float32) * np.float32(2.0 ** -32)))


def create_ray_f64(pinv, vinv, x, y, w, h):
    nx = x / (w / 2.0) - 1.0
    ny = y / (h / 2.0) - 1.0
    P = pinv @ np.array([nx, -ny, 1.0, 1.0])
    q = P / np.linalg.norm(P)
    d = (vinv @ np.array([q[0], q[1], q[2], 0.0]))[:3]
    o = (vinv @ np.array([0.0, 0.0, 0.0, 1.0]))[:3]
    return o, d


def test_create_ray_float64_closed_form():
    spec = scenes.procedural(10, width=320, height=180)
    cam = spec.camera()
    pinv = cam.projection_inverse.T.astype(np.float64)
    vinv = cam.view_inverse.T.astype(np.float64)
    rays = oracle.create_rays(cam.to_bytes(), 320, 180)
    rng = np.random.default_rng(3)
    for k in rng.integers(0, 320 * 180, size=400):
        x, y = k % 320, k // 320
        o, d = create_ray_f64(pinv, vinv, x, y, 320, 180)
        np.testing.assert_allclose(rays[k, :3], o, rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(rays[k, 4:7], d, rtol=2e-6, atol=2e-7)
    # |dir| = 1 / sqrt(1 + (w/|xyz|)^2) < 1: the vec4 normalize quirk
    lens = np.linalg.norm(rays[:, 4:7].astype(np.float64), axis=1)
    assert (lens < 1.0).all() and (lens > 0.9999).all()
    for k in rng.integers(0, 320 * 180, size=50):
        x, y = k % 320, k // 320
        P = pinv @ np.array([x / 160.0 - 1.0, -(y / 90.0 - 1.0), 1.0, 1.0])
        expect = 1.0 / np.sqrt(1.0 + (P[3] / np.linalg.norm(P[:3])) ** 2)
        assert lens[k] == pytest.approx(expect, abs=3e-7)
    # pixel (w/2, h/2) is the optical axis: direction +z
    c = (90 * 320) + 160
    np.testing.assert_allclose(rays[c, 4:7] / np.linalg.norm(rays[c, 4:7]), [0, 0, 1], atol=1e-6)


def moller_trumbore_f64(verts, idx, rays):
    v = verts[:, :3].astype(np.float64)
    tri = v[idx.reshape(-1, 3)]
    out_prim = np.full(len(rays), -1)
    out_t = np.full(len(rays), np.inf)
    for i, r in enumerate(rays.astype(np.float64)):
        o, d = r[:3], r[4:7]
        e1 = tri[:, 1] - tri[:, 0]
        e2 = tri[:, 2] - tri[:, 0]
        p = np.cross(d, e2)
        det = np.einsum("ij,ij->i", e1, p)
        ok = np.abs(det) > 1e-30
        inv = np.where(ok, 1.0 / np.where(ok, det, 1), 0)
        s = o - tri[:, 0]
        u = np.einsum("ij,ij->i", s, p) * inv
        q = np.cross(s, e1)
        vv = (q @ d) * inv
        t = np.einsum("ij,ij->i", e2, q) * inv
        good = ok & (u >= 0) & (vv >= 0) & (u + vv <= 1) & (t >= r[3]) & (t <= r[7])
        if good.any():
            j = np.argmin(np.where(good, t, np.inf))
            out_prim[i], out_t[i] = j, t[j]
    return out_prim, out_t


def random_soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, size=(n, 1, 3))
    v = (c + 0.3 * rng.normal(size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    verts = np.concatenate([v, np.zeros((len(v), 1), np.float32)], 1)
    return verts, np.arange(3 * n, dtype=np.uint32)


def random_rays(n, seed, spread=2.0):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = rng.uniform(-spread, spread, size=(n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 3] = 1e-4
    r[:, 7] = 100.0
    return r


def test_bruteforce_matches_float64_moller_trumbore():
    verts, idx = random_soup(300, 1)
    rays = random_rays(3000, 2)
    got = oracle.intersect_brute(verts, idx, rays)
    ref_prim, ref_t = moller_trumbore_f64(verts, idx, rays)
    gp = got["prim"].astype(np.int64)
    gp[got["prim"] == oracle.MISS] = -1
    agree = gp == ref_prim
    # disagreements only where two candidates are within rounding of a tie
    # or the hit grazes an edge; they must be rare
    assert agree.mean() > 0.995
    both = agree & (ref_prim >= 0)
    np.testing.assert_allclose(got["t"][both], ref_t[both], rtol=1e-5, atol=1e-6)


def test_t_range_is_closed_and_parametric():
    verts = np.array([[0, 0, 1, 0], [1, 0, 1, 0], [0, 1, 1, 0]], np.float32)
    idx = np.arange(3, dtype=np.uint32)
    base = np.array([0.2, 0.2, 0, 0, 0, 0, 2.0, 0], np.float32)  # |d| = 2 -> t = 0.5
    r = np.tile(base, (3, 1))
    r[0, 3], r[0, 7] = 0.0, 0.5     # tmax == t: hit
    r[1, 3], r[1, 7] = 0.5, 10.0    # tmin == t: hit
    r[2, 3], r[2, 7] = 0.0, 0.4999  # t > tmax: miss
    h = oracle.intersect_brute(verts, idx, r)
    assert list(h["prim"][:2]) == [0, 0] and h["prim"][2] == oracle.MISS
    assert h["t"][0] == 0.5
    np.testing.assert_allclose([h["u"][0], h["v"][0]], [0.2, 0.2], rtol=1e-6)


def test_fan_edges_are_watertight():
    ring = 12
    v = [(0.0, 0.0, 0.0, 0.0)] + [(np.cos(2 * np.pi * k / ring), np.sin(2 * np.pi * k / ring), 0, 0)
                                  for k in range(ring)]
    idx = []
    for k in range(ring):
        idx += [0, 1 + k, 1 + (k + 1) % ring]
    verts = np.asarray(v, np.float32)
    idx = np.asarray(idx, np.uint32)
    n = 20000
    rng = np.random.default_rng(4)
    k = rng.integers(0, ring, size=n)
    s = rng.uniform(0, 0.9, size=n)
    r = np.zeros((n, 8), np.float32)
    r[:, 0] = s * np.cos(2 * np.pi * k / ring)
    r[:, 1] = s * np.sin(2 * np.pi * k / ring)
    r[:, 2] = -1
    r[:, 4:7] = rng.normal(size=(n, 3)) * 0.05 + [0, 0, 1]
    r[:, 7] = 10
    h = oracle.intersect_brute(verts, idx, r)
    cross = r[:, :2] + (r[:, 4:6] / r[:, 6:7]) * 1.0  # z = -1 -> 0
    inside = np.hypot(cross[:, 0], cross[:, 1]) < 0.95  # inradius cos(pi/12) = 0.966
    assert inside.sum() > n * 0.8
    assert (h["prim"][inside] != oracle.MISS).all()


def test_oracle_bvh_equals_bruteforce():
    verts, idx = random_soup(3000, 5)
    rays = random_rays(20000, 6)
    a = oracle.intersect_brute(verts, idx, rays)
    b = oracle.Bvh(verts, idx).intersect(rays)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


# ------------------------------------------------------------ path tracer
def check_furnace(color, expect):
    """Every path inside a closed box carries exactly E * sum_k rho^k, except
    the rare path whose vertex lands within T_MIN of a box edge and whose next
    ray skips the adjacent wall (t < T_MIN) and escapes: it can only lose
    energy."""
    c = color[..., :3]
    exact = np.all(np.abs(c - expect) <= 2e-6 * expect, axis=-1)
    assert exact.mean() > 0.99, exact.mean()
    assert (c[~exact] < expect).all()


@pytest.mark.parametrize("quality", [0, 1])
@pytest.mark.parametrize("ty", [api.MaterialType.Diffuse, api.MaterialType.Metallic])
def test_oracle_furnace(quality, ty):
    spec = scenes.furnace_box(0.7, 3.0, ty=ty, width=24, height=20)
    c, n, a, v = oracle.Scene(spec).trace_plain(spec.camera().to_bytes(), 24, 20, quality=quality,
                                                samples=2, bounces=7)
    r = float(np.float32(round(0.7 * 255)) / np.float32(255))
    check_furnace(c, 3.0 * sum(r ** k for k in range(7)))
    assert v >= 0.99 * 24 * 20 * 2 * 7
    np.testing.assert_allclose(a[..., :3], r)


def test_oracle_background_only():
    spec = scenes.procedural(1, width=16, height=9)
    spec.vertices[:] = 1e6  # move the one triangle out of view
    c, _, a, v = oracle.Scene(spec).trace_plain(spec.camera().to_bytes(), 16, 9, samples=1)
    np.testing.assert_array_equal(c[..., :3], 1.0)  # white sky
    np.testing.assert_array_equal(a[..., :3], 0.0)
    assert v == 16 * 9


def test_oracle_cube_faces():
    spec = scenes.procedural(1, width=4, height=4)
    bg = np.zeros((6, 2, 2, 4), np.uint8)
    for f in range(6):
        bg[f] = (10 * (f + 1), 0, 0, 255)
    spec.background = bg
    o = oracle.Scene(spec)
    dirs = [(1, 0.1, 0.2), (-1, 0.1, 0.2), (0.1, 1, 0.2), (0.1, -1, 0.2), (0.1, 0.2, 1),
            (0.1, 0.2, -1)]
    for f, d in enumerate(dirs):
        assert o.background(d)[0] == pytest.approx(10 * (f + 1) / 255)


def test_oren_nayar_roughness_zero_is_one():
    rng = np.random.default_rng(7)
    for _ in range(100):
        a, b = rng.normal(size=3), rng.normal(size=3)
        np.testing.assert_array_equal(oracle.oren_nayar(a, b, 0.0, rng.uniform(size=3)), 1.0)


def test_oren_nayar_closed_form():
    inc, ref, rough, col = np.array([0.3, 0.5, -0.2]), np.array([0.1, 0.9, 0.4]), 0.6, np.array([0.2, 0.5, 0.8])
    s2 = (rough * np.pi / 2) ** 2
    a = 1 - 0.5 * s2 / (s2 + 0.33) + 0.17 * col * s2 / (s2 + 0.13)
    b = 0.45 * s2 / (s2 + 0.09)
    s = np.dot(ref, inc) - ref[1] * inc[1]
    t = max(ref[1], inc[1]) if s > 0 else 1.0
    np.testing.assert_allclose(oracle.oren_nayar(inc, ref, rough, col), a + b * s / t, rtol=1e-5)


def test_schlick_closed_form():
    for c in np.linspace(0, 1, 11):
        for n in (1 / 1.5, 1.5, 2.4):
            r0 = ((1 - n) / (1 + n)) ** 2
            assert oracle.reflectance(c, n) == pytest.approx(r0 + (1 - r0) * (1 - c) ** 5, rel=1e-5)


def test_vndf_zero_roughness_returns_normal():
    rng = np.random.default_rng(8)
    for i in range(200):
        view = rng.normal(size=3)
        view[1] = -abs(view[1]) - 0.01
        np.testing.assert_allclose(oracle.sample_vndf_ggx(i, view, 0.0, 0.0), [0, 1, 0])


def test_vndf_samples_upper_hemisphere_and_roughness_trend():
    rng = np.random.default_rng(9)
    ys = {}
    for rough in (0.1, 0.8):
        vals = []
        for i in range(3000):
            view = np.array([0.3, -0.9, 0.2])
            vals.append(oracle.sample_vndf_ggx(i * 7919, view, rough, rough))
        vals = np.array(vals)
        assert (vals[:, 1] > 0).all()
        np.testing.assert_allclose(np.linalg.norm(vals, axis=1), 1, rtol=1e-5)
        ys[rough] = vals[:, 1].mean()
    assert ys[0.1] > ys[0.8]  # rougher = wider lobe


def test_cosine_hemisphere_distribution():
    n, t = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])
    cos = np.array([oracle.cosine_hemisphere(i * 2654435761 % 2 ** 32, n, t) @ n
                    for i in range(20000)])
    # cosine-weighted: cos^2(theta) is uniform on [0, 1]
    assert stats.kstest(cos ** 2, "uniform").pvalue > 1e-3


def test_high_quality_colour_matches_float64_cie_fit():
    def g(x, u, t1, t2):
        tau = t1 if x < u else t2
        return np.exp(-((x - u) ** 2) * tau * tau / 2)

    M = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.204, 1.057]])
    for seed in range(300):
        rgb, u = oracle.high_colour(seed)
        lam = 760 * (1 - u) + 340 * u
        xyz = np.array([1.056 * g(lam, 599.8, 0.0264, 0.0323) + 0.362 * g(lam, 442.0, 0.0624, 0.0374)
                        - 0.065 * g(lam, 501.1, 0.049, 0.0382),
                        0.821 * g(lam, 568.8, 0.0213, 0.0248) + 0.286 * g(lam, 530.9, 0.0613, 0.0322),
                        1.217 * g(lam, 437.0, 0.0845, 0.0278) + 0.681 * g(lam, 459.0, 0.0385, 0.0725)])
        ref = np.clip(M @ xyz, 0, 1) * [3.910, 4.411, 5.511]
        np.testing.assert_allclose(rgb, ref, rtol=2e-4, atol=2e-5)
    # the wavelength weights average to ~1 per channel (the reference's
    # normalisation constants are the reciprocals of the integrals)
    cols = np.array([oracle.high_colour(s)[0] for s in range(20000)])
    np.testing.assert_allclose(cols.mean(axis=0), 1.0, atol=0.03)


def test_reference_test_scene_fixtures_decode():
    """The copied PNG fixtures of the reference's renderer test decode to
    64 x 64 RGBA8 through the TextureLoader path, and the oracle renders the
    scene: every primary ray hits one of its four cube faces (the camera is
    inside the cube), and the albedo output holds only texel values."""
    import ref_scene
    from lambent_amd import api
    loader = api.TextureLoader()
    for name in ("diffuse.png", "pure_white.png", "emission_partial.png", "emission.png",
                 "attributes.png", "plus_x.png", "minus_z.png"):
        loader.load(str(ref_scene.TEX / name))
        assert loader.images[-1].shape == (64, 64, 4)
    spec = ref_scene.reference_test_scene(64, 36)
    c, n, a, v = oracle.Scene(spec).trace_plain(spec.camera().to_bytes(), 64, 36, quality=1,
                                                samples=2, bounces=5)
    assert (a[..., 3] > 0).all()
    texels = {tuple(t) for im in spec.diffuse for t in im.reshape(-1, 4)[:, :3]}
    for px in np.unique(a[..., :3].reshape(-1, 3), axis=0):
        assert tuple(np.round(px * 255).astype(np.uint8)) in texels
    assert np.isfinite(c).all() and v >= 64 * 36 * 2


def _debug_scene(v, idx, ior=None):
    """Triangles in front of a camera at the origin looking down +z (60 deg,
    24 x 24 px), one material."""
    v4 = np.zeros((len(v), 4), np.float32)
    v4[:, :3] = v
    m = api.Material.new([0, 0], [1, 0], [1, 1], [0, 0], 0, None, None,
                         api.EmissionScale.from_emission_scale(0.0), ior,
                         api.MaterialType.Diffuse)
    proj = api.perspective(60.0, 1.0, 0.1, 100.0)
    view = api.look_to_rh((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0))
    return scenes.SceneSpec(v4, np.array(idx, np.uint32), np.zeros(len(idx) // 3, np.uint32),
                            api.materials_array([m]), [scenes._rgba(1, 1, 1)], [], [], None,
                            api.uniform_background(0.0), proj, view, 24, 24)


def _debug_quad(front: bool, ior=None):
    """A 3 x 3 quad in the plane z = 3; `front` picks the winding
    (counter-clockwise seen from the camera = front face)."""
    v = [[-1.5, -1.5, 3], [1.5, -1.5, 3], [1.5, 1.5, 3], [-1.5, 1.5, 3]]
    return _debug_scene(v, [0, 2, 1, 0, 3, 2] if front else [0, 1, 2, 0, 2, 3], ior)


@pytest.mark.parametrize("front", [True, False])
def test_debug_front_face_closed_form(front):
    """front_face.wesl:52-63: colour 1 / (t + 1) in green on a front face, in
    red on a back face, t the parametric distance (z = 3 over the direction's
    z), 0 on a miss."""
    spec = _debug_quad(front)
    cam = spec.camera()
    c = oracle.Scene(spec).trace_debug(cam.to_bytes(), 24, 24, 0)
    rays = oracle.create_rays(cam.to_bytes(), 24, 24).reshape(24, 24, 8).astype(np.float64)
    t = (3.0 - rays[..., 2]) / rays[..., 6]
    inside = (np.abs(rays[..., 0] + t * rays[..., 4]) < 1.5) & \
             (np.abs(rays[..., 1] + t * rays[..., 5]) < 1.5)
    lit = c[..., 1] if front else c[..., 0]
    dark = c[..., 0] if front else c[..., 1]
    assert inside.mean() > 0.5 and (~inside).any()
    np.testing.assert_allclose(lit[inside], 1.0 / (t[inside] + 1.0), rtol=2e-6)
    assert (dark == 0).all() and (c[..., 2] == 0).all() and (c[..., 3] == 1).all()
    assert (lit[~inside] == 0).all()


def test_debug_reflectance_closed_form():
    """reflectance.wesl:49-57: the material's refractive-index pair unpacked
    from f16, over 5, in red and green."""
    spec = _debug_quad(True, ior=api.refractive_indices.DIAMOND)
    c = oracle.Scene(spec).trace_debug(spec.camera().to_bytes(), 24, 24, 1)
    lo, hi = (np.float32(np.float16(x)) for x in api.refractive_indices.DIAMOND)
    hit = c[..., 0] > 0
    assert 0.5 < hit.mean() < 1.0
    assert (c[hit][:, 0] == lo / np.float32(5)).all() and (c[hit][:, 1] == hi / np.float32(5)).all()
    assert (c[~hit][:, :3] == 0).all()


@pytest.mark.parametrize("d,expect", [
    ((-3.0, 1.0, 0.5), (3.0, -1.0, -0.5)),   # x largest and negative: reversed
    ((0.5, -4.0, 1.0), (-0.5, 4.0, -1.0)),   # y largest and negative: reversed
    ((1.0, -1.0, 0.0), (1.0, -1.0, 0.0)),    # |x| = |y|: neither is largest, z >= 0 decides
    ((1.0, 1.0, -1.0), (-1.0, -1.0, 1.0)),   # all equal: z < 0 reverses
])
def test_debug_tangent_closed_form(d, expect):
    """tangent.wesl:52-72: correct_tangent(normalize(v0 - v1)) as
    (d + 1) / 2 on a triangle whose v0 - v1 is d (the largest axis made
    positive; ties leave the decision to z); a miss shows (0.5, 0.5, 0.5)."""
    d = np.asarray(d, np.float64)
    v0 = np.array([0.0, 0.0, 3.0])
    side = np.cross(d, [0.0, 0.0, 1.0])
    v = [v0, v0 - 0.5 * d / np.linalg.norm(d[:2]), v0 + side / np.linalg.norm(side)]
    spec = _debug_scene(v, [0, 1, 2])
    c = oracle.Scene(spec).trace_debug(spec.camera().to_bytes(), 24, 24, 2)
    e = np.asarray(expect) / np.linalg.norm(expect)
    want = (e + 1.0) / 2.0
    hit = np.all(np.abs(c[..., :3] - want) < 2e-7, axis=-1)
    miss = np.all(c[..., :3] == 0.5, axis=-1)
    assert hit.sum() >= 4 and miss.any()
    assert (hit | miss).all()


@pytest.mark.parametrize("front_type", [api.MaterialType.Diffuse, api.MaterialType.Transparent])
@pytest.mark.parametrize("opaque", [False, True])
def test_alpha_tested_any_hit_closed_form(front_type, opaque):
    """general.wesl:1057-1066 on non-opaque geometry: a camera ray passes the
    cut-out quad where its texel's alpha is zero and hits the quad behind
    (t = 4 / d_z instead of 2 / d_z); a TRANSPARENT material always counts
    for the path tracer but not for the debugging shaders, which test the
    alpha alone; opaque geometry ignores the alpha."""
    import cutout
    spec = cutout.cutout_scene(front_type, opaque)
    cam = spec.camera().to_bytes()
    sc = oracle.Scene(spec)
    rays = oracle.create_rays(cam, 48, 48)
    inside, opaque_texel, margin = cutout.expected_front(rays)
    clear = margin > 1e-3
    dz = rays[:, 6].astype(np.float64)
    ff = sc.trace_debug(cam, 48, 48, 0).reshape(-1, 4)
    shade = np.maximum(ff[:, 0], ff[:, 1]).astype(np.float64)
    front_debug = inside & (opaque | opaque_texel)
    t_expect = np.where(front_debug, 2.0 / dz, 4.0 / dz)
    np.testing.assert_allclose(shade[clear], 1.0 / (t_expect[clear] + 1.0), rtol=1e-5)
    assert (inside & ~opaque_texel & clear).sum() > 50 and (inside & opaque_texel & clear).sum() > 50
    # the path tracer's rule, seen in the albedo output: front texel colour or white
    _, _, alb, _ = sc.trace_plain(cam, 48, 48, quality=1, samples=1, bounces=1)
    alb = alb.reshape(-1, 4)
    front_pt = inside & (opaque | opaque_texel | (front_type == api.MaterialType.Transparent))
    white = np.all(alb[:, :3] == 1.0, axis=1)
    assert (white[clear & ~front_pt]).all()
    assert (~white[clear & front_pt]).all()
