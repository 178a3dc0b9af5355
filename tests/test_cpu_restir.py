"""Pins the oracle's reservoir algebra, octahedral packing and end-of-frame
processing with closed forms and statistics (no GPU)."""
import numpy as np
import pytest
from scipy import stats

import oracle

R = oracle.RESERVOIR_DTYPE


def make_res(n, rng, diffuse_frac=0.8, normal_spread=0.05, sky_frac=0.05):
    r = np.zeros(n, R)
    r["visible_point"] = rng.normal(size=(n, 3))
    r["sample_point"] = rng.normal(size=(n, 3)) + [0, 3, 0]
    r["out_radiance"] = rng.uniform(0, 2, size=(n, 3))
    for k in range(n):
        nv = np.array([0, 1, 0]) + normal_spread * rng.normal(size=3)
        ns = np.array([0, 1, 0]) + normal_spread * rng.normal(size=3)
        r["visible_normal"][k] = oracle.normalised_to_u32(nv / np.linalg.norm(nv))
        r["sample_normal"][k] = oracle.normalised_to_u32(ns / np.linalg.norm(ns))
    u = rng.uniform(size=n)
    r["ty"] = np.where(u < diffuse_frac, 0, np.where(u < 1 - sky_frac, 1, 0xFF))
    r["roughness"] = rng.choice([0.0, 0.05, 0.5], size=n)
    r["w"] = rng.uniform(0, 3, size=n)
    conf = rng.integers(1, 40, size=n).astype(np.uint32)
    valid = rng.integers(0, 2, size=n).astype(np.uint32)
    r["packed_confidence_valid"] = conf | (valid << np.uint32(31))
    r["W"] = rng.uniform(0, 2, size=n)
    return r


def p_hat(r):
    return max(float(np.linalg.norm(r["out_radiance"].astype(np.float64))), 1e-5)


def test_octahedral_roundtrip_and_axes():
    rng = np.random.default_rng(0)
    for _ in range(2000):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        back = oracle.u32_to_normalised(oracle.normalised_to_u32(n))
        assert np.degrees(np.arccos(np.clip(back @ n, -1, 1))) < 0.05
    for axis in np.eye(3).tolist() + (-np.eye(3)).tolist():
        np.testing.assert_array_equal(oracle.u32_to_normalised(oracle.normalised_to_u32(axis)),
                                      np.asarray(axis, np.float32))
    # u32 0 decodes to +y
    np.testing.assert_array_equal(oracle.u32_to_normalised(0), [0, 1, 0])


def test_merge_selection_probability():
    rng = np.random.default_rng(1)
    base = make_res(2, rng, sky_frac=0.0)
    res0, other = base[0:1].copy(), base[1:2].copy()
    res0["w"] = 1.0
    res0["packed_confidence_valid"] = 5
    other["W"] = 0.5
    other["packed_confidence_valid"] = 3
    p = 0.8
    w_new = p * 0.5 * 3
    picks = 0
    trials = 20000
    for s in range(trials):
        r = res0.copy()
        picked = oracle.merge(r, other, p, s * 2654435761 % 2 ** 32)
        picks += picked
        assert r["packed_confidence_valid"][0] & 0x7FFFFFFF == 8
        assert r["w"][0] == pytest.approx(1.0 + w_new, rel=1e-6)
        if picked:
            assert r["packed_confidence_valid"][0] >> 31 == 1
            np.testing.assert_array_equal(r["sample_point"], other["sample_point"])
    expect = w_new / (1.0 + w_new)
    assert stats.binomtest(picks, trials, expect).pvalue > 1e-4


def test_update_counts_and_selects():
    rng = np.random.default_rng(2)
    res = np.zeros(1, R)
    s = make_res(1, rng, sky_frac=0.0)
    oracle.update(res, s, 0.7, 123)
    assert res["packed_confidence_valid"][0] == (1 | (1 << 31))  # first sample always taken
    assert res["w"][0] == pytest.approx(0.7)


def test_temporal_closed_form_two_samples():
    rng = np.random.default_rng(3)
    g = make_res(1, rng, sky_frac=0.0)
    old = g.copy()
    old["out_radiance"] = [[1.0, 2.0, 2.0]]  # p_hat 3
    old["packed_confidence_valid"] = 3
    old["W"] = 0.25
    g["w"] = 0.5
    g["packed_confidence_valid"] = 1
    out = oracle.temporal(g, old, 1, 1, 9)[0]
    w = 0.5 + 3.0 * 0.25 * 3
    assert out["w"] == pytest.approx(w, rel=1e-6)
    assert out["packed_confidence_valid"] & 0x7FFFFFFF == 4
    assert out["W"] == pytest.approx(w / (4 * p_hat(out)), rel=1e-6)


def test_temporal_rejects_dissimilar_normals_and_skips_sky():
    rng = np.random.default_rng(4)
    g = make_res(2, rng, sky_frac=0.0)
    old = g.copy()
    old["sample_normal"][0] = oracle.normalised_to_u32([1.0, 0.0, 0.0])  # 90 degrees off
    old["packed_confidence_valid"] = 7
    g["ty"][1] = 0xFF
    out = oracle.temporal(g, old, 2, 1, 0)
    # old reservoir reset: only the current confidence remains
    assert out["packed_confidence_valid"][0] & 0x7FFFFFFF == g["packed_confidence_valid"][0] & 0x7FFFFFFF
    assert out[1].tobytes() == g[1].tobytes()  # sky: untouched


def test_spatial_without_neighbours_is_the_unbiased_weight():
    rng = np.random.default_rng(5)
    g = make_res(4 * 3, rng, sky_frac=0.0)
    info = rng.uniform(0, 1, size=(12, 9)).astype(np.float32)
    img, written = oracle.spatial(g, info, 4, 3, 0, 61, 0)
    assert written.all()
    for k in range(12):
        conf = int(g["packed_confidence_valid"][k]) & 0x7FFFFFFF
        W = g["w"][k] / (conf * p_hat(g[k]))
        expect = g["out_radiance"][k] * W * info[k, 6:9] + info[k, 3:6]
        np.testing.assert_allclose(img.reshape(-1, 4)[k, :3], expect, rtol=1e-5)


def test_spatial_merges_only_similar_diffuse_neighbours():
    rng = np.random.default_rng(6)
    w, h = 16, 16
    g = make_res(w * h, rng, diffuse_frac=1.0, sky_frac=0.0, normal_spread=0.0)
    info = np.zeros((w * h, 9), np.float32)
    info[:, 6:9] = 1.0
    img, written = oracle.spatial(g, info, w, h, 8, 5, 1)
    assert written.all() and np.isfinite(img).all()
    # metallic centres never merge: same result as with no neighbours
    g2 = g.copy()
    g2["ty"] = 1
    a, _ = oracle.spatial(g2, info, w, h, 8, 5, 1)
    b, _ = oracle.spatial(g2, info, w, h, 0, 5, 1)
    np.testing.assert_array_equal(a, b)


def test_eof_gi_confidence_cap():
    rng = np.random.default_rng(7)
    g = make_res(100, rng)
    out = oracle.eof_gi(g, 20)
    for a, b in zip(g, out):
        c = int(a["packed_confidence_valid"]) & 0x7FFFFFFF
        if c > 20:
            assert int(b["packed_confidence_valid"]) & 0x7FFFFFFF == 20
            assert b["packed_confidence_valid"] >> 31 == a["packed_confidence_valid"] >> 31
            assert b["w"] == pytest.approx(a["w"] * np.float32(20) / np.float32(c), rel=1e-6)
        else:
            assert a.tobytes() == b.tobytes()


def test_eof_world_decay_and_timeout():
    w = np.zeros(4, oracle.WORLD_DTYPE)
    w["secondary_hash"] = 77
    w["num_samples"] = [9, 10, 0, 5]
    w["luminance"] = [[9, 9, 9], [10, 20, 30], [1, 1, 1], [5, 5, 5]]
    w["timeout"] = [3, 1, 2, 0]
    out = oracle.eof_world(w)
    assert out["num_samples"][0] == 8 and out["timeout"][0] == 2          # ceil(72/9)
    np.testing.assert_allclose(out["luminance"][0], 8.0)                  # radiance kept
    assert out["num_samples"][1] == 9                                     # ceil(80/9)
    np.testing.assert_allclose(out["luminance"][1], np.array([10, 20, 30]) * np.float32(0.9),
                               rtol=1e-6)
    assert out["num_samples"][2] == 0 and out["timeout"][2] == 1          # empty: only timeout
    assert out["secondary_hash"][3] == 0 and out["num_samples"][3] == 0   # evicted
    assert (out["luminance"][3] == 0).all() and out["timeout"][3] == 0
