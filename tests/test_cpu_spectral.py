"""Four-wavelength spectral quality (no reference counterpart), pinned on the
CPU oracle: the colour weights average to 1 over 380-780 nm (so a white
furnace stays white), the IOR follows the Cauchy law through the material's
IOR-range endpoints, and collapsing to the hero wavelength at a dispersive
interface turns the mean weight into the hero's exactly."""
import numpy as np

import oracle
from lambent_amd import scenes

LO, HI = scenes.cauchy_range()


def test_cauchy_endpoints():
    a, b = scenes.CAUCHY_A, scenes.CAUCHY_B
    assert abs(LO - (a + b / 0.76 ** 2)) < 1e-12 and abs(HI - (a + b / 0.34 ** 2)) < 1e-12
    assert 1.5 < LO < HI < 1.54


def test_weights_average_to_one():
    seeds = np.arange(1, 200_001, dtype=np.uint32) * 7919
    mean_w = np.zeros(3)
    hero_w = np.zeros(3)
    for s in seeds[:40_000]:
        _, c, p = oracle.spectral_pick(-1.0, 1.5, 1.5, int(s))  # non-dispersive: 4-lambda mean
        assert 0.0 <= p <= 1.0
        mean_w += c
        _, h, p2 = oracle.spectral_pick(-1.0, LO, HI, int(s))  # dispersive: the hero alone
        assert 2.0 <= p2 < 3.0
        hero_w += h
    mean_w /= 40_000
    hero_w /= 40_000
    # E over u of the stratified mean = E over lambda of the weight = 1; the
    # hero's spread is larger (weights up to ~4), so its tolerance is wider
    np.testing.assert_allclose(mean_w, 1.0, atol=0.01)
    np.testing.assert_allclose(hero_w, 1.0, atol=0.03)


def test_cauchy_ior_and_collapse():
    rng = np.random.default_rng(3)
    inv_hi, inv_lo = 1 / 0.34 ** 2, 1 / 0.76 ** 2
    B = (np.float32(HI) - np.float32(LO)) / (inv_hi - inv_lo)
    A = np.float32(LO) - B * inv_lo
    for s in rng.integers(1, 2 ** 32 - 1, 200):
        ri0, c0, p0 = oracle.spectral_pick(-1.0, 1.5, 1.5, int(s))
        assert ri0 == np.float32(1.5)
        ri1, c1, p1 = oracle.spectral_pick(p0, LO, HI, int(s) + 5)
        lam_um = (380.0 + 400.0 * (p0 - np.floor(p0))) * 1e-3
        assert p1 == np.float32(p0 + 2.0)
        assert abs(ri1 - (A + B / lam_um ** 2)) < 2e-6
        # mean weight x collapse ratio = the hero's weight
        _, hero, _ = oracle.spectral_pick(-1.0, LO, HI, int(s))
        np.testing.assert_allclose(c0 * c1, hero, rtol=1e-5, atol=1e-6)
        # a second dispersive interface keeps the hero's IOR and adds no weight
        ri2, c2, p2 = oracle.spectral_pick(p1, LO, HI, int(s) + 9)
        assert ri2 == ri1 and p2 == p1
        np.testing.assert_array_equal(c2, 1.0)
