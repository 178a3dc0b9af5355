"""Pins the oracle's guided-mode functions (no GPU), each by a check that
does not restate the function:

* strip_NaN (general.wesl:120-135): IEEE special values;
* grid_loc / primary_hasher / secondary_hasher (general.wesl:148-188):
  hand-computed vectors that isolate each prime, the ``| 1``, the octant
  hash, XOR-linearity in the cell, floor(log2) at powers of two, the 30-shift
  cap, negative cells and f32 -> i32 saturation;
* vmf_pdf and its x / (e^x - 1) helper (general.wesl:1018-1047):
  normalisation over S^2 by quadrature and the closed-form peak
  kappa / (2 pi (1 - e^-2kappa));
* the vMF sampling steps (general.wesl:797-808): chi-square against the
  closed-form vMF CDF in cos(theta) and uniform azimuth;
* get_kappa (general.wesl:1000-1016): the Np = 0.2, rp = 0.99 prior, float64
  r(3 - r^2) / (1 - r^2), Banerjee's estimate recovering a known kappa from
  the exact mean resultant length coth k - 1/k, the upper clamp;
* expm1 helpers (general.wesl:1025-1047): float64 on both branches;
* update_marcov_state (general.wesl:464-492): float64 closed forms over 1-3
  updates, alpha = 1 at n = 1, the 0.01 floor, the 1024 clamp, the not_vmf
  reset, the zero-direction refusal, acceptance probability 21 p / sum_mc;
* resample_mcmm (general.wesl:404-425): selection probability proportional
  to weight_sum (chi-square), the weight sum, 42 seed advances;
* all_vmf_pdfs (general.wesl:984-998): the mixture integrates to one;
* grid_load / grid_store / radiance read / add (general.wesl:192-311):
  sequential semantics (key claim, refusal, old-table fallback, timeout,
  clamps 100 / 1000, lock, sample-count saturation, NaN scrubbing).

Parity unpinned (no runnable reference): the reference takes
u32(log2(len)) with the device's f32 log2, whose rounding just below a power
of two is not specified; here the exponent is exact.  NaN -> i32 (chosen: 0)
and clamp(NaN) in get_kappa (chosen: the lower bound) are
implementation-defined in WGSL.
"""
import math

import numpy as np
import pytest
from scipy import stats

import oracle

PRIMES = (1, 687603047, 571899763, 894271171)            # general.wesl:159
SECONDARY_PRIMES = (1699644091, 550105817, 956114303, 155096983)  # general.wesl:170
INCOMING_PRIME = 840057787                               # general.wesl:165
M32 = 0xFFFFFFFF


def pcg(s: int) -> int:
    state = (s * 747796405 + 2891336453) & M32
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M32
    return ((word >> 22) ^ word) & M32


def f32(x) -> float:
    return float(np.float32(x))


# ------------------------------------------------------------------ strip_NaN
def test_strip_nan_special_values():
    assert oracle.strip_nan(float("inf")) == 10000.0
    assert oracle.strip_nan(float("-inf")) == 10000.0
    for bits in (0x7FC00000, 0xFFC00000, 0x7F800001, 0xFF800001, 0x7FFFFFFF):
        v = float(np.array([bits], np.uint32).view(np.float32)[0])
        assert oracle.strip_nan(v) == 0.0
    for v in (0.0, -0.0, 1.5, -2.25, 3.4028234663852886e38, -3.4028234663852886e38, 1e-45, 1.17549435e-38):
        assert oracle.strip_nan(v) == f32(v)


# ------------------------------------------------------------------ hash grid
def key(pos, normal=(0.0, 1.0, 0.0), incoming=(-1.0, -1.0, -1.0), cam=(0.0, 0.0, 0.0), table=1000):
    return oracle.grid_key(pos, normal, incoming, cam, table)


def expect_hashes(loc, nbits, octant):
    lx, ly, lz = (int(v) & M32 for v in loc)
    to_hash = (lx, ly, lz, nbits)
    h = [(p * v) & M32 for p, v in zip(PRIMES, to_hash)]
    primary = (((h[0] ^ h[1]) ^ h[2]) ^ h[3]) ^ ((pcg(octant) * INCOMING_PRIME) & M32)
    s = [(p * v) & M32 for p, v in zip(SECONDARY_PRIMES, to_hash)]
    secondary = (pcg((s[0] ^ s[1]) ^ s[2]) ^ s[3]) | 1
    return primary, secondary


@pytest.mark.parametrize("pos,loc", [
    ((0.05, 0.05, 0.05), (0, 0, 0)),
    ((0.25, 0.05, 0.05), (2, 0, 0)),
    ((0.05, 0.35, 0.05), (0, 3, 0)),
    ((0.05, 0.05, 0.45), (0, 0, 4)),
    ((0.25, 0.35, 0.45), (2, 3, 4)),
    ((-0.35, 0.05, 0.05), (-3, 0, 0)),
])
def test_grid_key_isolates_each_prime(pos, loc):
    k = key(pos)
    np.testing.assert_array_equal(k["loc"], loc)
    primary, secondary = expect_hashes(loc, 0, 0)   # normal +y packs to 0; octant 0
    assert k["primary"] == primary
    assert k["secondary"] == secondary
    assert k["idx"] == primary % 1000
    assert k["secondary"] & 1 == 1


def test_grid_key_normal_and_octant_terms():
    # normal +x: pack2x16snorm(1, 0) = 0x00007FFF, masked by 0xFF00FF00
    n = oracle.normalised_to_u32((1.0, 0.0, 0.0))
    assert n == 0x00007FFF
    for incoming, octant in [((1, -1, -1), 1), ((-1, 1, -1), 2), ((-1, -1, 1), 4), ((1, 1, 1), 7),
                             ((0.0, 0.0, 0.0), 0)]:
        k = key((0.25, 0.35, 0.45), normal=(1.0, 0.0, 0.0), incoming=incoming)
        primary, secondary = expect_hashes((2, 3, 4), n & 0xFF00FF00, octant)
        assert (k["primary"], k["secondary"]) == (primary, secondary), incoming
    # the secondary hash ignores the incoming direction (general.wesl:174)
    a = key((0.25, 0.35, 0.45), incoming=(1, 1, 1))
    b = key((0.25, 0.35, 0.45), incoming=(-1, -1, -1))
    assert a["secondary"] == b["secondary"] and a["primary"] != b["primary"]


def test_grid_key_primary_is_xor_linear_in_the_cell():
    z = key((0.05, 0.05, 0.05))["primary"]
    x = key((0.25, 0.05, 0.05))["primary"]
    y = key((0.05, 0.35, 0.05))["primary"]
    w = key((0.05, 0.05, 0.45))["primary"]
    xyz = key((0.25, 0.35, 0.45))["primary"]
    assert xyz == x ^ y ^ w ^ z ^ z


def test_grid_loc_level_of_detail_shifts():
    # len = max |pos - cam| = L along y only; cell x = 73 (7.35 / 0.1)
    def lx(L):
        return int(key((7.35, 0.0, 0.0), cam=(7.35, -L, 0.0))["loc"][0])
    assert int(key((7.35, 0.0, 0.0), cam=(7.35, 0.0, 0.0))["loc"][0]) == 73   # len 0 -> 1
    assert lx(1.0) == 73                                  # log2 1 = 0
    assert lx(float(np.nextafter(np.float32(2.0), np.float32(0)))) == 73   # just below 2
    assert lx(2.0) == 72                                  # shift 1
    assert lx(3.99) == 72
    assert lx(4.0) == 72                                  # shift 2: 73 & ~3
    assert lx(16.0) == 64                                 # shift 4
    assert lx(128.0) == 0                                 # shift 7
    # the shift is capped at 30 (general.wesl:185): only the top two bits stay
    big = key((1e30, 0.0, 0.0))["loc"]
    assert int(big[0]) & M32 == 0x7FFFFFFF & 0xC0000000   # saturated, masked
    assert int(key((-1e30, 0.0, 0.0))["loc"][0]) & M32 == 0x80000000
    # negative cells mask in two's complement: -3 & ~1 = -4
    assert int(key((-0.35, 0.0, 0.0), cam=(-0.35, -2.5, 0.0))["loc"][0]) == -4
    # i32 conversion truncates toward zero
    assert int(key((-0.05, 0.0, 0.0))["loc"][0]) == 0


def test_grid_key_table_modulo():
    for table in (1, 7, 1000, 1 << 20, 0x80000001):
        k = key((0.25, 0.35, 0.45), table=table)
        assert k["idx"] == k["primary"] % table


# ------------------------------------------------------------------ vMF pdf
MEAN = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])


def basis(m):
    a = np.array([1.0, 0.0, 0.0]) if abs(m[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(m, a)
    u /= np.linalg.norm(u)
    return u, np.cross(m, u)


def sphere_integral(pdf_of_dirs, kappa):
    """Integral over S^2 in (s = 1 - cos theta, phi) about MEAN, float64."""
    u, v = basis(MEAN)
    edges = [0.0, min(2.0, 10.0 / kappa), min(2.0, 40.0 / kappa), 2.0]
    xg, wg = np.polynomial.legendre.leggauss(200)
    total = 0.0
    phis = np.arange(8) * (2 * np.pi / 8)
    for a, b in zip(edges[:-1], edges[1:]):
        if b <= a:
            continue
        s = 0.5 * (b - a) * xg + 0.5 * (b + a)
        ws = 0.5 * (b - a) * wg
        c = 1.0 - s
        sin = np.sqrt(np.clip(1.0 - c * c, 0.0, None))
        acc = np.zeros_like(s)
        for phi in phis:
            d = (c[:, None] * MEAN + sin[:, None] * (np.cos(phi) * u + np.sin(phi) * v))
            acc += pdf_of_dirs(d).astype(np.float64)
        total += float(np.sum(ws * acc / len(phis))) * 2 * np.pi
    return total


@pytest.mark.parametrize("kappa", [1e-3, 0.5, 8.0, 200.0])
def test_vmf_pdf_integrates_to_one(kappa):
    integral = sphere_integral(lambda d: oracle.vmf_pdf(MEAN, kappa, d), kappa)
    assert abs(integral - 1.0) < 2e-3, integral


@pytest.mark.parametrize("kappa", [1e-6, 1e-3, 0.5, 1.0, 8.0, 200.0, 1e4])
def test_vmf_pdf_peak_closed_form(kappa):
    k = float(np.float32(kappa))
    peak = k / (2 * np.pi * -math.expm1(-2 * k))
    got = float(oracle.vmf_pdf(MEAN, kappa, MEAN.reshape(1, 3))[0])
    assert abs(got / peak - 1.0) < 1e-5, (got, peak)
    # a direction at 90 degrees: peak * e^-kappa
    u, _ = basis(MEAN)
    side = float(oracle.vmf_pdf(MEAN, kappa, u.reshape(1, 3))[0])
    assert abs(side - peak * math.exp(-k)) <= 1e-5 * peak + 1e-38


# ------------------------------------------------------------------ vMF sampling
@pytest.mark.parametrize("kappa", [0.5, 8.0, 200.0])
def test_vmf_sampling_chi_square(kappa):
    rng = np.random.default_rng(7)
    seeds = rng.integers(0, 2 ** 32, size=200_000, dtype=np.uint64).astype(np.uint32)
    xs = oracle.vmf_sample(MEAN.astype(np.float32), kappa, seeds).astype(np.float64)
    norms = np.linalg.norm(xs, axis=1)
    assert np.all(np.abs(norms - 1.0) < 1e-5)
    d = xs / norms[:, None]
    c = np.clip(d @ MEAN, -1.0, 1.0)
    k = float(np.float32(kappa))
    # equal-probability bins of the vMF CDF in c: F(c) = (e^{k(c-1)} - e^{-2k}) / (1 - e^{-2k})
    nb = 50
    q = np.arange(1, nb) / nb
    e2 = math.exp(-2 * k)
    cuts = 1.0 + np.log(e2 + q * (1 - e2)) / k
    counts = np.bincount(np.searchsorted(cuts, c), minlength=nb)
    assert stats.chisquare(counts).pvalue > 1e-4, counts
    # azimuth about the mean is uniform
    u, v = basis(MEAN)
    phi = np.arctan2(d @ v, d @ u)
    keep = c < 1.0 - 1e-6
    counts = np.histogram(phi[keep], bins=32, range=(-np.pi, np.pi))[0]
    assert stats.chisquare(counts).pvalue > 1e-4, counts


# ------------------------------------------------------------------ kappa
def kappa_f64(n, mc, ws):
    n2 = float(np.float32(n * n))
    r = (n2 * (mc / ws) + 0.2 * 0.99) / (n2 + 0.2)
    return min(max((3 * r - r ** 3) / (1 - r * r), 0.001), 8.50705867e37)


def test_kappa_prior_at_zero_samples():
    r = 0.99
    expect = (3 * r - r ** 3) / (1 - r * r)
    for mc, ws in [(0.1, 1.0), (0.5, 2.0), (3.0, 4.0)]:
        got = oracle.get_kappa(oracle.chain_words(mean_cosine=mc, weight_sum=ws, num_samples=0))
        assert abs(got / expect - 1.0) < 1e-4, got


def test_kappa_against_float64():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        n = int(rng.integers(1, 1025))
        ws = f32(rng.uniform(0.01, 50.0))
        mc = f32(ws * rng.uniform(0.0, 0.995))
        got = oracle.get_kappa(oracle.chain_words(mean_cosine=mc, weight_sum=ws, num_samples=n))
        expect = kappa_f64(n, mc, ws)
        assert abs(got / expect - 1.0) < 2e-3, (n, mc, ws, got, expect)


@pytest.mark.parametrize("true_kappa", [2.0, 5.0, 20.0, 100.0])
def test_kappa_recovers_known_concentration(true_kappa):
    # the mean resultant length of a 3-D vMF is coth(k) - 1/k (Banerjee et al. 2005)
    rbar = 1.0 / math.tanh(true_kappa) - 1.0 / true_kappa
    got = oracle.get_kappa(oracle.chain_words(mean_cosine=f32(rbar), weight_sum=1.0, num_samples=1024))
    assert abs(got / true_kappa - 1.0) < 0.05, got


def test_kappa_upper_clamp_and_lower_clamp():
    # r rounds to exactly 1 in f32 at N = 1024: (3 - 1) / 0 = inf -> clamp
    got = oracle.get_kappa(oracle.chain_words(mean_cosine=1.0, weight_sum=1.0, num_samples=1024))
    assert got == f32(8.50705867e37)
    # r = 0 (mean cosine 0 and many samples): kappa -> 0, clamped at 0.001
    got = oracle.get_kappa(oracle.chain_words(mean_cosine=0.0, weight_sum=1.0, num_samples=1024))
    assert got == f32(0.001)


# ------------------------------------------------------------------ expm1 helpers
XS = [1e-30, -1e-30, 1e-8, -1e-8, 1e-4, -3e-3, 0.3, -0.7, 0.999, -0.999, 1.0, -1.0, 2.0, -2.0,
      10.0, -10.0, -80.0, -100.0, 50.0, 88.0]


@pytest.mark.parametrize("x", XS)
def test_expm1_helpers_against_float64(x):
    xf = f32(x)
    e = math.expm1(xf)
    got = oracle.expm1(xf)
    assert abs(got - e) <= 1e-6 * abs(e) + 1e-45, (x, got, e)
    got2 = oracle.x_div_expm1(xf)
    ref2 = xf / e if e != 0 else 1.0
    assert abs(got2 - ref2) <= 1e-6 * abs(ref2), (x, got2, ref2)


def test_expm1_helpers_on_a_dense_range():
    for x in np.linspace(-60.0, 60.0, 4001, dtype=np.float32):
        x = float(x)
        if x == 0.0:
            continue
        e = math.expm1(x)
        assert abs(oracle.expm1(x) - e) <= 2e-6 * abs(e)
        assert abs(oracle.x_div_expm1(x) - x / e) <= 2e-6 * abs(x / e)


# ------------------------------------------------------------------ update_marcov_state
X1 = np.array([1.0, 2.0, 3.0], np.float32)


def nrm(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def upd(chain, x2, p, conf=1.0, sum_mc=0.0, not_vmf=False, x1=X1, seed=12345):
    return oracle.update_state(x1, np.asarray(x2, np.float32), p, seed, not_vmf, chain, conf, sum_mc)


def step_f64(state, x1, x2, p, conf):
    """float64 of one accepted update (general.wesl:478-490)."""
    ls, mc, ws, n = state
    n = min(n + 1, 1024)
    alpha = 1.0 if n == 1 else max(1.0 / n, 0.01) * conf
    ws = ws * (1 - alpha) + p * alpha
    ls = ls * (1 - alpha) + p * np.asarray(x2, np.float64) * alpha
    mean = nrm(ls / ws - x1)
    mc = mc * (1 - alpha) + p * float(nrm(np.asarray(x2, np.float64) - x1) @ mean) * alpha
    return ls, mc, ws, n


def close_to(words, ref, tol=2e-6):
    f = oracle.chain_fields(words)
    ls, mc, ws, n = ref
    assert f["num_samples"] == n
    assert abs(f["weight_sum"] - ws) <= tol * abs(ws)
    assert abs(f["mean_cosine"] - mc) <= tol * max(abs(mc), 1.0)
    np.testing.assert_allclose(f["light_source"], ls, rtol=tol, atol=tol)


def test_update_three_steps_closed_form():
    x1 = X1.astype(np.float64)
    steps = [((4.0, -1.0, 2.0), 0.7, 1.0), ((0.0, 5.0, 1.0), 0.3, 1.0), ((2.0, 2.0, -3.0), 1.9, 0.5)]
    w = oracle.chain_words()
    ref = (np.zeros(3), 0.0, 0.0, 0)
    for x2, p, conf in steps:
        w = upd(w, x2, p, conf)
        ref = step_f64(ref, x1, x2, float(np.float32(p)), conf)
        close_to(w, ref)
    # n = 1: alpha = 1, so the first update takes the sample whole
    first = oracle.chain_fields(upd(oracle.chain_words(), (4.0, -1.0, 2.0), 0.7))
    assert first["weight_sum"] == f32(0.7)
    np.testing.assert_array_equal(first["light_source"], np.float32(0.7) * np.float32([4, -1, 2]))


def test_update_floor_clamp_and_confidence():
    x1 = X1.astype(np.float64)
    ls0 = np.array([3.0, 1.0, 0.5], np.float32)
    for n0, conf, alpha in [(1023, 1.0, 0.01), (1024, 1.0, 0.01), (1023, 0.5, 0.005), (9, 0.5, 0.05)]:
        c = oracle.chain_words(light_source=ls0, mean_cosine=0.2, weight_sum=1.0, num_samples=n0)
        out = oracle.chain_fields(upd(c, (0.0, 5.0, 1.0), 2.0, conf))
        assert out["num_samples"] == min(n0 + 1, 1024)
        assert abs(out["weight_sum"] - (1.0 * (1 - alpha) + 2.0 * alpha)) < 1e-6


def test_update_not_vmf_resets_and_zero_direction_refuses():
    c = oracle.chain_words(light_source=(3.0, 1.0, 0.5), mean_cosine=0.2, weight_sum=1.0, num_samples=40)
    a = upd(c, (4.0, -1.0, 2.0), 0.7, not_vmf=True)
    b = upd(oracle.chain_words(), (4.0, -1.0, 2.0), 0.7)
    np.testing.assert_array_equal(a, b)
    # get_light_dir of the zero chain from the origin is the zero vector: no update
    z = upd(oracle.chain_words(), (4.0, -1.0, 2.0), 0.7, x1=np.zeros(3, np.float32))
    np.testing.assert_array_equal(z, oracle.chain_words())


def test_update_acceptance_probability():
    # accepted iff rand * sum_mc < 21 p: probability 21 p / sum_mc when below one
    rng = np.random.default_rng(5)
    seeds = rng.integers(0, 2 ** 32, size=20000, dtype=np.uint64)
    c0 = oracle.chain_words()
    acc = sum(oracle.chain_fields(upd(c0, (4.0, -1.0, 2.0), 1.0, sum_mc=84.0, seed=int(s)))["num_samples"]
              for s in seeds)
    assert stats.binomtest(acc, len(seeds), 0.25).pvalue > 1e-4, acc
    # sum_mc of zero always accepts (0 < 21 p for p > 0)
    assert oracle.chain_fields(upd(c0, (4.0, -1.0, 2.0), 1e-6, sum_mc=0.0))["num_samples"] == 1


# ------------------------------------------------------------------ resample_mcmm
def chains_with(weights):
    out = np.zeros((21, 8), np.uint32)
    for k, w in enumerate(weights):
        out[k] = oracle.chain_words(light_source=(k, 1.0, 2.0), mean_cosine=0.1, weight_sum=w,
                                    num_samples=k)
    return out


def test_resample_selects_proportionally_to_weight():
    w = np.zeros(21, np.float32)
    w[[0, 3, 4, 10, 16, 19, 20]] = [1.0, 2.0, 0.5, 4.0, 1.5, 3.0, 0.25]
    ch = chains_with(w)
    counts = np.zeros(22, np.int64)
    for s in range(30000):
        k, _, sm = oracle.resample_mcmm(ch, s * 2654435761 & M32)
        counts[k + 1] += 1
    assert counts[0] == 0
    assert counts[1:][w == 0].sum() == 0
    p = w[w > 0] / w.sum()
    assert stats.chisquare(counts[1:][w > 0], p * counts.sum()).pvalue > 1e-4, counts
    # the running sum is the f32 sum in candidate order
    s = np.float32(0)
    for v in w:
        s = np.float32(s + v)
    assert sm == float(s)


def test_resample_zero_weights_and_seed_advance():
    ch = chains_with(np.zeros(21))
    for seed in (0, 1, 0xDEADBEEF):
        k, after, sm = oracle.resample_mcmm(ch, seed)
        assert k == -1 and sm == 0.0
        s = seed
        for _ in range(42):   # two rand_u32 per candidate
            s = pcg(s)
        assert after == s


def test_all_vmf_pdfs_mixture_integrates_to_one():
    ch = np.zeros((21, 8), np.uint32)
    hit = np.array([0.5, -0.25, 1.0], np.float32)
    rng = np.random.default_rng(11)
    total = 0.0
    for k in (2, 7, 17, 20):
        d = nrm(rng.normal(size=3))
        ws = float(rng.uniform(0.5, 2.0))
        ls = (hit + 3.0 * d) * ws
        ch[k] = oracle.chain_words(light_source=ls, mean_cosine=ws * 0.6, weight_sum=ws, num_samples=4)
        total += np.float32(ws)
    pts = 200_000
    i = np.arange(pts) + 0.5
    z = 1 - 2 * i / pts
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    dirs = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    vals = np.array([oracle.all_vmf_pdfs(ch, total, hit, d) for d in dirs[::20]])
    integral = vals.mean() * 4 * np.pi
    assert abs(integral - 1.0) < 0.02, integral
    # no valid candidate or sum_mc of zero: pdf 0
    assert oracle.all_vmf_pdfs(np.zeros((21, 8), np.uint32), 1.0, hit, (0, 1, 0)) == 0.0
    assert oracle.all_vmf_pdfs(ch, 0.0, hit, (0, 1, 0)) == 0.0


# ------------------------------------------------------------------ world grid semantics
CAM = (0.0, 0.0, 0.0)
P1 = ((0.25, 0.35, 0.45), (0.0, 1.0, 0.0), (1.0, 1.0, 1.0))
P2 = ((0.55, 0.35, 0.45), (0.0, 1.0, 0.0), (1.0, 1.0, 1.0))   # another key


def fresh(table=1):
    return np.zeros(table, oracle.WORLD_DTYPE), np.zeros(table, oracle.WORLD_DTYPE)


def test_grid_store_claims_and_refuses():
    cur, old = fresh()
    c = oracle.chain_words(light_source=(1, 2, 3), mean_cosine=0.5, weight_sum=2.0, num_samples=3)
    np.testing.assert_array_equal(oracle.grid_op(oracle.GRID_LOAD, cur, old, CAM, *P1),
                                  oracle.chain_words())
    oracle.grid_op(oracle.GRID_STORE, cur, old, CAM, *P1, chain=c)
    sec = oracle.grid_key(*P1, CAM, 1)["secondary"]
    assert cur[0]["secondary_hash"] == sec and cur[0]["timeout"] == 30
    cur[0]["timeout"] = 5
    np.testing.assert_array_equal(oracle.grid_op(oracle.GRID_LOAD, cur, old, CAM, *P1), c)
    assert cur[0]["timeout"] == 30
    # table of one cell: P2 collides on the index with another secondary hash
    before = cur.copy()
    oracle.grid_op(oracle.GRID_STORE, cur, old, CAM, *P2, chain=oracle.chain_words(weight_sum=9.0))
    assert cur.tobytes() == before.tobytes()
    np.testing.assert_array_equal(oracle.grid_op(oracle.GRID_LOAD, cur, old, CAM, *P2),
                                  oracle.chain_words())
    # the previous frame's table answers when the current one holds another key
    old[0] = cur[0]
    cur[0]["secondary_hash"] = oracle.grid_key(*P2, CAM, 1)["secondary"]
    np.testing.assert_array_equal(oracle.grid_op(oracle.GRID_LOAD, cur, old, CAM, *P1), c)
    # force_old reads the old table even when the current one matches
    cur2, old2 = fresh()
    oracle.grid_op(oracle.GRID_STORE, cur2, old2, CAM, *P1, chain=c)
    np.testing.assert_array_equal(oracle.grid_op(oracle.GRID_LOAD, cur2, old2, CAM, *P1, force_old=True),
                                  oracle.chain_words())


def test_grid_radiance_add_clamps_lock_and_saturation():
    cur, old = fresh()
    z = oracle.grid_op(oracle.GRID_RADIANCE, cur, old, CAM, *P1)
    np.testing.assert_array_equal(z, 0)
    oracle.grid_op(oracle.GRID_ADD, cur, old, CAM, *P1, radiance=(1.0, 2.0, 3.0))
    oracle.grid_op(oracle.GRID_ADD, cur, old, CAM, *P1, radiance=(3.0, 2.0, 1.0))
    assert cur[0]["num_samples"] == 2
    np.testing.assert_array_equal(oracle.grid_op(oracle.GRID_RADIANCE, cur, old, CAM, *P1), [2, 2, 2])
    # add clamps each channel to [0, 100]
    oracle.grid_op(oracle.GRID_ADD, cur, old, CAM, *P1, radiance=(500.0, -5.0, 50.0))
    np.testing.assert_array_equal(cur[0]["luminance"], np.float32([104.0, 4.0, 54.0]))
    # the read clamps the mean to [0, 1000]
    cur[0]["luminance"] = (5000.0, 1.0, 2.0)
    cur[0]["num_samples"] = 1
    np.testing.assert_array_equal(oracle.grid_op(oracle.GRID_RADIANCE, cur, old, CAM, *P1), [1000, 1, 2])
    # a held lock makes the add give up
    cur[0]["lock"] = 1
    before = cur.copy()
    oracle.grid_op(oracle.GRID_ADD, cur, old, CAM, *P1, radiance=(1.0, 1.0, 1.0))
    assert cur.tobytes() == before.tobytes()
    cur[0]["lock"] = 0
    # a saturated sample count adds nothing but refreshes the timeout
    cur[0]["num_samples"] = 0x0FFFFFFF
    cur[0]["timeout"] = 3
    oracle.grid_op(oracle.GRID_ADD, cur, old, CAM, *P1, radiance=(1.0, 1.0, 1.0))
    assert cur[0]["num_samples"] == 0x0FFFFFFF and cur[0]["timeout"] == 30
    np.testing.assert_array_equal(cur[0]["luminance"], np.float32([5000.0, 1.0, 2.0]))
    # another key on the cell: nothing is added
    cur[0]["num_samples"] = 1
    before = cur.copy()
    oracle.grid_op(oracle.GRID_ADD, cur, old, CAM, *P2, radiance=(1.0, 1.0, 1.0))
    assert cur.tobytes() == before.tobytes()


def test_grid_scrubs_nan_and_inf():
    cur, old = fresh()
    bad = oracle.chain_words(light_source=(float("nan"), float("inf"), 1.0), mean_cosine=float("-inf"),
                             weight_sum=float("nan"), num_samples=2, score=float("inf"))
    oracle.grid_op(oracle.GRID_STORE, cur, old, CAM, *P1, chain=bad)
    f = oracle.chain_fields(oracle.grid_op(oracle.GRID_LOAD, cur, old, CAM, *P1))
    np.testing.assert_array_equal(f["light_source"], [0.0, 10000.0, 1.0])
    assert (f["mean_cosine"], f["weight_sum"], f["score"], f["num_samples"]) == (10000.0, 0.0, 10000.0, 2)
    cur[0]["luminance"] = (float("inf"), 1.0, 1.0)
    cur[0]["num_samples"] = 1
    oracle.grid_op(oracle.GRID_ADD, cur, old, CAM, *P1, radiance=(1.0, 1.0, 1.0))
    np.testing.assert_array_equal(cur[0]["luminance"], np.float32([10000.0, 2.0, 2.0]))


# ------------------------------------------------------------------ screen-space candidates
def test_screen_draw_lane_repeats_and_window():
    # general.wesl:369-372: the lane is drawn from the seed without advancing
    # it; the 12 neighbours then chain the hash twice each (x, then y)
    for seed, x, y, width, sub in [(5, 10, 20, 100, 64), (0xFFFFFFFF, 70, 0, 100, 64), (77, 3, 9, 40, 8)]:
        d = oracle.screen_draw(seed, x, y, width, sub)
        base = (x // sub) * sub
        lane = pcg(seed) % sub
        assert d[0] == y * width + base + lane
        assert d[1] == int(base + lane < width)
        s = seed
        for k in range(12):
            xr = pcg(s)
            yr = pcg(xr)
            s = yr
            dx, dy = xr % 91 - 45, yr % 91 - 45
            assert d[2 + k] == ((x + dx) + (y + dy) * width) & M32
        assert d[14] == s


def test_screen_draw_offsets_uniform_in_the_window():
    width = 1000
    x, y = 500, 500
    counts = np.zeros((91, 91), np.int64)
    for seed in range(4000):
        d = oracle.screen_draw(seed * 7919, x, y, width)
        idx = d[2:14].astype(np.int64)
        dy = idx // width - y
        dx = idx % width - x
        assert (np.abs(dx) <= 45).all() and (np.abs(dy) <= 45).all()
        np.add.at(counts, (dy + 45, dx + 45), 1)
    assert stats.chisquare(counts.sum(axis=0)).pvalue > 1e-4
    assert stats.chisquare(counts.sum(axis=1)).pvalue > 1e-4
