"""The cached tracer's stratified first-scatter sampler (dev_qmc.h), built
for the CPU (tests/native/libqmc_host.so).

It has no counterpart in the reference, so it is pinned by mathematics: the
first Sobol values of each dimension by hand from the direction numbers
(m = 1, 3, 5 for x + 1; m = 1, 3, 3 for x^2 + x + 1), the (t, m, 2)-net
property of every pair of dimensions over prefixes of 2^m points (t = 0 for
dimensions 0 and 1, which a wrong recurrence breaks), the same after
scrambling (nested uniform scrambling keeps nets), the nesting of the
scramble itself (values that share their top k bits still do afterwards),
and the uniformity of a single point over many scrambling seeds (what makes
every sample unbiased).
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest

LIB_PATH = Path(__file__).resolve().parent / "native" / "libqmc_host.so"


@pytest.fixture(scope="module")
def lib():
    if not LIB_PATH.exists():
        pytest.skip("tests/native/libqmc_host.so is not built")
    lib = ctypes.CDLL(str(LIB_PATH))
    lib.qmc_sobol.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.qmc_sobol.restype = ctypes.c_uint32
    lib.qmc_owen.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.qmc_owen.restype = ctypes.c_uint32
    lib.qmc_points.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.POINTER(ctypes.c_float)]
    return lib


def points(lib, n0, count, stream):
    out = np.empty((count, 3), np.float32)
    lib.qmc_points(n0, count, stream, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out


def test_first_sobol_values_by_hand(lib):
    # dimension 0: radical inverse; 1: v = 1/2, 3/4, 5/8; 2: v = 1/2, 3/4, 3/8
    expect = {
        0: [0, 0.5, 0.25, 0.75, 0.125, 0.625, 0.375, 0.875],
        1: [0, 0.5, 0.75, 0.25, 0.625, 0.125, 0.375, 0.875],
        2: [0, 0.5, 0.75, 0.25, 0.375, 0.875, 0.625, 0.125],
    }
    for d, vals in expect.items():
        got = [lib.qmc_sobol(d, n) / 2.0 ** 32 for n in range(8)]
        assert got == vals, (d, got)


def _net_t(a, b, m):
    """Smallest t such that the 2^m pairs (a, b) (32-bit fractions) put
    2^t points in every elementary box of area 2^(t - m)."""
    n = 1 << m
    for t in range(m + 1):
        k = m - t
        ok = True
        for bits_a in range(k + 1):
            bits_b = k - bits_a
            ia = (a >> np.uint64(32 - bits_a)) if bits_a else np.zeros(n, np.uint64)
            ib = (b >> np.uint64(32 - bits_b)) if bits_b else np.zeros(n, np.uint64)
            cell = (ia << np.uint64(bits_b)) | ib
            if not (np.bincount(cell.astype(np.int64), minlength=1 << k) == (1 << t)).all():
                ok = False
                break
        if ok:
            return t
    return m


@pytest.mark.parametrize("scramble", [False, True])
def test_prefixes_are_nets(lib, scramble):
    seeds = (0x2545F491, 0x9E3779B9, 0x6A09E667)
    for m in range(1, 13):
        n = 1 << m
        cols = []
        for d in range(3):
            v = np.array([lib.qmc_sobol(d, i) for i in range(n)], np.uint64)
            if scramble:
                v = np.array([lib.qmc_owen(int(x), seeds[d]) for x in v], np.uint64)
            cols.append(v)
        assert _net_t(cols[0], cols[1], m) == 0, m
        assert _net_t(cols[0], cols[2], m) <= 1, m
        assert _net_t(cols[1], cols[2], m) <= 1, m


def test_scramble_is_nested(lib):
    """Nested uniform scrambling: the scrambled value's top k bits depend
    only on the input's top k bits, and the map is a bijection on them."""
    rng = np.random.default_rng(3)
    for seed in rng.integers(0, 2 ** 32, 4):
        for k in (1, 3, 8):
            tops = {}
            for x in rng.integers(0, 2 ** 32, 4000):
                y = lib.qmc_owen(int(x), int(seed))
                top_x, top_y = int(x) >> (32 - k), y >> (32 - k)
                assert tops.setdefault(top_x, top_y) == top_y
            if k <= 8:
                assert len(set(tops.values())) == len(tops)


def test_points_of_a_pixel_prefix_stratify_the_unit_square(lib):
    # the 8 samples of a frame (an aligned block of 8 indices) fall one per
    # 1/2 x 1/4, 1/4 x 1/2, 1/8 x 1 and 1 x 1/8 box of dimensions 0 and 1
    for stream in (0, 17, 123456):
        for block in (0, 8, 64):
            p = points(lib, block, 8, stream)
            for bx, by in ((2, 4), (4, 2), (8, 1), (1, 8)):
                cell = np.floor(p[:, 0] * bx).astype(int) * by + np.floor(p[:, 1] * by).astype(int)
                assert sorted(cell) == list(range(8)), (stream, block, bx, by)


def test_single_points_are_uniform_over_pixels(lib):
    """Point n of 40,000 pixel streams: each dimension uniform (chi^2 over
    32 bins), and dimension pairs uncorrelated."""
    for n in (0, 5, 200):
        p = np.array([points(lib, n, 1, s)[0] for s in range(40000)])
        assert (p >= 0).all() and (p < 1).all()
        for d in range(3):
            h = np.bincount(np.floor(p[:, d] * 32).astype(int), minlength=32)
            chi2 = float(((h - 1250.0) ** 2 / 1250.0).sum())
            assert chi2 < 70.0, (n, d, chi2)  # 31 dof: p < 1e-4
        c = np.corrcoef(p.T)
        assert np.abs(c[np.triu_indices(3, 1)]).max() < 0.03
