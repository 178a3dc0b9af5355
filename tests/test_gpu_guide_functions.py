"""The guided mode's device functions (lambent_amd/csrc/dev_guide.h) on the
GPU against the CPU oracle, one function at a time, through the test-only
probe library tests/native/libguide_probe.so, which calls the same inline
device code that the guided kernels use:

* grid_key (general.wesl:148-188): index and secondary hash bit-exact;
* get_kappa (general.wesl:1000-1016): bit-exact (no transcendental);
* update_marcov_state (general.wesl:464-492): bit-exact chains;
* resample_mcmm's step (general.wesl:404-425): selection, seed and weight
  sum bit-exact;
* screen-space candidate draws (general.wesl:367-385): bit-exact;
* world-grid operations (general.wesl:192-311), sequential: bit-exact cells;
* expm1 helpers, vmf_pdf, vMF sampling, the mixture pdf
  (general.wesl:797-808, 984-1047): within a few ulp (the device's
  expf / logf / sinf / cosf against the host's);
* the per-wave writer election of the final world-grid update elects one
  lane per (index, secondary) key among the live lanes.
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

LIB = Path(__file__).resolve().parent / "native" / "libguide_probe.so"
_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)


_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_vp = ctypes.c_void_p
SIGNATURES = {
    "probe_grid_key": [_f32p, _f32p, _f32p, _u64, _f32p, _u32, _u32p],
    "probe_kappa": [_u32p, _u64, _f32p],
    "probe_expm1": [_f32p, _u64, _f32p],
    "probe_vmf_pdf": [_f32p, _f32p, _f32p, _u64, _f32p],
    "probe_vmf_sample": [_f32p, _f32p, _u32p, _u64, _f32p],
    "probe_update_state": [_f32p, _f32p, _f32p, _u32p, _u32p, _u32p, _f32p, _f32p, _u64, _u32p],
    "probe_resample": [_u32p, _u32p, _u64, _i32p, _u32p, _f32p],
    "probe_mixture_pdf": [_u32p, _f32p, _f32p, _f32p, _u64, _f32p],
    "probe_screen_candidates": [_u32p, _u32p, _u64, _u32, _u32, _u32p],
    "probe_first_writer": [_u32p, _u64, _u32p],
    "probe_grid_ops": [_vp, _vp, _u32, _f32p, _u32p, _f32p, _u32p, _u32p, _u32, _u32p, _f32p],
}


@pytest.fixture(scope="module")
def probe():
    if not LIB.exists():
        pytest.fail(f"{LIB} is missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(str(LIB))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    return lib


def F(a):
    return np.ascontiguousarray(a, np.float32).ctypes.data_as(_f32p)


def U(a):
    return np.ascontiguousarray(a, np.uint32).ctypes.data_as(_u32p)


def ok(rc):
    assert rc == 0, f"probe call failed with hipError {rc}"


def rand_unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def random_chains(rng, n):
    c = np.zeros((n, 8), np.uint32)
    ws = rng.uniform(0.01, 20.0, n).astype(np.float32)
    ws[rng.random(n) < 0.1] = 0.0
    f = np.zeros((n, 7), np.float32)
    f[:, :3] = rng.normal(scale=10.0, size=(n, 3)) * ws[:, None]
    f[:, 3] = ws * rng.uniform(0.0, 0.999, n)
    f[:, 4] = ws
    f[:, 6] = rng.uniform(0, 1, n)
    c[:, :7] = f.view(np.uint32)
    c[:, 5] = rng.integers(0, 1025, n)
    return c


def test_grid_key_bit_exact(probe):
    rng = np.random.default_rng(1)
    n = 20000
    pos = np.concatenate([rng.uniform(-30, 30, (n // 2, 3)), rng.normal(scale=0.7, size=(n // 2, 3))])
    pos = pos.astype(np.float32)
    nrm = rand_unit(rng, n)
    nrm[::7] = (0.0, 1.0, 0.0)
    nrm[1::7] = (0.0, -1.0, 0.0)
    inc = rand_unit(rng, n)
    cam = np.array([0.3, -25.0, 1.5], np.float32)
    for table in (1_500_007, 1000, 1):
        out = np.zeros((n, 2), np.uint32)
        ok(probe.probe_grid_key(F(pos), F(nrm), F(inc), n, F(cam), table, U(out)))
        ref = np.array([[k["idx"], k["secondary"]] for k in
                        (oracle.grid_key(pos[i], nrm[i], inc[i], cam, table) for i in range(n))], np.uint32)
        np.testing.assert_array_equal(out, ref)


def test_kappa_bit_exact(probe):
    rng = np.random.default_rng(2)
    c = random_chains(rng, 20000)
    out = np.zeros(len(c), np.float32)
    ok(probe.probe_kappa(U(c), len(c), F(out)))
    ref = np.array([oracle.get_kappa(w) for w in c], np.float32)
    np.testing.assert_array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_update_state_bit_exact(probe):
    rng = np.random.default_rng(3)
    n = 20000
    x1 = rng.uniform(-10, 10, (n, 3)).astype(np.float32)
    x2 = rng.uniform(-10, 10, (n, 3)).astype(np.float32)
    p = rng.uniform(0, 5, n).astype(np.float32)
    seed = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    not_vmf = (rng.random(n) < 0.2).astype(np.uint32)
    chains = random_chains(rng, n)
    conf = np.where(rng.random(n) < 0.5, 1.0, 0.5).astype(np.float32)
    sum_mc = np.where(rng.random(n) < 0.3, 0.0, rng.uniform(0, 200, n)).astype(np.float32)
    out = np.zeros((n, 8), np.uint32)
    ok(probe.probe_update_state(F(x1), F(x2), F(p), U(seed), U(not_vmf), U(chains), F(conf),
                                F(sum_mc), n, U(out)))
    ref = np.stack([oracle.update_state(x1[i], x2[i], float(p[i]), int(seed[i]), bool(not_vmf[i]),
                                        chains[i], float(conf[i]), float(sum_mc[i])) for i in range(n)])
    changed = (ref != chains).any(axis=1)
    assert changed.mean() > 0.3   # the cases include accepted updates
    np.testing.assert_array_equal(out, ref)


def test_resample_bit_exact(probe):
    rng = np.random.default_rng(4)
    n = 4000
    chains = random_chains(rng, n * 21).reshape(n, 21, 8)
    chains[: n // 10] = 0   # all-zero candidate sets select nothing
    seeds = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    sel = np.zeros(n, np.int32)
    after = np.zeros(n, np.uint32)
    sm = np.zeros(n, np.float32)
    ok(probe.probe_resample(U(chains), U(seeds), n, sel.ctypes.data_as(_i32p), U(after), F(sm)))
    for i in range(n):
        k, s, m = oracle.resample_mcmm(chains[i], int(seeds[i]))
        assert (sel[i], after[i]) == (k, s), i
        assert np.float32(m).view(np.uint32) == sm[i].view(np.uint32), i


def test_screen_candidates_bit_exact(probe):
    rng = np.random.default_rng(5)
    n = 20000
    for width, sub in [(1920, 64), (100, 64), (37, 8)]:
        xy = np.stack([rng.integers(0, width, n), rng.integers(0, 200, n)], axis=1).astype(np.uint32)
        seeds = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        out = np.zeros((n, 15), np.uint32)
        ok(probe.probe_screen_candidates(U(seeds), U(xy), n, width, sub, U(out)))
        ref = np.stack([oracle.screen_draw(int(seeds[i]), int(xy[i, 0]), int(xy[i, 1]), width, sub)
                        for i in range(n)])
        np.testing.assert_array_equal(out, ref)


def test_expm1_helpers_within_ulps(probe):
    x = np.concatenate([np.linspace(-100, 60, 20001), np.geomspace(1e-30, 0.999, 2000),
                        -np.geomspace(1e-30, 0.999, 2000)]).astype(np.float32)
    out = np.zeros((len(x), 2), np.float32)
    ok(probe.probe_expm1(F(x), len(x), F(out)))
    ref = np.array([[oracle.expm1(float(v)), oracle.x_div_expm1(float(v))] for v in x], np.float32)
    np.testing.assert_allclose(out, ref, rtol=4e-6, atol=0)


def test_vmf_pdf_and_sampling_within_ulps(probe):
    rng = np.random.default_rng(6)
    n = 20000
    mean = rand_unit(rng, n)
    kappa = np.exp(rng.uniform(np.log(1e-3), np.log(2e3), n)).astype(np.float32)
    dirs = rand_unit(rng, n)
    dirs[::3] = mean[::3] + 0.01 * dirs[::3]
    out = np.zeros(n, np.float32)
    ok(probe.probe_vmf_pdf(F(mean), F(kappa), F(dirs), n, F(out)))
    ref = np.array([oracle.vmf_pdf(mean[i], float(kappa[i]), dirs[i:i + 1])[0] for i in range(n)],
                   np.float32)
    # exp(kappa (c - 1)): an ulp of the argument is an ulp-scaled factor
    arg = np.abs(kappa.astype(np.float64) * (np.sum(dirs * mean, axis=1) - 1.0))
    tol = 4e-6 * (1.0 + arg)
    assert (np.abs(out - ref) <= tol * np.abs(ref) + 1e-37).all()
    seeds = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    xs = np.zeros((n, 3), np.float32)
    ok(probe.probe_vmf_sample(F(mean), F(kappa), U(seeds), n, F(xs)))
    ref = np.stack([oracle.vmf_sample(mean[i], float(kappa[i]), seeds[i:i + 1])[0] for i in range(n)])
    err = np.abs(xs - ref).max(axis=1)
    assert np.quantile(err, 0.999) < 1e-5, np.quantile(err, [0.5, 0.99, 0.999, 1.0])
    assert err.max() < 1e-3


def test_mixture_pdf_within_ulps(probe):
    rng = np.random.default_rng(7)
    n = 3000
    chains = random_chains(rng, n * 21).reshape(n, 21, 8)
    hit = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    d = rand_unit(rng, n)
    ws = chains[:, :, 4].view(np.float32)
    sum_mc = ws.sum(axis=1, dtype=np.float32)
    sum_mc[::10] = 0.0
    out = np.zeros(n, np.float32)
    ok(probe.probe_mixture_pdf(U(chains), F(sum_mc), F(hit), F(d), n, F(out)))
    ref = np.array([oracle.all_vmf_pdfs(chains[i], float(sum_mc[i]), hit[i], d[i]) for i in range(n)],
                   np.float32)
    np.testing.assert_allclose(out, ref, rtol=5e-5, atol=1e-30)


def test_world_grid_operations_bit_exact(probe):
    rng = np.random.default_rng(8)
    table = 61
    cur = np.zeros(table, oracle.WORLD_DTYPE)
    old = np.zeros(table, oracle.WORLD_DTYPE)
    # a previous frame's table with some of the same keys
    cam = np.array([0.0, 0.0, -3.0], np.float32)
    pts = rng.uniform(-2, 2, (40, 3)).astype(np.float32)
    nrm = rand_unit(rng, 40)
    inc = rand_unit(rng, 40)
    for i in range(0, 40, 3):
        oracle.grid_op(oracle.GRID_STORE, old, cur, cam, pts[i], nrm[i], inc[i],
                       chain=random_chains(rng, 1)[0])
        oracle.grid_op(oracle.GRID_ADD, old, cur, cam, pts[i], nrm[i], inc[i], radiance=(1, 2, 3))
    cur[:] = 0
    nops = 800
    ops = rng.integers(0, 5, nops).astype(np.uint32)  # 4: fused add and train
    which = rng.integers(0, 40, nops)
    args = np.zeros((nops, 12), np.float32)
    args[:, 0:3] = pts[which]
    args[:, 3:6] = nrm[which]
    args[:, 6:9] = inc[which]
    args[:, 9:12] = rng.uniform(-20, 150, (nops, 3))
    force = (rng.random(nops) < 0.3).astype(np.uint32)
    chains = random_chains(rng, nops)
    gpu_cur = cur.copy()
    cout = np.zeros((nops, 8), np.uint32)
    rout = np.zeros((nops, 3), np.float32)
    ok(probe.probe_grid_ops(gpu_cur.ctypes.data_as(ctypes.c_void_p), old.ctypes.data_as(ctypes.c_void_p),
                            table, F(cam), U(ops), F(args), U(force), U(chains), nops, U(cout), F(rout)))
    for i in range(nops):
        if ops[i] == 4:
            # grid_add_and_train: the add, then (flag set) load, update, store
            geo = (cam, args[i, 0:3], args[i, 3:6], args[i, 6:9])
            oracle.grid_op(oracle.GRID_ADD, cur, old, *geo, radiance=args[i, 9:12])
            if force[i]:
                wc = oracle.grid_op(oracle.GRID_LOAD, cur, old, *geo)
                np.testing.assert_array_equal(cout[i], wc, err_msg=f"op {i}")
                st = wc.copy()
                st[0:3] = chains[i][0:3]
                st[3] = (wc[3:4].view(np.float32) * np.float32(0.5)).view(np.uint32)[0]
                st[4] = (wc[4:5].view(np.float32) + np.float32(1.0)).view(np.uint32)[0]
                st[5] = (int(wc[5]) + 3) & 0xFFFFFFFF
                st[6] = chains[i][6]
                oracle.grid_op(oracle.GRID_STORE, cur, old, *geo, chain=st)
            continue
        r = oracle.grid_op(int(ops[i]), cur, old, cam, args[i, 0:3], args[i, 3:6], args[i, 6:9],
                           force_old=bool(force[i]), chain=chains[i], radiance=args[i, 9:12])
        if ops[i] == oracle.GRID_LOAD:
            np.testing.assert_array_equal(cout[i], r, err_msg=f"op {i}")
        elif ops[i] == oracle.GRID_RADIANCE:
            np.testing.assert_array_equal(rout[i].view(np.uint32), r.view(np.uint32), err_msg=f"op {i}")
    assert gpu_cur.tobytes() == cur.tobytes()
    assert (cur["secondary_hash"] != 0).sum() > 5


def test_writer_election_per_key(probe):
    def run(keys, live):
        out = np.zeros(64, np.uint32)
        ok(probe.probe_first_writer(U(np.asarray(keys, np.uint32).reshape(-1)), live, U(out)))
        return out

    all_live = (1 << 64) - 1
    same = np.tile([[17, 0x1001]], (64, 1))
    np.testing.assert_array_equal(run(same, all_live), [1] + [0] * 63)
    # lane 0 not live: lane 1 is elected
    live = all_live & ~1
    np.testing.assert_array_equal(run(same, live), [0, 1] + [0] * 62)
    # one index, two secondary hashes: one writer per key
    k = same.copy()
    k[1::2, 1] = 0x2003
    out = run(k, all_live)
    np.testing.assert_array_equal(out, [1, 1] + [0] * 62)
    # random keys: exactly one writer per distinct live key, the lowest lane
    rng = np.random.default_rng(9)
    for _ in range(20):
        k = np.stack([rng.integers(0, 4, 64), rng.choice([1, 3, 5], 64)], axis=1).astype(np.uint32)
        live = int(rng.integers(0, 2 ** 63)) | (1 << 63)
        out = run(k, live)
        seen = set()
        for lane in range(64):
            if not (live >> lane) & 1:
                assert out[lane] == 0
                continue
            key = (int(k[lane, 0]), int(k[lane, 1]))
            assert out[lane] == (key not in seen), lane
            seen.add(key)
