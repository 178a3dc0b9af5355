"""GPU ReSTIR passes and end-of-frame processing against the oracle.

Inputs are written into the reservoir / info / chain buffers through the C
ABI; every arithmetic operation involved is correctly rounded (add, mul,
div, sqrt, fma) so the GPU results must equal the oracle's bit for bit.
"""
import numpy as np
import pytest

import oracle
from lambent_amd import api
from tests.test_cpu_restir import make_res

pytestmark = pytest.mark.gpu

R = oracle.RESERVOIR_DTYPE
SB = api.StateBuffer


def _bufs(w, h, world=1):
    return api.DataBuffers(w, h, world, api.BufferType.all())


def _read(b, which, dtype):
    return np.frombuffer(b.state_bytes(which), dtype).copy()


@pytest.mark.parametrize("seed_offset", [0, 12345])
def test_temporal_bit_exact(seed_offset):
    w, h = 96, 40
    rng = np.random.default_rng(seed_offset + 1)
    gi = make_res(w * h, rng, normal_spread=0.3)
    old = make_res(w * h, rng, normal_spread=0.3)
    b = _bufs(w, h)
    b.write_state(SB.GI_RESERVOIRS, gi.tobytes())
    b.write_state(SB.OLD_GI_RESERVOIRS, old.tobytes())
    b.temporal_resample(seed_offset)
    got = _read(b, SB.GI_RESERVOIRS, R)
    ref = oracle.temporal(gi, old, w, h, seed_offset)
    assert got.tobytes() == ref.tobytes()


@pytest.mark.parametrize("is_samples,is_space", [(8, 61), (5, 61), (4, 31), (0, 61)])
def test_spatial_bit_exact(is_samples, is_space):
    w, h = 80, 48
    rng = np.random.default_rng(is_samples * 100 + is_space)
    gi = make_res(w * h, rng, normal_spread=0.1)
    info = rng.uniform(0, 1, size=(w * h, 9)).astype(np.float32)
    b = _bufs(w, h)
    b.write_state(SB.GI_RESERVOIRS, gi.tobytes())
    b.write_state(SB.INFO, info.tobytes())
    b.spatial_resample(is_samples, is_space, 7)
    got = b.output(api.Output.COLOR)
    ref, written = oracle.spatial(gi, info, w, h, is_samples, is_space, 7)
    assert written.mean() > 0.9
    assert got[written].tobytes() == ref[written].tobytes()
    # the pass never writes the reservoirs (spatial:179 is commented out)
    assert b.state_bytes(SB.GI_RESERVOIRS) == gi.tobytes()


def test_advance_frame_matches_reference_order():
    w, h, world = 40, 30, 5000
    rng = np.random.default_rng(9)
    gi = make_res(w * h, rng)
    gi["packed_confidence_valid"] = (rng.integers(1, 100, size=w * h).astype(np.uint32)
                                     | (np.uint32(1) << np.uint32(31)))
    mc = rng.integers(0, 2 ** 32, size=(w * h, 8), dtype=np.uint64).astype(np.uint32)
    wmc = np.zeros(world, oracle.WORLD_DTYPE)
    wmc["secondary_hash"] = rng.integers(1, 2 ** 31, size=world)
    wmc["num_samples"] = rng.integers(0, 5000, size=world)
    wmc["luminance"] = rng.uniform(0, 50, size=(world, 3))
    wmc["timeout"] = rng.integers(0, 31, size=world)
    wmc["chain_weight_sum"] = rng.uniform(0, 1, size=world)
    b = _bufs(w, h, world)
    b.write_state(SB.GI_RESERVOIRS, gi.tobytes())
    b.write_state(SB.MARKOV_CHAINS, mc.tobytes())
    b.write_state(SB.WORLD_MARKOV_CHAINS, wmc.tobytes())
    b.advance_frame(api.BufferType.all(), api.AdvanceOptions(20))
    assert b.state_bytes(SB.OLD_GI_RESERVOIRS) == oracle.eof_gi(gi, 20).tobytes()
    assert b.state_bytes(SB.GI_RESERVOIRS) == bytes(64 * w * h)
    assert b.state_bytes(SB.OLD_MARKOV_CHAINS) == mc.tobytes()
    assert b.state_bytes(SB.MARKOV_CHAINS) == bytes(32 * w * h)
    exp_world = oracle.eof_world(wmc).tobytes()
    assert b.state_bytes(SB.WORLD_MARKOV_CHAINS) == exp_world
    assert b.state_bytes(SB.OLD_WORLD_MARKOV_CHAINS) == exp_world
    # without the flags every buffer is cleared
    b.advance_frame(api.BufferType(0), api.AdvanceOptions.RESPONSIVE)
    for sb in SB:
        if sb == SB.INFO:
            continue
        data = b.state_bytes(sb)
        assert data == bytes(len(data)), sb


def test_resampling_needs_spatial_buffers():
    b = api.DataBuffers(8, 8, 1, api.BufferType.MARKOV_CHAIN_SCREEN_SPACE)
    with pytest.raises(api.LambentError):
        b.temporal_resample(0)
    with pytest.raises(api.LambentError):
        b.advance_frame(api.BufferType.all(), api.AdvanceOptions(0))
