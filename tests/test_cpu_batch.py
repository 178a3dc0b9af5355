"""Batched tracer host logic that needs no GPU: the autotuner's default
sweep and the cfg4 workload generator."""
import numpy as np

from lambent_amd import api, scenes


def test_default_sweep_is_the_full_grid():
    sweep = api.BatchConfig.default_sweep()
    assert len(sweep) == 3 * 3 * 3 * 2
    keys = {(c.tile_w, c.tile_h, c.blocks_per_cu, c.refill, c.pass_pixels) for c in sweep}
    assert len(keys) == len(sweep)
    assert all(c.use_graph for c in sweep)
    assert {(c.tile_w, c.tile_h) for c in sweep} == {(8, 8), (16, 16), (64, 1)}


def test_batch_workload_distribution():
    specs = scenes.batch_workload(n_scenes=12, tri_min=1_000, tri_max=5_000)
    assert len(specs) == 12
    for sp in specs:
        assert sp.width == sp.height and sp.width in scenes.BATCH_RESOLUTIONS
        n = len(sp.vertices) // 3
        assert 1_000 <= n <= 5_000
        ty = sp.materials[:n, 5] >> 16
        assert set(np.unique(ty)) <= {0, 1}  # diffuse and GGX metal only
