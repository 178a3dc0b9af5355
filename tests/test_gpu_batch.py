"""Batched tracer on the GPU: every image of a batch must equal, bit for bit,
what the single-scene tracer renders for the same scene, camera, options and
seed - for every launch configuration, with and without the captured graph,
across graph replays interleaved with other GPU work, and after autotuning."""
import numpy as np
import pytest

from lambent_amd import api, scenes

pytestmark = pytest.mark.gpu

OPTS = api.PathTracerOpts(samples=2, bounces=6, rr_start=3, guiding=api.Guiding.OFF,
                          restir_temporal=api.RestirTemporalMode.Disabled)


@pytest.fixture(scope="module")
def items():
    specs = [
        scenes.cornell_box(64, 48),
        scenes.procedural(2_000, seed=11, width=96, height=64),
        scenes.procedural(20_000, seed=12, width=128, height=128, mix=(0.67, 0.33, 0.0)),
        scenes.furnace_box(0.5, 1.0, width=37, height=29),
        scenes.procedural(1_000, seed=13, width=50, height=70),
    ]
    return [(sp.build(), sp.camera(), sp.width, sp.height) for sp in specs]


def single(items, seed, quality=api.Medium):
    pipe = api.RayTracer(quality).create_pipeline(OPTS)
    outs, verts = [], 0
    for scene, cam, w, h in items:
        b = api.DataBuffers(w, h, 1, api.BufferType(0))
        pipe.dispatch(scene, cam, b, seed)
        outs.append(b.output(api.Output.COLOR))
        verts += b.path_vertices()  # both tracers reuse the first sample's camera-ray hits
    return outs, verts


def check(batch, ref):
    for k, r in enumerate(ref):
        np.testing.assert_array_equal(batch.output(k), r, err_msg=f"item {k}")


CONFIGS = [
    api.BatchConfig(),
    api.BatchConfig(use_graph=False),
    api.BatchConfig(tile_w=16, tile_h=4, blocks_per_cu=2, refill=4),
    api.BatchConfig(tile_w=64, tile_h=1, pass_pixels=4096, blocks_per_cu=16, refill=32),
    api.BatchConfig(tile_w=3, tile_h=5, pass_pixels=1000, use_graph=False),
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: f"t{c.tile_w}x{c.tile_h}-b{c.blocks_per_cu}"
                         f"-r{c.refill}-p{c.pass_pixels}-g{int(c.use_graph)}")
def test_batch_equals_single_scene_tracer(items, cfg):
    ref, verts = single(items, 5)
    batch = api.Batch(items, OPTS, api.Medium, cfg)
    batch.trace(5)
    check(batch, ref)
    assert batch.path_vertices() == verts


def test_graph_replays_interleaved_with_other_work(items):
    batch = api.Batch(items, OPTS, api.Medium, api.BatchConfig(use_graph=True))
    for seed in (7, 8, 9, 7):
        batch.trace(seed)
        api.sync()
        ref, _ = single(items, seed)  # eager single-scene launches between replays
        check(batch, ref)


def test_spectral_quality_batched(items):
    ref, _ = single(items, 3, api.Spectral)
    batch = api.Batch(items, OPTS, api.Spectral)
    batch.trace(3)
    check(batch, ref)


def test_camera_update(items):
    batch = api.Batch(items, OPTS, api.Medium)
    spec = scenes.cornell_box(64, 48)
    proj = api.perspective(60.0, 64 / 48, 0.1, 200.0)
    cam = api.Camera.from_proj_view(proj, api.look_to_rh((0.5, 0.4, -0.8), (0.0, 0.1, 1.0),
                                                         (0.0, 1.0, 0.0)))
    batch.trace(1)
    batch.set_camera(0, cam)
    batch.trace(1)
    b = api.DataBuffers(64, 48, 1, api.BufferType(0))
    api.RayTracer(api.Medium).create_pipeline(OPTS).dispatch(items[0][0], cam, b, 1)
    np.testing.assert_array_equal(batch.output(0), b.output(api.Output.COLOR))


def test_autotune_picks_a_candidate_and_keeps_the_images(items):
    ref, _ = single(items, 2)
    batch = api.Batch(items, OPTS, api.Medium)
    cands = [api.BatchConfig(tile_w=8, tile_h=8), api.BatchConfig(tile_w=32, tile_h=2, refill=8),
             api.BatchConfig(blocks_per_cu=4, pass_pixels=8192)]
    best, ms = batch.autotune(cands, frames=2)
    assert best in cands
    assert len(ms) == 3 and all(m > 0 for m in ms)
    assert batch.config == best
    batch.trace(2)
    check(batch, ref)


def test_sample_groups_with_a_remainder(items, monkeypatch):
    """5 samples traced in groups of 2, 2 and 1 (the group size capped by
    LMB_BATCH_GROUP_PATHS): the start seeds carry over from group to group
    as from sample to sample, so the images and the path-vertex count equal
    the single-scene tracer's."""
    opts = api.PathTracerOpts(samples=5, bounces=5, rr_start=2, guiding=api.Guiding.OFF,
                              restir_temporal=api.RestirTemporalMode.Disabled)
    npix = sum(w * h for _, _, w, h in items)
    pipe = api.RayTracer(api.Medium).create_pipeline(opts)
    ref, verts = [], 0
    for scene, cam, w, h in items:
        b = api.DataBuffers(w, h, 1, api.BufferType(0))
        pipe.dispatch(scene, cam, b, 4)
        ref.append(b.output(api.Output.COLOR))
        verts += b.path_vertices()
    for cap in (2 * npix, 0):  # groups of 2; one sample at a time
        monkeypatch.setenv("LMB_BATCH_GROUP_PATHS", str(cap))
        batch = api.Batch(items, opts, api.Medium)
        batch.trace(4)
        check(batch, ref)
        assert batch.path_vertices() == verts
