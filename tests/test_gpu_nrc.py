"""Neural radiance cache on the GPU against the CPU oracle.

Inference and the loss gradient are compared with the oracle's bf16 mode,
which rounds at the same points as the kernels (weights, encoded features,
hidden activations, gradients); the kernels accumulate in fp32 in matrix-core
order, so agreement is within stated tolerances, not bitwise.  The Adam step
is float arithmetic in the same order on both sides and must agree exactly.
"""
import numpy as np
import pytest

import oracle
from lambent_amd import api, scenes

pytestmark = pytest.mark.gpu

NW = api.NRC_NUM_WEIGHTS
WH, W5 = 320 * 64, 320 * 64 + 4 * 64 * 64
LAYERS = [(0, WH)] + [(WH + l * 4096, WH + (l + 1) * 4096) for l in range(4)] + [(W5, NW)]


def he_weights(seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(NW, np.float64)
    a0 = np.sqrt(6.0 / 313)
    w0 = rng.uniform(-a0, a0, (320, 64))
    w0[313:] = 0.0
    w[:WH] = w0.ravel()
    w[WH:] = rng.uniform(-np.sqrt(6.0 / 64), np.sqrt(6.0 / 64), NW - WH)
    return w.astype(np.float32)


def test_weights_roundtrip_and_default_init():
    cache = api.NeuralCache()
    w0 = cache.weights()
    assert np.isfinite(w0).all() and np.abs(w0).max() > 0
    assert (w0[313 * 64:320 * 64] == 0).all()
    w = he_weights(1)
    cache.set_weights(w)
    np.testing.assert_array_equal(cache.weights(), w)


def test_infer_matches_oracle():
    cache = api.NeuralCache()
    w = he_weights(3)
    cache.set_weights(w)
    x = np.random.default_rng(4).random((2000, 13)).astype(np.float32)
    y = cache.infer(x)
    ref = oracle.nrc_forward(w.astype(np.float64), x, oracle.NRC_BF16)
    scale = max(1.0, float(np.abs(ref).max()))
    err = np.abs(y - ref)
    assert err.max() < 2e-2 * scale, err.max()
    assert err.mean() < 2e-3 * scale, err.mean()
    exact = oracle.nrc_forward(w.astype(np.float64), x, oracle.NRC_EXACT)
    assert np.abs(y - exact).max() < 5e-2 * scale


@pytest.mark.parametrize("n", [1, 31, 33, 64, 65, 257])
def test_infer_ragged_sizes_are_independent_of_batch(n):
    cache = api.NeuralCache()
    cache.set_weights(he_weights(5))
    x = np.random.default_rng(n).random((300, 13)).astype(np.float32)
    full = cache.infer(x)
    part = cache.infer(x[:n])
    np.testing.assert_array_equal(part, full[:n])  # per-sample arithmetic, no batch coupling


def test_gradient_matches_oracle():
    cache = api.NeuralCache()
    w = he_weights(6)
    cache.set_weights(w)
    rng = np.random.default_rng(7)
    n = 700  # not a multiple of the 256-sample chunk
    x = rng.random((n, 13)).astype(np.float32)
    t = (rng.random((n, 3)) * 2).astype(np.float32)
    g, loss = cache.gradient(x, t)
    ref, ref_loss = oracle.nrc_grad(w.astype(np.float64), x, t, oracle.NRC_BF16)
    assert abs(loss - ref_loss) < 1e-3 * ref_loss + 1e-6, (loss, ref_loss)
    for a, b in LAYERS:
        ga, ra = g[a:b].astype(np.float64), ref[a:b]
        rel = np.linalg.norm(ga - ra) / np.linalg.norm(ra)
        assert rel < 3e-2, (a, rel)
    assert (g[313 * 64:320 * 64] == 0).all()
    np.testing.assert_array_equal(cache.weights(), w)  # gradient only, no step
    exact, _ = oracle.nrc_grad(w.astype(np.float64), x, t, oracle.NRC_EXACT)
    cos = g.astype(np.float64) @ exact / (np.linalg.norm(g) * np.linalg.norm(exact))
    assert cos > 0.99, cos


def test_optimiser_steps_keep_the_fragments_current():
    """The Adam kernel writes the bf16 fragment elements of the weights it
    updates (forward and backward sets): after optimiser steps, inference and
    the gradient match the oracle at the new master weights."""
    cache = api.NeuralCache(api.NeuralCacheOpts(learning_rate=0.02))
    cache.set_weights(he_weights(10))
    rng = np.random.default_rng(11)
    n = 600
    x = rng.random((n, 13)).astype(np.float32)
    t = (rng.random((n, 3)) * 2).astype(np.float32)
    for _ in range(3):
        cache.train(x, t)
    w = cache.weights()
    assert np.abs(w - he_weights(10)).max() > 0.01  # the steps moved the weights
    y = cache.infer(x)
    ref = oracle.nrc_forward(w.astype(np.float64), x, oracle.NRC_BF16)
    scale = max(1.0, float(np.abs(ref).max()))
    assert np.abs(y - ref).max() < 2e-2 * scale
    g, _ = cache.gradient(x, t)
    gref, _ = oracle.nrc_grad(w.astype(np.float64), x, t, oracle.NRC_BF16)
    for a, b in LAYERS:
        rel = np.linalg.norm(g[a:b].astype(np.float64) - gref[a:b]) / np.linalg.norm(gref[a:b])
        assert rel < 3e-2, (a, rel)


def test_adam_step_matches_oracle_exactly():
    cache = api.NeuralCache()
    w = he_weights(8)
    cache.set_weights(w)
    rng = np.random.default_rng(9)
    m = np.zeros(NW, np.float32)
    v = np.zeros(NW, np.float32)
    ref = w.copy()
    o = cache.opts
    p1 = p2 = 1.0
    for step in range(3):
        g = (rng.standard_normal(NW) * 10.0 ** rng.integers(-6, 0)).astype(np.float32)
        g[::17] = 0.0
        cache.apply_gradient(g)
        p1 *= float(np.float32(o.beta1))
        p2 *= float(np.float32(o.beta2))
        oracle.nrc_adam(ref, m, v, g, np.float32(o.learning_rate), np.float32(o.beta1),
                        np.float32(o.beta2), np.float32(o.epsilon), np.float32(1 - p1),
                        np.float32(1 - p2))
        np.testing.assert_array_equal(cache.weights(), ref)
    assert cache.stats()["steps"] == 3


def test_training_fits_a_smooth_field():
    cache = api.NeuralCache(api.NeuralCacheOpts(learning_rate=3e-3, seed=2))
    rng = np.random.default_rng(10)

    def field(x):
        r = 0.6 + 0.4 * np.sin(2 * np.pi * x[:, 0]) * x[:, 10]
        g = 0.5 + 0.3 * np.cos(np.pi * x[:, 1]) * x[:, 11]
        b = 0.2 + 0.5 * x[:, 8] * x[:, 12]
        return np.stack([r, g, b], 1).astype(np.float32)

    xv = rng.random((4096, 13)).astype(np.float32)
    tv = field(xv)
    first = None
    for it in range(300):
        x = rng.random((4096, 13)).astype(np.float32)
        loss = cache.train(x, field(x))
        first = first if first is not None else loss
    assert np.isfinite(loss)
    assert loss < first / 10, (first, loss)
    y = cache.infer(xv)
    rel = np.mean((y - tv) ** 2 / (tv ** 2 + 0.01))
    assert rel < 0.05, rel  # 300 steps of online Adam: about 0.02 measured


def _rel_mse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 0.01)))


@pytest.mark.parametrize("overlap", [True, False])
def test_cached_tracer_learns_the_indirect_light(overlap):
    spec = scenes.cornell_box(64, 64)
    scene = spec.build()
    cam = spec.camera()
    opts = api.PathTracerOpts(samples=1, bounces=8, rr_start=3, guiding=api.Guiding.OFF)
    pipe = api.RayTracer(api.Medium).create_pipeline(opts)
    ref_pipe = api.RayTracer(api.Medium).create_pipeline(
        api.PathTracerOpts(samples=2048, bounces=8, rr_start=3, guiding=api.Guiding.OFF))
    rb = api.DataBuffers(64, 64, 1)
    ref_pipe.dispatch(scene, cam, rb, 777)
    ref = rb.output(api.Output.COLOR)[..., :3].astype(np.float64)

    def run(cache, frames, skip):
        bufs = api.DataBuffers(64, 64, 1)
        acc = np.zeros((64, 64, 3))
        for f in range(frames):
            pipe.dispatch_cached(scene, cam, bufs, cache, f)
            if f >= skip:
                acc += bufs.output(api.Output.COLOR)[..., :3]
        return acc / (frames - skip)

    cache = api.NeuralCache(api.NeuralCacheOpts(train_paths=1024, train_steps=4, seed=3,
                                                overlap_training=overlap))
    img = run(cache, 300, 100)
    st = cache.stats()
    assert st["steps"] == 300 * 4
    assert 0.6 * 1024 < st["records"] < 1.4 * 1024, st
    assert st["queries"] > 64 * 64 * 0.3, st
    assert np.isfinite(img).all()
    # the same frames with an untrained zero cache: paths truncated at vertex 2
    zero = api.NeuralCache(api.NeuralCacheOpts(train_paths=0, seed=3))
    zero.set_weights(np.zeros(NW, np.float32))
    trunc = run(zero, 200, 0)
    e_nrc, e_trunc = _rel_mse(img, ref), _rel_mse(trunc, ref)
    assert e_nrc < 0.5 * e_trunc, (e_nrc, e_trunc)
    assert abs(img.mean() - ref.mean()) / ref.mean() < 0.1, (img.mean(), ref.mean())
    assert trunc.mean() < img.mean()


def test_stratified_first_scatter_is_unbiased_and_lowers_the_noise():
    """A trained, then frozen cache: frames with the first scatter from each
    pixel's scrambled Sobol sequence (consecutive seed offsets) against the
    same frames from the reference's hash draws, both judged against a long
    hash-draw render of the same cache.  The stratified image has the same
    expectation (image mean within 1%, no per-pixel bias beyond the noise)
    and less noise at equal samples."""
    spec = scenes.cornell_box(64, 64)
    scene = spec.build()
    cam = spec.camera()
    pipe = api.RayTracer(api.Medium).create_pipeline(
        api.PathTracerOpts(samples=4, bounces=8, rr_start=3, guiding=api.Guiding.OFF))
    bufs = api.DataBuffers(64, 64, 1)
    o = api.NeuralCacheOpts(train_paths=1024, train_steps=4, seed=9)
    cache = api.NeuralCache(o)
    for f in range(60):
        pipe.dispatch_cached(scene, cam, bufs, cache, f)

    def render(stratify, frames, first):
        cache.set_opts(api.NeuralCacheOpts(**{**o.__dict__, "train_paths": 0,
                                              "stratify": stratify}))
        acc = np.zeros((64, 64, 3))
        for f in range(frames):
            pipe.dispatch_cached(scene, cam, bufs, cache, first + f)
            acc += bufs.output(api.Output.COLOR)[..., :3]
        return acc / frames

    w = cache.weights()
    conv = render(False, 1024, 100_000)
    strat, hashed = render(True, 32, 0), render(False, 32, 50_000)
    np.testing.assert_array_equal(cache.weights(), w)  # frozen throughout
    e_s, e_h = _rel_mse(strat, conv), _rel_mse(hashed, conv)
    assert e_s < 0.6 * e_h, (e_s, e_h)
    assert abs(strat.mean() - conv.mean()) / conv.mean() < 0.01, (strat.mean(), conv.mean())
    # per-pixel bias: the stratified error has no mean beyond the noise of
    # the hash-draw error's spread
    d = (strat - conv).mean(axis=2)
    assert abs(d.mean()) < 4 * (hashed - conv).mean(axis=2).std() / 64, d.mean()


def test_ema_inference_weights():
    d = 0.75
    cache = api.NeuralCache(api.NeuralCacheOpts(ema_decay=d, learning_rate=0.05))
    w = he_weights(12)
    cache.set_weights(w)
    x = np.random.default_rng(13).random((512, 13)).astype(np.float32)
    ref0 = oracle.nrc_forward(w.astype(np.float64), x, oracle.NRC_BF16)
    assert np.abs(cache.infer(x) - ref0).max() < 2e-2 * max(1.0, np.abs(ref0).max())
    g = (np.random.default_rng(14).standard_normal(NW) * 1e-3).astype(np.float32)
    cache.apply_gradient(g)
    w1 = cache.weights()
    assert np.abs(w1 - w).max() > 0.04  # an lr-0.05 Adam step moved the masters
    ema = (np.float32(d) * w + np.float32(1 - d) * w1).astype(np.float32)
    y = cache.infer(x)
    ref_ema = oracle.nrc_forward(ema.astype(np.float64), x, oracle.NRC_BF16)
    ref_raw = oracle.nrc_forward(w1.astype(np.float64), x, oracle.NRC_BF16)
    scale = max(1.0, np.abs(ref_ema).max())
    assert np.abs(y - ref_ema).max() < 2e-2 * scale
    assert np.abs(y - ref_ema).mean() < 0.2 * np.abs(y - ref_raw).mean()


def test_set_opts_freezes_training_in_the_tracer():
    spec = scenes.cornell_box(32, 32)
    scene = spec.build()
    pipe = api.RayTracer(api.Medium).create_pipeline(
        api.PathTracerOpts(samples=1, bounces=8, rr_start=3, guiding=api.Guiding.OFF))
    bufs = api.DataBuffers(32, 32, 1)
    o = api.NeuralCacheOpts(train_paths=256, train_steps=2)
    cache = api.NeuralCache(o)
    for f in range(3):
        pipe.dispatch_cached(scene, spec.camera(), bufs, cache, f)
    assert cache.stats()["steps"] == 6
    w = cache.weights()
    cache.set_opts(api.NeuralCacheOpts(**{**o.__dict__, "train_paths": 0}))
    for f in range(3, 6):
        pipe.dispatch_cached(scene, spec.camera(), bufs, cache, f)
    st = cache.stats()
    assert st["steps"] == 6 and st["records"] == 0, st
    np.testing.assert_array_equal(cache.weights(), w)
    assert np.isfinite(bufs.output(api.Output.COLOR)).all()
    with pytest.raises(api.LambentError):
        cache.set_opts(api.NeuralCacheOpts(train_steps=0))


@pytest.mark.parametrize("overlap", [True, False])
def test_cached_tracer_several_samples_per_trace(overlap):
    """Training paths come from the first sample only (or from the training
    pass beside the samples); later samples reuse its camera-ray hits and end
    at their cache vertices."""
    spec = scenes.cornell_box(48, 48)
    scene = spec.build()
    pipe = api.RayTracer(api.Medium).create_pipeline(
        api.PathTracerOpts(samples=4, bounces=8, rr_start=3, guiding=api.Guiding.OFF))
    bufs = api.DataBuffers(48, 48, 1)
    cache = api.NeuralCache(api.NeuralCacheOpts(train_paths=512, train_steps=2, seed=4,
                                                overlap_training=overlap))
    for f in range(40):
        pipe.dispatch_cached(scene, spec.camera(), bufs, cache, f)
    st = cache.stats()
    assert st["steps"] == 80
    assert 0.6 * 512 < st["records"] < 1.4 * 512, st
    assert st["queries"] > 48 * 48 * 0.5, st  # the last sample: no training paths
    img = bufs.output(api.Output.COLOR)[..., :3]
    assert np.isfinite(img).all() and img.mean() > 0


def test_training_frames_train_without_an_image():
    """lmb_trace_nrc_train: about train_paths records from random pixels, the
    optimiser steps taken, the output buffers untouched; a cache warmed up by
    training frames alone renders the Cornell box much closer to a 4096-spp
    reference than the truncated paths alone (measured: relMSE 0.093 vs 0.242;
    300 full cached frames give 0.090)."""
    spec = scenes.cornell_box(64, 64)
    scene = spec.build()
    cam = spec.camera()
    opts = api.PathTracerOpts(samples=1, bounces=8, rr_start=3, guiding=api.Guiding.OFF)
    pipe = api.RayTracer(api.Medium).create_pipeline(opts)
    ref_pipe = api.RayTracer(api.Medium).create_pipeline(
        api.PathTracerOpts(samples=4096, bounces=8, rr_start=3, guiding=api.Guiding.OFF))
    rb = api.DataBuffers(64, 64, 1)
    ref_pipe.dispatch(scene, cam, rb, 999)
    ref = rb.output(api.Output.COLOR)[..., :3].astype(np.float64)
    bufs = api.DataBuffers(64, 64, 1)
    pipe.dispatch(scene, cam, bufs, 1)
    before = bufs.output(api.Output.COLOR)
    cache = api.NeuralCache(api.NeuralCacheOpts(train_paths=1024, train_steps=4, seed=5))
    for f in range(300):
        pipe.train_cached(scene, cam, bufs, cache, 10 + f)
    st = cache.stats()
    assert st["steps"] == 1200
    assert 0.6 * 1024 < st["records"] < 1.4 * 1024, st
    np.testing.assert_array_equal(bufs.output(api.Output.COLOR), before)
    frozen = api.NeuralCacheOpts(train_paths=0, seed=5)

    def render(c):
        c.set_opts(frozen)
        acc = np.zeros((64, 64, 3))
        for f in range(200):
            pipe.dispatch_cached(scene, cam, bufs, c, 5000 + f)
            acc += bufs.output(api.Output.COLOR)[..., :3]
        return acc / 200

    untrained = api.NeuralCache(api.NeuralCacheOpts(seed=5))
    untrained.set_weights(np.zeros(NW, np.float32))  # the truncated paths alone
    img = render(cache)
    e_trained, e_untrained = _rel_mse(img, ref), _rel_mse(render(untrained), ref)
    assert e_trained < 0.5 * e_untrained, (e_trained, e_untrained)
    assert abs(img.mean() - ref.mean()) / ref.mean() < 0.06, (img.mean(), ref.mean())
    with pytest.raises(api.LambentError):
        pipe.train_cached(scene, cam, bufs, api.NeuralCache(frozen), 0)


def _frozen_render(pipe, scene, cam, w, h, weights, seed):
    frozen = api.NeuralCache(api.NeuralCacheOpts(train_paths=0))
    frozen.set_weights(weights)
    fb = api.DataBuffers(w, h, 1)
    pipe.dispatch_cached(scene, cam, fb, frozen, seed)
    return fb.output(api.Output.COLOR), fb


def _same_render(a, b):
    """Images of the same paths traced one sample at a time and several
    samples at a time: equal but for the order of the per-pixel float sums."""
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("samples", [3, 11])
def test_overlapped_training_renders_with_the_previous_frames_weights(samples):
    """With overlap_training the samples of frame k read the inference weights
    published by frame k - 1's last optimiser step while frame k's own
    training pass and steps run on the cache's stream: frame k's image equals
    a frozen cache's render with the weights after frame k - 1.  The frozen
    cache traces one sample at a time; the overlapped frame traces up to 8
    together (11 samples: a group of 8 and one of 3), so the images agree to
    the rounding of the per-pixel sums."""
    spec = scenes.cornell_box(48, 40)
    scene = spec.build()
    cam = spec.camera()
    pipe = api.RayTracer(api.Medium).create_pipeline(
        api.PathTracerOpts(samples=samples, bounces=8, rr_start=3, guiding=api.Guiding.OFF))
    bufs = api.DataBuffers(48, 40, 1)
    cache = api.NeuralCache(api.NeuralCacheOpts(train_paths=512, train_steps=3, seed=6,
                                                learning_rate=3e-3))
    for f in range(4):
        pipe.dispatch_cached(scene, cam, bufs, cache, f)
    w = cache.weights()
    pipe.dispatch_cached(scene, cam, bufs, cache, 4)
    img = bufs.output(api.Output.COLOR)
    st = cache.stats()
    assert st["steps"] == 15
    assert 0.6 * 512 < st["records"] < 1.4 * 512, st
    assert np.abs(cache.weights() - w).max() > 0  # frame 4 trained on
    assert np.isfinite(img).all() and img[..., :3].mean() > 0
    frozen, fb = _frozen_render(pipe, scene, cam, 48, 40, w, 4)
    _same_render(img, frozen)
    np.testing.assert_array_equal(bufs.output(api.Output.NORMAL), fb.output(api.Output.NORMAL))
    np.testing.assert_array_equal(bufs.output(api.Output.ALBEDO), fb.output(api.Output.ALBEDO))
    assert bufs.path_vertices() == fb.path_vertices()


def test_switching_training_modes_keeps_inference_current():
    """Host inference after overlapped frames, synchronous frames between
    overlapped ones, and new weights set between overlapped frames all see
    the current weights."""
    spec = scenes.cornell_box(40, 32)
    scene = spec.build()
    cam = spec.camera()
    pipe = api.RayTracer(api.Medium).create_pipeline(
        api.PathTracerOpts(samples=2, bounces=8, rr_start=3, guiding=api.Guiding.OFF))
    bufs = api.DataBuffers(40, 32, 1)
    o = api.NeuralCacheOpts(train_paths=256, train_steps=2, seed=7, learning_rate=3e-3)
    cache = api.NeuralCache(o)
    for f in range(3):
        pipe.dispatch_cached(scene, cam, bufs, cache, f)
    x = np.random.default_rng(21).random((256, 13)).astype(np.float32)
    ref = oracle.nrc_forward(cache.weights().astype(np.float64), x, oracle.NRC_BF16)
    assert np.abs(cache.infer(x) - ref).max() < 2e-2 * max(1.0, np.abs(ref).max())
    cache.set_opts(api.NeuralCacheOpts(**{**o.__dict__, "overlap_training": False}))
    pipe.dispatch_cached(scene, cam, bufs, cache, 3)
    cache.set_opts(o)
    pipe.dispatch_cached(scene, cam, bufs, cache, 4)
    assert cache.stats()["steps"] == 10
    w = he_weights(31)
    cache.set_weights(w)
    pipe.dispatch_cached(scene, cam, bufs, cache, 5)
    _same_render(bufs.output(api.Output.COLOR), _frozen_render(pipe, scene, cam, 40, 32, w, 5)[0])


@pytest.mark.parametrize("quality", [api.Low, api.High, api.Spectral])
def test_overlapped_training_other_qualities(quality):
    """The grouped samples of an overlapped frame against a frozen cache's
    one-sample-at-a-time render with the same weights, for the qualities
    other than Medium (Low: no roughness, every metal and dielectric scatter
    specular; High: one wavelength per path; Spectral: four)."""
    spec = scenes.cornell_box(40, 40)
    scene = spec.build()
    cam = spec.camera()
    pipe = api.RayTracer(quality).create_pipeline(
        api.PathTracerOpts(samples=5, bounces=8, rr_start=3, guiding=api.Guiding.OFF))
    bufs = api.DataBuffers(40, 40, 1)
    cache = api.NeuralCache(api.NeuralCacheOpts(train_paths=512, train_steps=2, seed=8,
                                                learning_rate=3e-3))
    for f in range(3):
        pipe.dispatch_cached(scene, cam, bufs, cache, f)
    w = cache.weights()
    pipe.dispatch_cached(scene, cam, bufs, cache, 3)
    img = bufs.output(api.Output.COLOR)
    assert np.isfinite(img).all() and img[..., :3].mean() > 0
    _same_render(img, _frozen_render(pipe, scene, cam, 40, 40, w, 3)[0])
