"""Neural radiance cache oracle (oracle/oracle_nrc.cpp), pinned on the CPU.

The cache has no counterpart in the reference, so the oracle is pinned by
mathematics: central finite differences of the exact-mode loss for the
gradient of every layer, closed-form forward values of networks with
hand-set weights (encoding order, the constant feature, ReLU placement,
weight layout [in][out]), sin^2 + cos^2 = 1 and dyadic-input values of the
encoding, and the Adam recurrences in closed form.
"""
import numpy as np
import pytest

import oracle

NW = oracle.NRC_NUM_WEIGHTS
W0, WH, W5 = 0, 320 * 64, 320 * 64 + 4 * 64 * 64


def he_weights(seed=0):
    rng = np.random.default_rng(seed)
    w = np.zeros(NW, np.float64)
    a0 = np.sqrt(6.0 / 313)
    w0 = rng.uniform(-a0, a0, (320, 64))
    w0[313:] = 0.0
    w[:WH] = w0.ravel()
    ah = np.sqrt(6.0 / 64)
    w[WH:] = rng.uniform(-ah, ah, NW - WH)
    return w


def test_layout_constants():
    assert NW == 37056 == 320 * 64 + 4 * 64 * 64 + 64 * 3


def test_encoding_identities():
    rng = np.random.default_rng(1)
    for _ in range(20):
        x = rng.random(13).astype(np.float32)
        f = oracle.nrc_encode(x).reshape(-1)
        s, c = f[0:312:2], f[1:312:2]
        np.testing.assert_allclose(s * s + c * c, 1.0, atol=1e-12)
        assert f[312] == 1.0 and (f[313:] == 0).all()
    # dyadic inputs: x = 1/2 -> sin(2^k pi / 2) = 1, 0, 0, ...; cos = 0, -1, 1, 1, ...
    x = np.full(13, 0.5, np.float32)
    f = oracle.nrc_encode(x)
    for s in range(13):
        sin_k = f[s * 24 + 0:s * 24 + 24:2]
        cos_k = f[s * 24 + 1:s * 24 + 24:2]
        np.testing.assert_allclose(sin_k, [1.0] + [0.0] * 11, atol=1e-9)
        np.testing.assert_allclose(cos_k, [0.0, -1.0] + [1.0] * 10, atol=1e-9)
    # x = 1/8 on input 4 only: feature 4*24 + 2k + c = sin/cos(2^k pi / 8)
    x = np.zeros(13, np.float32)
    x[4] = 0.125
    f = oracle.nrc_encode(x)
    k = np.arange(12)
    np.testing.assert_allclose(f[96:120:2], np.sin(2.0 ** k * np.pi / 8), atol=1e-12)
    np.testing.assert_allclose(f[97:120:2], np.cos(2.0 ** k * np.pi / 8), atol=1e-12)


def _chain_weights(row0):
    """Feature row0 -> hidden unit 0 of every layer (weight 1) -> output 0,
    1.5 x it -> output 1; everything else zero."""
    w = np.zeros(NW, np.float64)
    w[W0 + row0 * 64 + 0] = 1.0
    for l in range(4):
        w[WH + l * 4096 + 0 * 64 + 0] = 1.0
    w[W5 + 0 * 3 + 0] = 1.0
    w[W5 + 0 * 3 + 1] = 1.5
    return w


@pytest.mark.parametrize("s,k,c", [(0, 0, 0), (3, 5, 1), (12, 11, 0), (7, 2, 1)])
def test_forward_single_feature_chain(s, k, c):
    rng = np.random.default_rng(s * 100 + k)
    x = rng.random((64, 13)).astype(np.float32)
    y = oracle.nrc_forward(_chain_weights(s * 24 + 2 * k + c), x)
    a = np.ldexp(np.pi, k) * x[:, s].astype(np.float64)
    feat = np.sin(a) if c == 0 else np.cos(a)
    expect = np.maximum(feat, 0.0)  # ReLU after the first layer (and the identity chain)
    np.testing.assert_allclose(y[:, 0], expect, atol=1e-12)
    np.testing.assert_allclose(y[:, 1], 1.5 * expect, atol=1e-12)
    np.testing.assert_allclose(y[:, 2], 0.0, atol=0)


def test_forward_bias_row_and_negative_relu():
    x = np.random.default_rng(3).random((8, 13)).astype(np.float32)
    w = _chain_weights(312)  # the constant feature: y0 = relu(1) = 1
    np.testing.assert_allclose(oracle.nrc_forward(w, x)[:, 0], 1.0, atol=1e-15)
    w[W0 + 312 * 64] = -2.0  # negative pre-activation is cut by the ReLU
    np.testing.assert_allclose(oracle.nrc_forward(w, x)[:, 0], 0.0, atol=0)
    # a negative hidden weight after a positive unit is cut at the next ReLU
    w = _chain_weights(312)
    w[WH + 2 * 4096] = -1.0
    np.testing.assert_allclose(oracle.nrc_forward(w, x)[:, 0], 0.0, atol=0)


def test_forward_output_linear_in_last_layer():
    w = he_weights(4)
    x = np.random.default_rng(5).random((32, 13)).astype(np.float32)
    y = oracle.nrc_forward(w, x)
    w2 = w.copy()
    w2[W5:] *= -3.0
    np.testing.assert_allclose(oracle.nrc_forward(w2, x), -3.0 * y, rtol=1e-12, atol=1e-14)


def _fd_loss(w, x, t, den):
    y = oracle.nrc_forward(w, x)
    return float(np.mean((y - t) ** 2 / den))


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    w = he_weights(7)
    x = rng.random((12, 13)).astype(np.float32)
    t = rng.random((12, 3)).astype(np.float32)
    g, loss = oracle.nrc_grad(w, x, t)
    y0 = oracle.nrc_forward(w, x)
    den = y0 ** 2 + 0.01  # held constant, as the loss definition states
    assert abs(loss - np.mean((y0 - t) ** 2 / den)) < 1e-12
    # weights from every layer, including the bias row and the output layer
    picks = np.concatenate([
        rng.choice(np.arange(0, 312 * 64), 40, replace=False),
        np.arange(312 * 64, 312 * 64 + 64, 9),
        WH + rng.choice(4 * 4096, 60, replace=False),
        W5 + np.arange(0, 192, 5),
    ])
    h = 1e-6
    gmax = np.abs(g).max()
    bad = []
    for i in picks:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (_fd_loss(wp, x, t, den) - _fd_loss(wm, x, t, den)) / (2 * h)
        if abs(fd - g[i]) > 1e-6 * gmax + 1e-4 * abs(g[i]):
            bad.append((int(i), fd, g[i]))
    # a ReLU kink inside +-h can spoil a difference quotient; allow one
    assert len(bad) <= 1, bad[:5]
    assert np.abs(g[313 * 64:320 * 64]).max() == 0.0  # unused rows get no gradient


def test_bf16_mode_tracks_exact_mode():
    rng = np.random.default_rng(11)
    w = he_weights(11)
    x = rng.random((256, 13)).astype(np.float32)
    t = rng.random((256, 3)).astype(np.float32)
    ye = oracle.nrc_forward(w, x, oracle.NRC_EXACT)
    yb = oracle.nrc_forward(w, x, oracle.NRC_BF16)
    scale = np.abs(ye).max()
    assert np.abs(yb - ye).max() < 0.05 * scale
    ge, _ = oracle.nrc_grad(w, x, t, oracle.NRC_EXACT)
    gb, _ = oracle.nrc_grad(w, x, t, oracle.NRC_BF16)
    cos = ge @ gb / (np.linalg.norm(ge) * np.linalg.norm(gb))
    assert cos > 0.99, cos


def test_adam_closed_form():
    rng = np.random.default_rng(2)
    n = 1000
    w = rng.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    g = (rng.standard_normal(n) * 0.1).astype(np.float32)
    g[:10] = 0.0
    w0 = w.copy()
    b1, b2, lr, eps = np.float32(0.9), np.float32(0.99), np.float32(1e-3), np.float32(1e-8)
    p1, p2 = 1.0, 1.0
    for step in range(1, 4):
        p1 *= float(b1)
        p2 *= float(b2)
        oracle.nrc_adam(w, m, v, g, lr, b1, b2, eps, np.float32(1 - p1), np.float32(1 - p2))
        # constant gradient: bias-corrected moments equal g and g^2, so each
        # step moves by lr * g / (|g| + eps)
        gd = g[10:].astype(np.float64)
        expect = -step * float(lr) * gd / (np.abs(gd) + float(eps))
        np.testing.assert_allclose(w[10:] - w0[10:], expect, rtol=1e-4, atol=1e-9)
        np.testing.assert_array_equal(w[:10], w0[:10])
    np.testing.assert_allclose(m, (1 - 0.9 ** 3) * g, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(v, (1 - 0.99 ** 3) * g * g, rtol=1e-5, atol=1e-15)
