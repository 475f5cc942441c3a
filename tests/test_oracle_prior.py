"""CPU checks that pin the learned-prior oracle before it judges the library.

* the vectorised convolution equals its loop definition;
* pooling and upsampling are paired with their exact adjoints;
* the whole backward pass equals central finite differences of the loss
  (float64, bf16 rounding off), for parameters drawn from every conv;
* bf16 rounding is round-to-nearest-even on float32 bit patterns;
* board features and targets on a hand-made board.
"""
import numpy as np
import pytest

from oracle import prior as op


def random_params(rng, scale=1.0):
    lay, total = op.param_layout()
    p = np.zeros(total)
    for w_off, b_off, cin, cout in lay:
        p[w_off:b_off] = rng.normal(0, scale * np.sqrt(2.0 / (9 * cin)), b_off - w_off)
        p[b_off:b_off + cout] = rng.normal(0, 0.05, cout)
    return p


def test_conv_matches_definition():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 5, 4, 3))
    w = rng.normal(size=(4, 9, 3))
    b = rng.normal(size=4)
    assert np.allclose(op.conv(x, w, b, False), op.conv_direct(x, w, b), atol=1e-12)


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 8, 6, 3))
    g = rng.normal(size=(2, 4, 3, 3))
    assert np.isclose((op.pool(x) * g).sum(), (x * op.pool_adjoint(g)).sum())
    c = rng.normal(size=(2, 4, 3, 3))
    f = rng.normal(size=(2, 8, 6, 3))
    assert np.isclose((op.up(c) * f).sum(), (c * op.up_adjoint(f)).sum())
    # pooling a constant keeps it; upsampling then pooling is the identity
    assert np.allclose(op.pool(op.up(c)), c)


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even (1).
    vals = np.array([1.0, 1 + 2**-8, 1 + 3 * 2**-8, 1 + 2**-7 + 2**-9, -2.5, 0.0])
    got = op.bf16(vals)
    assert got.tolist() == [1.0, 1.0, 1 + 2**-6, 1 + 2**-7, -2.5, 0.0]


@pytest.mark.parametrize("layers", [1, 3])
def test_backward_matches_finite_differences(layers):
    rng = np.random.default_rng(10 + layers)
    params = random_params(rng)
    x = rng.random(size=(2, 2 * layers, 8, 8))
    t = rng.normal(size=(2, layers, 8, 8))
    loss, g = op.loss_and_gradients(params, x, t, round_ops=False)
    lay, total = op.param_layout()
    # Live parameters only: weights on padded input planes and on unused head
    # outputs have zero gradient by construction.
    picks = []
    for k, (w_off, b_off, cin, cout) in enumerate(lay):
        for _ in range(3):
            co = rng.integers(0, layers if k == 14 else cout)
            ci = rng.integers(0, 2 * layers if k == 0 else cin)
            picks.append(w_off + (co * 9 + rng.integers(0, 9)) * cin + ci)
        picks.append(b_off + rng.integers(0, layers if k == 14 else cout))
    eps = 1e-6
    for i in picks:
        pp, pm = params.copy(), params.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (op.loss_and_gradients(pp, x, t, False)[0] - op.loss_and_gradients(pm, x, t, False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])
    # the padded-plane and unused-output parameters get exactly zero
    w_off, b_off, cin, cout = lay[0]
    w0 = g[w_off:b_off].reshape(cout, 9, cin)
    assert not w0[:, :, 2 * layers:].any()
    w_off, b_off, cin, cout = lay[14]
    assert not g[w_off:b_off].reshape(cout, 9, cin)[layers:].any()
    assert not g[b_off + layers:b_off + cout].any()


def test_loss_is_mean_squared_error():
    rng = np.random.default_rng(3)
    params = random_params(rng)
    x = rng.random(size=(1, 4, 8, 8))
    out, _ = op.forward(params, x, round_ops=False)
    loss, _ = op.loss_and_gradients(params, x, out + 0.5, round_ops=False)
    assert np.isclose(loss, 0.25)


def test_features_and_targets_hand_board():
    L, H, W = 2, 100, 70      # grid 2 x 2 of 64-cell tiles, ragged at the far edges
    cost = np.ones((L, H, W), dtype=np.uint32)
    cost[1, 0:64, 0:32] = 0   # half of tile (0,0) on layer 1 blocked
    pins = np.array([0, 65, 64 * W + 66, H * W + 5], dtype=np.uint32)  # l0:(0,0),(65,0),(66,64); l1:(5,0)
    f = op.features(cost, pins)
    assert f.shape == (4, 2, 2)
    assert f[0, 0, 0] == 1 / 64 and f[0, 0, 1] == 1 / 64 and f[0, 1, 1] == 1 / 64 and f[2, 0, 0] == 1 / 64
    assert f[3, 0, 0] == 0.5 and f[1, 0, 0] == 0.0
    # tile (0,1) covers x 64..69 only: 58 of 64 columns lie beyond the lattice
    assert np.isclose(f[1, 0, 1], 58 / 64)
    assert np.isclose(f[1, 1, 1], 1 - (6 * 36) / 4096)
    hist = np.zeros((L, H, W), dtype=np.uint32)
    hist[0, :64, :64] = 30
    hist[1, 64:, 64:] = 10
    t = op.targets(hist)
    assert np.isclose(t[0, 0, 0], np.log1p(3.0)) and t[0, 1, 1] == 0.0
    assert np.isclose(t[1, 1, 1], np.log1p(1.0))  # mean over the 36 x 6 cells inside the lattice
