"""GPU parity of the learned history-cost prior against its CPU oracle.

Bar (written per test): features and targets to float32 rounding; forward
outputs and every parameter gradient against the oracle with the same bf16
operand rounding, to fp32 accumulation-order error; training is deterministic
(graph replay equals direct launches bit for bit) and lowers the loss; the
decoded history plane follows the decode rule exactly.
"""
import numpy as np
import pytest

from oracle import prior as op

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def random_params(rng, layers):
    lay, total = op.param_layout()
    p = np.zeros(total, dtype=np.float32)
    for k, (w_off, b_off, cin, cout) in enumerate(lay):
        w = rng.normal(0, np.sqrt(2.0 / (9 * cin)), (cout, 9, cin))
        if k == 0:
            w[:, :, 2 * layers:] = 0
        if k == 14:
            w[layers:] = 0
        p[w_off:b_off] = w.reshape(-1)
        p[b_off:b_off + cout] = rng.normal(0, 0.05, cout)
    return p


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("layers,grid,n", [(2, 8, 3), (6, 24, 2)])
def test_forward_and_gradients_match_oracle(apgx, gpu, layers, grid, n):
    _say(f"[case] prior gradients layers={layers} grid={grid} n={n}")
    rng = np.random.default_rng(layers * 100 + grid)
    net = apgx.Prior(layers, grid, grid, seed=5)
    params = random_params(rng, layers)
    net.set_params(params)
    x = rng.random((n, 2 * layers, grid, grid)).astype(np.float32)
    t = rng.normal(0, 1, (n, layers, grid, grid)).astype(np.float32)
    out = net.forward(x)
    ref, _ = op.forward(params.astype(np.float64), x, round_ops=True)
    # fp32 accumulation of bf16 products vs float64: relative to the output scale
    assert np.abs(out - ref).max() <= 2e-3 * np.abs(ref).max() + 1e-5
    loss, g = net.gradients(x, t)
    rloss, rg = op.loss_and_gradients(params.astype(np.float64), x, t, round_ops=True)
    assert abs(loss - rloss) <= 1e-4 * rloss
    lay, _ = op.param_layout()
    for k, (w_off, b_off, cin, cout) in enumerate(lay):
        assert rel(g[w_off:b_off], rg[w_off:b_off]) < 5e-3, (k, rel(g[w_off:b_off], rg[w_off:b_off]))
        assert rel(g[b_off:b_off + cout], rg[b_off:b_off + cout]) < 5e-3, k
    net.close()


def test_training_is_deterministic_and_learns(apgx, gpu):
    """A smooth synthetic target (a blurred combination of the input planes):
    the loss must drop, and graph replay must equal direct launches bit for bit."""
    _say("[case] prior training determinism")
    rng = np.random.default_rng(7)
    L, G, S = 2, 16, 48
    x = rng.random((S, 2 * L, G, G)).astype(np.float32)
    k = np.array([0.25, 0.5, 0.25])
    blur = lambda a: np.apply_along_axis(lambda r: np.convolve(r, k, "same"), -1,
                                         np.apply_along_axis(lambda r: np.convolve(r, k, "same"), -2, a))
    t = np.stack([blur(x[:, 0] + 0.5 * x[:, 1]), blur(x[:, 2] - x[:, 3])], axis=1).astype(np.float32)
    results = []
    for graph in (1, 0):
        net = apgx.Prior(L, G, G, seed=11)
        rep = net.train(x, t, epochs=30, batch=16, learning_rate=2e-3, use_graph=graph, seed=3)
        assert rep["graph"] == graph
        results.append((rep, net.get_params()))
        net.close()
    (r1, p1), (r0, p0) = results
    assert np.array_equal(p1, p0)
    assert r1["last_epoch_loss"] == r0["last_epoch_loss"]
    assert r1["last_epoch_loss"] < 0.25 * r1["first_epoch_loss"], r1


def test_features_targets_and_decode(apgx, gpu):
    _say("[case] prior features / targets / decode")
    from tests.lattice_cases import SIX_LAYER_1536, make_board

    cost, nets, pins = make_board(apgx, SIX_LAYER_1536, 900, width=512, height=512, nets=120, bga_fields=1,
                                  scatter_radius=160)
    L, H, W = cost.shape
    lat = apgx.Lattice(cost, via_cost=20, cost_bytes=2)
    f = apgx.prior_features(lat, pins)
    assert np.abs(f - op.features(cost, pins)).max() < 1e-6
    rng = np.random.default_rng(2)
    hist = (rng.integers(0, 50, size=cost.shape) * 10).astype(np.uint32)
    t = apgx.prior_targets(hist)
    assert np.abs(t - op.targets(hist)).max() < 1e-6
    net = apgx.Prior(L, H // 64, W // 64, seed=9)
    params = random_params(rng, L)
    net.set_params(params)
    plane = net.seed_history(lat, pins, gain=2.0, max_value=400)
    out = net.forward(f[None])[0]
    tiles = np.minimum(np.rint(2.0 * 10.0 * np.expm1(np.maximum(out, 0.0))), 400)
    want = np.repeat(np.repeat(tiles, 64, axis=1), 64, axis=2).astype(np.uint32)
    # expm1f / rounding of float32 values may differ by one unit at most
    assert np.abs(plane.astype(np.int64) - want.astype(np.int64)).max() <= 1
    assert (plane == want).mean() > 0.99
    lat.close()
    net.close()


def test_trained_prior_gain_on_held_out_boards(apgx, gpu):
    """End to end at test scale: plain allocator runs on 16 training boards give
    the targets, the network trains on the GPU, and on 4 held-out boards it must
    predict the allocator's history better than zero history does.  Both arms of
    the allocator comparison run the same exact-order policy and legality checks;
    their passes and times are printed (the full study is bench.py's
    learned_prior workload)."""
    from apgar_amd import prior_study as ps
    from tests.lattice_cases import SIX_LAYER_512

    spec = dict(SIX_LAYER_512, nets=100)
    _say("[case] learned prior end to end: 16 training boards, 4 held out")
    x, y, runs = ps.collect(spec, range(16), log=_say)
    assert all(r["routed"] > 0 for r in runs)
    net, rep = ps.train(x, y, spec["layers"], epochs=150, batch=16)
    assert rep["last_epoch_loss"] < rep["first_epoch_loss"]
    xh, yh, _ = ps.collect(spec, range(200, 204), log=_say)
    mse, zero_mse = ps.prediction_error(net, xh, yh)
    _say(f"[info] held-out MSE {mse:.4f} vs zero history {zero_mse:.4f}")
    assert mse < 0.5 * zero_mse
    res = [ps.seeded_run(net, spec, s, 1.0) for s in range(200, 204)]
    for r in res:
        for arm in ("plain", "learned"):
            # zero overflow is reported only with every net routed
            if r[arm]["time_to_zero_overflow_ms"] is not None:
                assert r[arm]["overuse"] == 0 and r[arm]["routed"] == spec["nets"]
    _say(f"[info] held-out comparison: {ps.compare(res)}")
    net.close()
