"""GPU parity of the candidate-allocator price epoch against its oracle.

Bar: bit-exact planes and summary, and the same outcome and cell for every
bound, on random small planes; at six_layer_1536 size (14.2 M cells) the
planes against a vectorised restatement and the aggregate identity."""
import numpy as np
import pytest

from oracle import prices as op

pytestmark = pytest.mark.gpu
NONE = (1 << 64) - 1


def _say(msg):
    print(msg, flush=True)


def run_both(apgx, epoch, usage, cap, hist, **settings):
    pol = {k: v for k, v in settings.items() if k in op.DEFAULT_POLICY}
    lim = {k: v for k, v in settings.items() if k in op.DEFAULT_LIMITS}
    g, present, h, total = apgx.price_update(epoch, usage, hist, capacity=cap, **settings)
    oc, cell, r = op.price_epoch(epoch, usage.tolist(), cap.tolist(), hist.tolist(), pol, lim)
    assert g["outcome"] == oc, (g, oc, cell)
    assert g["outcome_cell"] == (NONE if cell is None else cell)
    if oc == op.OK:
        assert g["present_factor"] == r["present_factor"] and g["hot_resources"] == r["hot_resources"]
        assert g["price_entries"] == r["price_entries"] and g["aggregate_price"] == r["aggregate_price"]
        assert present.tolist() == r["present"] and h.tolist() == r["history"] and total.tolist() == r["total"]
    else:
        assert np.array_equal(h, hist)  # a failing epoch leaves the planes alone
    return g


@pytest.mark.parametrize("seed", range(6))
def test_random_planes_match_oracle(apgx, gpu, seed):
    _say(f"[case] price epoch seed={seed}")
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 5000))
    usage = rng.integers(0, 4, n).astype(np.uint32)
    cap = rng.integers(0, 3, n).astype(np.uint8)
    hist = np.where(rng.random(n) < 0.3, rng.integers(1, 200, n) * 10, 0).astype(np.uint32)
    epoch = int(rng.integers(0, 6))
    g = run_both(apgx, epoch, usage, cap, hist)
    assert g["outcome"] == op.OK
    over = int((usage > cap).sum())
    entries = int(((hist != 0) | (usage > cap)).sum())
    # bounds at, just below and just above each running quantity
    for lim in (dict(maximum_hot_resources=max(over - 1, 0)), dict(maximum_hot_resources=over),
                dict(maximum_price_entries=max(entries - 1, 0)), dict(maximum_price_entries=entries),
                dict(maximum_price_entries=entries // 2),
                dict(maximum_price_value=int(rng.integers(100, 2000))),
                dict(maximum_aggregate_price=max(g["aggregate_price"] - 1, 0)),
                dict(maximum_aggregate_price=g["aggregate_price"] // 3),
                dict(maximum_aggregate_price=g["aggregate_price"]),
                dict(initial_present_factor=1 << 62),
                dict(present_factor_increment=1 << 63, initial_present_factor=1 << 63)):
        run_both(apgx, epoch, usage, cap, hist, **lim)


def test_reference_known_answers(apgx, gpu):
    usage = np.array([2, 1, 2, 0, 2], dtype=np.uint32)
    g0, p0, h0, t0 = apgx.price_update(0, usage, np.zeros(5, np.uint32))
    assert g0["present_factor"] == 100 and list(p0[[0, 2, 4]]) == [100] * 3
    assert list(h0[[0, 2, 4]]) == [10] * 3 and list(t0[[0, 2, 4]]) == [110] * 3
    g1, p1, h1, t1 = apgx.price_update(1, usage, h0)
    assert list(p1[[0, 2, 4]]) == [200] * 3 and list(h1[[0, 2, 4]]) == [20] * 3 and list(t1[[0, 2, 4]]) == [220] * 3
    gc, pc, hc, tc = apgx.price_update(1, np.ones(5, np.uint32), h0)
    assert list(pc[[0, 2, 4]]) == [0] * 3 and list(hc[[0, 2, 4]]) == [10] * 3 and list(tc[[0, 2, 4]]) == [10] * 3


def test_full_size_plane(apgx, gpu):
    """14.2 M cells (6 x 1536 x 1536): vectorised restatement of the update."""
    _say("[case] price epoch at six_layer_1536 size")
    rng = np.random.default_rng(9)
    n = 6 * 1536 * 1536
    usage = (rng.random(n) < 0.01).astype(np.uint32) + (rng.random(n) < 0.15).astype(np.uint32)
    hist = np.where(rng.random(n) < 0.02, 10 * rng.integers(1, 50, n), 0).astype(np.uint32)
    g, present, h, total = apgx.price_update(3, usage, hist, maximum_price_entries=2_000_000,
                                             maximum_hot_resources=2_000_000)
    assert g["outcome"] == op.OK and g["present_factor"] == 400
    over = np.maximum(usage.astype(np.int64) - 1, 0)
    want_p = 400 * over
    want_h = hist + 10 * over
    assert np.array_equal(present, want_p) and np.array_equal(h, want_h)
    assert np.array_equal(total, want_p + want_h)
    entries = (hist != 0) | (over > 0)
    assert g["price_entries"] == int(entries.sum()) and g["hot_resources"] == int((over > 0).sum())
    assert g["aggregate_price"] == int((want_p + want_h)[entries].sum())
