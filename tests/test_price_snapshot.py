"""Price snapshots on cell resources (apgx_price_snapshot_*).

CPU: the policy and snapshot identities against the oracle's restatement of
the reference's field order, and every check of ValidatePriorSnapshot on a
snapshot built from the reference's known-answer epoch.  GPU: device capture
(ballot / prefix-scan compaction) against the oracle's entry map on random and
14.2 M-cell planes; capture -> validate -> restore -> next epoch equals the
epoch run on the original planes."""
import numpy as np
import pytest

from oracle import prices as op

INF64 = (1 << 64) - 1


def _assoc(apgx, w=4, h=3, layers=2):
    return apgx.PriceAssociations(11, 12, 1, w, h, layers)


def _assoc_dict(a):
    return {k: getattr(a, k) for k, _ in a._fields_}


def _entries(apgx, rows):
    e = np.zeros(len(rows), apgx.PRICE_ENTRY)
    for i, r in enumerate(rows):
        e[i] = r
    return e


def known_answer_snapshot(apgx, epoch=0, prior=0):
    """The reference's known answers (negotiated_regeneration_plan_test.cc:
    815-872: epoch 0 prices 100 / 10 / 110, epoch 1 prices 200 / 20 / 220) on
    cell 5 of a 4 x 3 x 2 lattice overused by one in every epoch; from epoch 1
    on, cell 17 also carries history 30 (three units of earlier overuse)."""
    a = _assoc(apgx)
    f = 100 + 100 * epoch
    rows = [(5, f, 10 * (epoch + 1), f + 10 * (epoch + 1))]
    if epoch:
        rows.append((17, 0, 30, 30))
    snap = apgx.PriceSnapshot()
    snap.associations = a
    snap.capacity_units = 1
    snap.policy_identity = op.policy_identity()
    snap.prior_snapshot_identity = prior
    snap.epoch_index = epoch
    snap.present_factor = 100 + 100 * epoch
    snap.price_count = len(rows)
    e = _entries(apgx, rows)
    snap.snapshot_identity = apgx.price_snapshot_identity(snap, e)
    return snap, e


@pytest.mark.parametrize("policy", [{}, dict(initial_present_factor=0), dict(present_factor_increment=7,
                                                                              historical_price_increment=3)])
def test_policy_identity(apgx, policy):
    assert apgx.price_policy_identity(**policy) == op.policy_identity(policy)


def test_snapshot_identity_restatement(apgx):
    snap, e = known_answer_snapshot(apgx, epoch=2, prior=99)
    want = op.snapshot_identity(_assoc_dict(snap.associations), 1, snap.policy_identity, 99, 2, 300,
                                [tuple(int(v) for v in r) for r in e])
    assert snap.snapshot_identity == want != 0


def test_known_answer_snapshot_validates(apgx):
    snap, e = known_answer_snapshot(apgx)
    assert apgx.price_snapshot_validate(snap, e, _assoc(apgx), 1, maximum_overuse_per_epoch=4) == (True, "", INF64)


def _mutate(apgx, name):
    snap, e = known_answer_snapshot(apgx, epoch=1, prior=5)
    kw = dict(maximum_overuse_per_epoch=4)
    expected = _assoc(apgx)
    resign = True
    if name == "association":
        expected.board_content_hash += 1
    elif name == "capacity":
        snap.capacity_units = 2
    elif name == "policy":
        kw["historical_price_increment"] = 20
    elif name == "chain":
        snap.prior_snapshot_identity = 0
    elif name == "present factor":
        snap.present_factor = 100
    elif name == "price count":
        kw["maximum_price_entries"] = 1
    elif name == "order":
        e = e[::-1].copy()
    elif name == "zero entry":
        e[1] = (17, 0, 0, 0)
    elif name == "total":
        e[0]["total"] = 221
    elif name == "history schedule":
        e[0]["history"], e[0]["total"] = 25, 225
    elif name == "present schedule":
        e[0]["present"], e[0]["total"] = 300, 320
    elif name == "current overuse":
        e[0]["present"], e[0]["history"], e[0]["total"] = 1000, 50, 1050
    elif name == "history contribution":
        e[0]["present"], e[0]["history"], e[0]["total"] = 400, 10, 410
    elif name == "cumulative overuse":
        e[1]["history"], e[1]["total"] = 90, 90  # 9 units of prior overuse > 1 epoch x 4
    elif name == "value bound":
        kw["maximum_price_value"] = 100
    elif name == "aggregate bound":
        kw["maximum_aggregate_price"] = 240
    elif name == "identity":
        resign = False
        snap.snapshot_identity ^= 1
    elif name == "epoch bound":
        kw["maximum_epoch_index"] = 1
    if resign:
        snap.snapshot_identity = apgx.price_snapshot_identity(snap, e)
    return snap, e, expected, kw


CASES = {
    "association": ("allocator.negotiated_plan.prior_association.v1", INF64),
    "capacity": ("allocator.negotiated_plan.prior_association.v1", INF64),
    "policy": ("allocator.negotiated_plan.prior_policy.v1", INF64),
    "chain": ("allocator.negotiated_plan.prior_chain.v1", INF64),
    "present factor": ("allocator.negotiated_plan.prior_present_factor.v1", INF64),
    "price count": ("allocator.negotiated_plan.prior_price_count.v1", INF64),
    "order": ("allocator.negotiated_plan.prior_price_order.v1", 1),
    "zero entry": ("allocator.negotiated_plan.prior_price_order.v1", 1),
    "total": ("allocator.negotiated_plan.prior_price_total.v1", 0),
    "history schedule": ("allocator.negotiated_plan.prior_history_schedule.v1", 0),
    "present schedule": ("allocator.negotiated_plan.prior_present_schedule.v1", 0),
    "current overuse": ("allocator.negotiated_plan.prior_current_overuse_bound.v1", 0),
    "history contribution": ("allocator.negotiated_plan.prior_history_contribution.v1", 0),
    "cumulative overuse": ("allocator.negotiated_plan.prior_cumulative_overuse_bound.v1", 1),
    "value bound": ("allocator.negotiated_plan.prior_price_bound.v1", 0),
    "aggregate bound": ("allocator.negotiated_plan.prior_aggregate_price_bound.v1", 1),
    "identity": ("allocator.negotiated_plan.prior_identity.v1", INF64),
    "epoch bound": ("allocator.negotiated_plan.epoch_bound.v1", INF64),
}


@pytest.mark.parametrize("name", list(CASES))
def test_each_prior_check(apgx, name):
    snap, e, expected, kw = _mutate(apgx, name)
    ok, inv, idx = apgx.price_snapshot_validate(snap, e, expected, 1, **kw)
    assert (ok, inv, idx) == (False, *CASES[name])


def test_unmutated_epoch_one_snapshot_validates(apgx):
    snap, e, expected, kw = _mutate(apgx, "none")
    assert apgx.price_snapshot_validate(snap, e, expected, 1, **kw)[0]


@pytest.mark.gpu
@pytest.mark.parametrize("shape, density", [((2, 37, 53), 0.05), ((3, 64, 64), 0.6), ((6, 1536, 1536), 0.03)])
def test_capture_matches_oracle(apgx, gpu, shape, density):
    print(f"[case] price snapshot capture {shape} {density}", flush=True)
    rng = np.random.default_rng(sum(shape))
    L, H, W = shape
    n = L * H * W
    present = np.where(rng.random(n) < density, rng.integers(1, 50, n) * 100, 0).astype(np.uint32)
    history = np.where(rng.random(n) < density, rng.integers(1, 50, n) * 10, 0).astype(np.uint32)
    a = apgx.PriceAssociations(3, 4, 1, W, H, L)
    snap, e = apgx.price_snapshot_capture(a, 1, 2, 77, present, history)
    nz = np.flatnonzero((present | history) != 0)
    assert snap.price_count == nz.size
    assert np.array_equal(e["cell"], nz)
    assert np.array_equal(e["present"], present[nz]) and np.array_equal(e["history"], history[nz])
    assert np.array_equal(e["total"], present[nz].astype(np.uint64) + history[nz])
    assert (snap.epoch_index, snap.present_factor, snap.prior_snapshot_identity) == (2, 300, 77)
    if n <= 20000:
        want = op.snapshot_identity(_assoc_dict(a), 1, op.policy_identity(), 77, 2, 300,
                                    op.snapshot_entries(present, history))
        assert snap.snapshot_identity == want
    else:
        assert snap.snapshot_identity == apgx.price_snapshot_identity(snap, e)


@pytest.mark.gpu
def test_snapshot_chain_resumes_the_next_epoch(apgx, gpu):
    """epoch 0 on random usage -> capture -> validate -> restore -> epoch 1 from
    the restored history equals epoch 1 run on the original planes."""
    print("[case] price snapshot chain", flush=True)
    rng = np.random.default_rng(21)
    L, H, W = 2, 96, 80
    n = L * H * W
    usage = rng.integers(0, 3, n).astype(np.uint32)
    s0, present0, history0, _ = apgx.price_update(0, usage, np.zeros(n, np.uint32))
    assert s0["outcome"] == 0
    a = apgx.PriceAssociations(8, 9, 1, W, H, L)
    snap, e = apgx.price_snapshot_capture(a, 1, 0, 0, present0, history0)
    ok, inv, _ = apgx.price_snapshot_validate(snap, e, a, 1, maximum_overuse_per_epoch=1)
    assert ok, inv
    p_back, h_back = apgx.price_snapshot_restore(snap, e)
    assert np.array_equal(p_back, present0) and np.array_equal(h_back, history0)
    usage1 = rng.integers(0, 3, n).astype(np.uint32)
    direct = apgx.price_update(1, usage1, history0)
    resumed = apgx.price_update(1, usage1, h_back)
    for x, y in zip(direct[1:], resumed[1:]):
        assert np.array_equal(x, y)
    snap1, e1 = apgx.price_snapshot_capture(a, 1, 1, snap.snapshot_identity, resumed[1], resumed[2])
    ok, inv, _ = apgx.price_snapshot_validate(snap1, e1, a, 1, maximum_overuse_per_epoch=1)
    assert ok, inv
