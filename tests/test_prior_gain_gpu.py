"""What the trained prior gains, asserted (MI355X).

The network the learned-prior study trains (six_layer_512 seeds 0-179, a
fixture written by tests/golden/make_prior_weights.py) seeds the history of a
held-out board of the denser six_layer_512_dense family, where the plain
allocator needs several passes.  Both arms run the same exact-order allocator
with the same legality and overflow accounting; the seeded one must reach zero
overflow in fewer passes (pass counts are deterministic: exact searches, integer
arithmetic, a deterministic network)."""
from pathlib import Path

import numpy as np
import pytest

from tests.lattice_cases import SIX_LAYER_512_DENSE, make_board

pytestmark = pytest.mark.gpu

WEIGHTS = Path(__file__).parent / "golden" / "prior_six_layer_512_weights.npz"


def test_trained_prior_reaches_zero_overflow_in_fewer_passes(apgx, gpu):
    print("[case] six_layer_512_dense seed 201: plain vs learned-prior seed (gain 4)", flush=True)
    w = np.load(WEIGHTS, allow_pickle=False)
    spec = SIX_LAYER_512_DENSE
    cost, nets, pins = make_board(apgx, spec, 201)
    L, H, W = cost.shape
    lat = apgx.Lattice(cost, via_cost=spec["via"], cost_bytes=2)
    net = apgx.Prior(int(w["layers"]), H // 64, W // 64)
    net.set_params(w["params"])
    plain = lat.negotiate(nets, pins, maximum_passes=40, batch_nets=0)
    hist = net.seed_history(lat, pins, gain=4.0)
    learned = lat.negotiate(nets, pins, initial_history=hist, maximum_passes=40, batch_nets=0)
    print(f"[info] plain {plain['passes']} passes {plain['wall_ms']:.0f} ms, learned {learned['passes']} passes "
          f"{learned['wall_ms']:.0f} ms", flush=True)
    assert plain["termination"] == apgx.NEGOTIATION_FEASIBLE
    assert learned["termination"] == apgx.NEGOTIATION_FEASIBLE
    assert learned["passes"] < plain["passes"]
    assert learned["overuse_per_pass"][3] < plain["overuse_per_pass"][3]
    net.close()
    lat.close()
