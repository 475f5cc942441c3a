"""Writes tests/golden/prior_six_layer_512_weights.npz: the learned prior's
parameters as the study trains them (needs an MI355X; run under the GPU runner
and copy the file here).

The training is the learned-prior study's, step for step: plain exact-order
allocator runs on six_layer_512 seeds 0-199 (or a dataset .npz of them written
by `bench.py --workload learned_prior --collect-only --dataset ...`), then 300
epochs of Adam at batch 32 on seeds 0-179.  Training is deterministic, so the
file is a fixture of this library's own output, used by
tests/test_prior_gain_gpu.py to assert what the trained prior gains.

    python tests/golden/make_prior_weights.py OUT.npz [DATASET.npz]
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from apgar_amd import prior_study as ps  # noqa: E402
from tests.lattice_cases import SIX_LAYER_512  # noqa: E402


def main():
    out = Path(sys.argv[1])
    data = Path(sys.argv[2]) if len(sys.argv) > 2 else None
    if data is not None and data.exists():
        d = np.load(data, allow_pickle=False)
        x, y = d["x"], d["y"]
    else:
        x, y, _ = ps.collect(SIX_LAYER_512, list(range(200)), log=lambda m: print(m, flush=True))
    net, rep = ps.train(x[:180], y[:180], SIX_LAYER_512["layers"], epochs=300, batch=32)
    print(f"trained: {rep}", flush=True)
    np.savez_compressed(out, params=net.get_params(), layers=np.uint32(SIX_LAYER_512["layers"]),
                        training_boards=np.uint32(180), epochs=np.uint32(300), batch=np.uint32(32))
    net.close()


if __name__ == "__main__":
    main()
