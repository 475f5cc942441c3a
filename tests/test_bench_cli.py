"""CPU checks of bench.py's command line: every workload runs on one GPU, and
asking for more exits with status 2 and a message before any GPU work."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.parametrize("workload", ["four_layer_1024", "eight_layer_2048", "allocator", "learned_prior"])
def test_more_than_one_gpu_is_refused(workload):
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--workload", workload, "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""
