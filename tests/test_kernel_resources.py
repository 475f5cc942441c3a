"""CPU check of the hot kernels' register allocation: the register-resident
tile kernels must not spill (a spill puts scratch-memory traffic inside every
relaxation sweep).  Compiles lattice.hip to gfx950 assembly with the library's
own device flags and reads the kernel metadata."""
import re
import subprocess
import tempfile
from pathlib import Path

from apgar_amd import build as apgx_build


def test_register_tile_kernels_do_not_spill():
    src = apgx_build.CSRC / "lattice.hip"
    with tempfile.TemporaryDirectory() as d:
        out = Path(d) / "lattice.s"
        cmd = [apgx_build.HIPCC, "-x", "hip", *apgx_build.DEVICE, *apgx_build.COMMON, "--cuda-device-only",
               "-S", str(src), "-o", str(out)]
        subprocess.run(cmd, check=True, capture_output=True)
        text = out.read_text()
    meta = re.findall(r"\.name:\s+(\S+)\n(?:.*\n)*?\s+\.vgpr_spill_count:\s+(\d+)", text)
    reg = [(n, int(c)) for n, c in meta if "RegRoundKernel" in n]
    assert len(reg) == 24  # 1..8 layers: plain, via-pays-cell, via-pays-cell with wide costs
    assert all(c == 0 for _, c in reg), reg
