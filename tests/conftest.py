import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import py as oracle_py

    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="session")
def apgx():
    import apgar_amd

    apgar_amd.lib()
    return apgar_amd


def gpu_available() -> bool:
    try:
        import apgar_amd

        apgar_amd.device_info()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu(apgx):
    info = apgx.device_info()
    print(f"[gpu] {info.device_name.decode()} {info.gcn_arch_name.decode()} CUs={info.compute_units}",
          flush=True)
    return info
