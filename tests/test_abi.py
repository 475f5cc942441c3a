"""Checks of the C-ABI library: it loads and exports every declared symbol.

The CPU tests make no compute call (there may be no GPU); the one GPU test
reads the device class."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = header.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(apgx_[a-z0-9_]+)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(apgx):
    names = declared_functions()
    assert len(names) >= 15
    handle = ctypes.CDLL(str(apgx.LIB_PATH))
    missing = [n for n in sorted(names) if not hasattr(handle, n)]
    assert not missing, missing


def test_python_binding_covers_the_header(apgx):
    from apgar_amd import replay, v1

    bound = set(apgx.exported_symbols()) | set(v1.exported_symbols()) | set(replay.exported_symbols())
    assert declared_functions() <= bound


def test_last_error_reports_invalid_arguments(apgx):
    lib = apgx.lib()
    assert lib.apgx_abi_version() >> 16 == 1
    rc = lib.apgx_generate_lattice(None, None)
    assert rc == 4
    assert b"invalid" in lib.apgx_last_error()


@pytest.mark.gpu
def test_device_class_names_the_gfx_architecture(apgx, gpu):
    """apgx_device_class: "hip-" + the gfx name of the device (the HIP
    counterpart of the reference's "cuda-cc-<major>.<minor>")."""
    info = apgx.device_info()
    arch = info.gcn_arch_name.decode().split(":")[0]
    assert arch.startswith("gfx")
    assert apgx.device_class() == "hip-" + arch
