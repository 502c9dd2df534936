"""The C-ABI library loads without a GPU and exports every declared entry point."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "sparc_hip.h"
LIB = ROOT / "sparc_flows_amd" / "lib" / "libsparc_hip.so"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sparc_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("sparc_create", "sparc_destroy", "sparc_run", "sparc_kernels", "sparc_ddx", "sparc_ddz"):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        pytest.skip("libsparc_hip.so not built")
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_create_rejects_bad_config_without_gpu_work():
    if not LIB.exists():
        pytest.skip("libsparc_hip.so not built")
    lib = ctypes.CDLL(str(LIB))
    lib.sparc_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.sparc_last_error.restype = ctypes.c_char_p
    lib.sparc_last_error.argtypes = [ctypes.c_void_p]
    out = ctypes.c_void_p()
    assert lib.sparc_create(None, ctypes.byref(out)) == -1
    assert b"null" in lib.sparc_last_error(None)
