"""The C-ABI library loads without a GPU and exports every declared entry point."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADERS = [ROOT / "include" / "sparc_hip.h", ROOT / "include" / "sparc_hip3d.h"]
LIB = ROOT / "sparc_flows_amd" / "lib" / "libsparc_hip.so"


def declared_functions():
    text = "\n".join(h.read_text() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sparc(?:3d)?_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("sparc_create", "sparc_destroy", "sparc_run", "sparc_kernels", "sparc_ddx", "sparc_ddz",
              "sparc3d_create", "sparc3d_run", "sparc3d_step", "sparc3d_ddy"):
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


def test_create3d_rejects_bad_config_without_gpu_work():
    if not LIB.exists():
        pytest.skip("libsparc_hip.so not built")
    lib = ctypes.CDLL(str(LIB))
    lib.sparc3d_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.sparc3d_last_error.restype = ctypes.c_char_p
    lib.sparc3d_last_error.argtypes = [ctypes.c_void_p]
    out = ctypes.c_void_p()
    assert lib.sparc3d_create(None, ctypes.byref(out)) == -1
    assert b"null" in lib.sparc3d_last_error(None)
