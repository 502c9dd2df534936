"""sparc_ddz_noalloc (the allocation-free z-derivative) through the C ABI:
equal to sparc_ddz / sparc_ddz_kern on the same inputs, on two calls in a row
(the second smaller), and sizes beyond its scratch are rejected."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from sparc_flows_amd import background as B  # noqa: E402

LIB = Path(__file__).resolve().parents[1] / "sparc_flows_amd" / "lib" / "libsparc_hip.so"
SCRATCH_DOUBLES = 1 << 20


class Config(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("nz", ctypes.c_int), ("npmlbot", ctypes.c_int),
                ("npmltop", ctypes.c_int), ("npmlhor", ctypes.c_int), ("nsrc", ctypes.c_int),
                ("adjoint", ctypes.c_int), ("xlength", ctypes.c_double), ("timestep", ctypes.c_double),
                ("cadforcing", ctypes.c_double), ("outputcad", ctypes.c_double),
                ("obsheight", ctypes.c_double), ("excitdep", ctypes.c_double),
                ("nupeak", ctypes.c_double), ("nuwidth", ctypes.c_double),
                ("model", ctypes.c_void_p), ("c2d", ctypes.c_void_p), ("psi", ctypes.c_void_p)]


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(str(LIB))
    L.sparc_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_void_p)]
    L.sparc_destroy.argtypes = [ctypes.c_void_p]
    L.sparc_last_error.argtypes = [ctypes.c_void_p]
    L.sparc_last_error.restype = ctypes.c_char_p
    for name in ("sparc_ddz", "sparc_ddz_kern"):
        getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_int]
    L.sparc_ddz_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int]
    L.sparc_query_int.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
    return L


def _handle(L, nx, nz):
    bg = np.ascontiguousarray(B.polytrope(nz))
    cfg = Config(nx, nz, 8, 8, 5, 1, 0, 200e8, B.stable_timestep(bg), 30.0, 30.0, 2e7, -1.5e7,
                 0.0017, 0.0055, bg.ctypes.data, None, None)
    h = ctypes.c_void_p()
    torch.cuda.synchronize()
    rc = L.sparc_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == 0, L.sparc_last_error(None)
    return h


@pytest.mark.parametrize("filt", [0, 1])
def test_noalloc_equals_allocating_entry_point(lib, filt):
    nx, nz = 128, 150
    h = _handle(lib, nx, nz)
    try:
        ref_fn = lib.sparc_ddz_kern if filt else lib.sparc_ddz
        rng = np.random.default_rng(7 + filt)
        for nbatch in (5, 2):                      # second call smaller than the first
            for ibc in (0, 1, 2, 3):
                f = torch.as_tensor(rng.standard_normal((nbatch, nz, nx)), device="cuda")
                pre = torch.as_tensor(rng.standard_normal((nbatch, nz, nx)), device="cuda")
                a = pre.clone()
                b = pre.clone()
                torch.cuda.synchronize()
                assert ref_fn(h, f.data_ptr(), a.data_ptr(), nbatch, ibc) == 0
                rc = lib.sparc_ddz_noalloc(h, f.data_ptr(), b.data_ptr(), nbatch, ibc, filt)
                assert rc == 0, lib.sparc_last_error(h)
                torch.cuda.synchronize()
                print("noalloc filt", filt, "nbatch", nbatch, "ibc", ibc, "max |diff|",
                      float((a - b).abs().max()))
                assert torch.equal(a, b), (nbatch, ibc)
    finally:
        lib.sparc_destroy(h)


def test_noalloc_rejects_sizes_beyond_its_scratch(lib):
    nx, nz = 256, 300
    h = _handle(lib, nx, nz)
    try:
        nbatch = SCRATCH_DOUBLES // (nx * nz) + 1
        f = torch.zeros((1, nz, nx), dtype=torch.float64, device="cuda")
        g = torch.zeros_like(f)
        # the size check runs before any device work, so no buffer of that size is needed
        rc = lib.sparc_ddz_noalloc(h, f.data_ptr(), g.data_ptr(), nbatch, 1, 0)
        assert rc == -1
        assert b"exceeds the library scratch" in lib.sparc_last_error(h)
        ok = lib.sparc_ddz_noalloc(h, f.data_ptr(), g.data_ptr(), 1, 1, 0)
        torch.cuda.synchronize()
        assert ok == 0
    finally:
        lib.sparc_destroy(h)
