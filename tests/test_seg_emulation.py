"""The column pass's segmented z-derivative, run on the CPU from the kernel's
own source: tests/seg_emul.cpp drives seg_rhs / seg_solve / seg_correct
(sparc_flows_amd/csrc/sparc_seg.h) with the library's tables
(sparc_segtab.cpp) for every segment of one column, as the lanes of
k_colsplit do, and the result is compared with the oracle's dbyd2 (itself
bit-identical to the reference's dbyd2.f + mtridag.f) for every segment
length the autotuner can choose, ibc 1, 2 and 3 (the three the column pass
uses), with the kept first / last rows of ibc 2 / 3.

The segmented solve differs from the Thomas sweep by rounding only: the
bar is 1e-12 relative (max norm) on smooth columns and 1e-10 on white noise,
whose boundary closure amplifies rounding."""
import ctypes
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "sparc_flows_amd" / "csrc"
SIZES = (4, 6, 8, 10, 12, 16, 19, 20, 24)


def usable(nz, ms):
    p = -(-nz // ms)
    return p <= 64 and nz - (p - 1) * ms >= 3


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("seg") / "libsegemul.so"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wno-unknown-pragmas", f"-I{CSRC}",
           str(ROOT / "tests" / "seg_emul.cpp"), str(CSRC / "sparc_segtab.cpp"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    L = ctypes.CDLL(str(out))
    dp = ctypes.POINTER(ctypes.c_double)
    L.seg_ddz_column.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double,
                                 dp, ctypes.POINTER(ctypes.c_int)]
    return L


def seg_ddz(L, B, ms, ibc, pre0, preN):
    B = np.ascontiguousarray(B, dtype=np.float64)
    out = np.zeros_like(B)
    hw = ctypes.c_int()
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.seg_ddz_column(B.size, ms, ibc, B.ctypes.data_as(dp), pre0, preN, out.ctypes.data_as(dp),
                          ctypes.byref(hw))
    assert rc == 0
    return out, hw.value


def reference(B, ibc, pre0, preN):
    A = np.zeros((B.size, 1))
    A[0, 0], A[-1, 0] = pre0, preN
    return oracle.dbyd2(B[:, None], ibc, A)[:, 0]


@pytest.mark.parametrize("nz", [16, 23, 40, 60, 97, 150, 300, 301, 512])
def test_segmented_derivative_matches_oracle(emul, nz):
    rng = np.random.default_rng(nz)
    z = np.linspace(0.0, 1.0, nz)
    smooth = np.sin(2.3 * np.pi * z + 0.4) + 0.3 * z ** 3
    noise = rng.standard_normal(nz)
    worst = {}
    for ms in [m for m in SIZES if usable(nz, m)]:
        for ibc in (1, 2, 3):
            pre0, preN = (0.0, 0.0) if ibc == 1 else ((0.7, 0.0) if ibc == 2 else (0.0, -1.3))
            for name, B, tol in (("smooth", smooth, 1e-12), ("noise", noise, 1e-10)):
                got, hw = seg_ddz(emul, B, ms, ibc, pre0, preN)
                ref = reference(B, ibc, pre0, preN)
                err = np.abs(got - ref).max() / np.abs(ref).max()
                worst[(ms, ibc, name)] = err
                assert err < tol, (nz, ms, ibc, name, err, hw)
    print(f"nz={nz}: worst {max(worst.values()):.2e} over {len(worst)} cases")


def test_each_boundary_row_is_exercised(emul):
    """The first two and last two rows differ from a derivative made with the
    interior formula alone, so a wrong one-sided row cannot pass unnoticed."""
    nz, ms = 150, 10
    B = np.cos(np.linspace(0.0, 3.0, nz)) ** 3
    for ibc in (1, 2, 3):
        got, _ = seg_ddz(emul, B, ms, ibc, 0.0, 0.0)
        ref = reference(B, ibc, 0.0, 0.0)
        assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    # a column whose only nonzero values sit in the rows the one-sided
    # formulas alone read: B(0..4) and B(nz-5..nz-1)
    for rows in (range(0, 5), range(nz - 5, nz)):
        for r in rows:
            B = np.zeros(nz)
            B[r] = 1.0
            for ibc in (1, 2, 3):
                got, _ = seg_ddz(emul, B, ms, ibc, 0.0, 0.0)
                ref = reference(B, ibc, 0.0, 0.0)
                assert np.abs(got - ref).max() < 1e-12 * max(np.abs(ref).max(), 1.0), (r, ibc)


def test_column_pass_fits_lds_up_to_nz_1024(emul):
    """Some segment length the autotuner can choose fits the column pass's
    LDS (colsplit_lds_bytes, TC = 1) in gfx950's 160 KB for every nz up to
    1024: the segment tables are held for three segments only, and the
    profiles, p, the RHS v_z parts and the source profile per row."""
    kzs, kzf, kpl, nzpml = 15, 6, 7, 40
    worst = 0
    for nz in list(range(16, 1025, 7)) + [1023, 1024]:
        best = None
        for ms in [m for m in SIZES if usable(nz, m)]:
            _, hw = seg_ddz(emul, np.ones(nz), ms, 1, 0.0, 0.0)
            P = -(-nz // ms)
            nzp = P * ms
            staged = 3 * 2 * P * (4 * hw + 4) + kzs * 3 * ms + kzf * nzp + 12 + kpl * nzpml + nzp
            lds = 8 * (staged + 4 * (2 * P + 4 * hw + 2) + 3 * nzp)
            best = lds if best is None else min(best, lds)
        if best is None:
            continue                         # one of the segmentation gaps (test_colseg_bounds)
        assert best <= 160 * 1024, (nz, best)
        worst = max(worst, best)
    print(f"largest minimum LDS {worst / 1024:.1f} KB")
