"""File formats of the drop-in (FITS images, direct-access records, params.i)."""
import numpy as np

from sparc_flows_amd.fitsio import read_fits, write_direct_records, write_fits


def test_fits_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((7, 1, 16))
    write_fits(tmp_path / "a.fits", a, time=42)
    raw = (tmp_path / "a.fits").read_bytes()
    assert len(raw) % 2880 == 0
    hdr = raw[:2880].decode("ascii")
    assert "BITPIX  =                  -64" in hdr
    assert "NAXIS1  =                   16" in hdr and "NAXIS3  =                    7" in hdr
    assert "TIME    =                   42" in hdr
    b = read_fits(tmp_path / "a.fits")
    assert b.shape == a.shape and np.array_equal(a, b)


def test_fits_reads_float32_and_scaled_int(tmp_path):
    # a BITPIX=16 image with BSCALE/BZERO, written by hand
    def card(s):
        return s.ljust(80).encode()
    hdr = b"".join(card(c) for c in ["SIMPLE  =                    T", "BITPIX  =                   16",
                                      "NAXIS   =                    2", "NAXIS1  =                    3",
                                      "NAXIS2  =                    2", "BSCALE  =                  0.5",
                                      "BZERO   =                 10.0", "END"])
    hdr += b" " * ((-len(hdr)) % 2880)
    data = np.array([[1, 2, 3], [4, 5, -6]], dtype=">i2").tobytes()
    (tmp_path / "b.fits").write_bytes(hdr + data + b"\0" * ((-len(data)) % 2880))
    b = read_fits(tmp_path / "b.fits")
    assert np.array_equal(b, np.array([[1, 2, 3], [4, 5, -6]]) * 0.5 + 10.0)


def test_direct_access_records(tmp_path):
    rows = np.arange(12, dtype=np.float64).reshape(3, 4)
    write_direct_records(tmp_path / "vz_cc.bin", rows)
    back = np.fromfile(tmp_path / "vz_cc.bin", dtype="<f8").reshape(3, 4)
    assert np.array_equal(rows, back)
