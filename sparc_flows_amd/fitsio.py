"""Minimal FITS image I/O for the reference's file formats.

The reference writes every array with CFITSIO as a primary image of BITPIX -64
and axes (nx, ny=1, n) plus a TIME keyword (writefits_3D,
all_modules.f90:296-378), and reads images with ftgpvd, which converts any
BITPIX to double (readfits, all_modules.f90:99-171).  Python callers
(grad.py, misfit.py) read and write the same files with numpy arrays of shape
(n, 1, nx).  astropy is not required.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

BLOCK = 2880
_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


def _card(key: str, value, comment: str = "") -> bytes:
    if isinstance(value, bool):
        v = "T" if value else "F"
        s = f"{key:<8}= {v:>20}"
    elif isinstance(value, (int, np.integer)):
        s = f"{key:<8}= {int(value):>20}"
    elif isinstance(value, float):
        s = f"{key:<8}= {value:>20.13E}"
    else:
        s = f"{key:<8}= '{value}'"
    if comment:
        s += f" / {comment}"
    return s[:80].ljust(80).encode("ascii")


def write_fits(path, data: np.ndarray, time: int | None = None, overwrite: bool = True) -> None:
    """Write data (numpy shape (n3, n2, n1), FITS axes (n1, n2, n3)) as BITPIX -64."""
    path = Path(path)
    if path.exists() and not overwrite:
        raise FileExistsError(path)
    a = np.ascontiguousarray(data, dtype=">f8")
    cards = [_card("SIMPLE", True, "conforms to FITS standard"), _card("BITPIX", -64),
             _card("NAXIS", a.ndim)]
    for i, n in enumerate(a.shape[::-1]):
        cards.append(_card(f"NAXIS{i + 1}", int(n)))
    cards.append(_card("EXTEND", False))
    if time is not None:
        cards.append(_card("TIME", int(time), "Also, init = time+1"))
    cards.append(b"END".ljust(80))
    hdr = b"".join(cards)
    hdr += b" " * ((-len(hdr)) % BLOCK)
    body = a.tobytes()
    body += b"\0" * ((-len(body)) % BLOCK)
    tmp = path.with_suffix(path.suffix + ".tmp")
    tmp.write_bytes(hdr + body)
    tmp.replace(path)


def read_fits(path) -> np.ndarray:
    """Read the primary image as float64, numpy axis order (reversed FITS axes)."""
    raw = Path(path).read_bytes()
    hdr = {}
    pos = 0
    while True:
        if pos + 80 > len(raw):
            raise ValueError(f"{path}: no END card")
        card = raw[pos:pos + 80].decode("ascii", "replace")
        pos += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] == "= ":
            val = card[10:].split("/")[0].strip()
            hdr[key] = val
    pos = ((pos + BLOCK - 1) // BLOCK) * BLOCK
    bitpix = int(hdr["BITPIX"])
    naxis = int(hdr["NAXIS"])
    shape = [int(hdr[f"NAXIS{i + 1}"]) for i in range(naxis)][::-1]
    n = int(np.prod(shape)) if shape else 0
    dt = np.dtype(_DTYPES[bitpix])
    a = np.frombuffer(raw, dtype=dt, count=n, offset=pos).astype(np.float64)
    bscale = float(hdr.get("BSCALE", "1").replace("D", "E"))
    bzero = float(hdr.get("BZERO", "0").replace("D", "E"))
    if bscale != 1.0 or bzero != 0.0:
        a = a * bscale + bzero
    return a.reshape(shape)


def write_direct_records(path, rows: np.ndarray) -> None:
    """Fortran direct-access unformatted records, one (nx, ny) slab per record
    (write_binary_seq_2D, all_modules.f90:947-1006): raw native doubles."""
    np.ascontiguousarray(rows, dtype="<f8").tofile(str(path))
