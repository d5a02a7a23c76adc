"""Minimal FITS image reader for the saved reference fixtures (test infrastructure).

Reads 2-D primary or IMAGE-extension HDUs of BITPIX 8/16/32/64/-32/-64 from
plain or gzip files, applying BSCALE/BZERO like the reference readers
(src/io.c:21-75; mod/cusmic/io.py:21-45). No checksum verification.
"""

import gzip
from pathlib import Path

import numpy as np

BLOCK = 2880
DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


def _cards(raw, pos):
    header = {}
    while True:
        block = raw[pos:pos + BLOCK]
        if len(block) < BLOCK:
            raise ValueError("truncated FITS header")
        pos += BLOCK
        for i in range(0, BLOCK, 80):
            card = block[i:i + 80].decode("ascii")
            key = card[:8].strip()
            if key == "END":
                return header, pos
            if card[8:10] == "= ":
                value = card[10:].split("/")[0].strip()
                if value.startswith("'"):
                    value = card[10:].strip()[1:].split("'")[0].strip()
                header[key] = value


def read_fits(path, ext=None):
    """Return (array, header) for the primary HDU or the extension named ``ext``."""
    raw = Path(path).read_bytes()
    if raw[:2] == b"\x1f\x8b":
        raw = gzip.decompress(raw)
    pos = 0
    while pos < len(raw):
        header, pos = _cards(raw, pos)
        bitpix = int(header["BITPIX"])
        naxis = int(header["NAXIS"])
        shape = [int(header[f"NAXIS{i}"]) for i in range(naxis, 0, -1)]
        count = int(np.prod(shape)) if naxis else 0
        nbytes = count * abs(bitpix) // 8
        name = header.get("EXTNAME")
        if (ext is None and "SIMPLE" in header) or (ext is not None and name == ext):
            data = np.frombuffer(raw, DTYPES[bitpix], count, pos).reshape(shape)
            scale = float(header.get("BSCALE", 1))
            zero = float(header.get("BZERO", 0))
            if scale != 1 or zero != 0:
                data = data.astype(np.float64) * scale + zero
            else:
                data = data.astype(data.dtype.newbyteorder("="))
            return data, header
        pos += -(-nbytes // BLOCK) * BLOCK
    raise KeyError(f"no HDU {ext!r} in {path}")
