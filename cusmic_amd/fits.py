"""FITS images on the host: the on-disk side of the command-line cleaner.

The reference reads and writes FITS through Astropy (mod/cusmic/io.py:21-58)
and CFITSIO (src/io.c:21-121); neither is available here, so this module
implements the parts of the FITS standard (version 4.0) that contract uses:

* reading: 2-D images in the primary HDU or a named IMAGE extension, BITPIX
  8/16/32/64/-32/-64, plain or gzip-compressed; integer pixels equal to
  BLANK become NaN, BSCALE/BZERO are applied as ``raw * scale + zero`` in
  float64 (IEEE pixels are never substituted);
* writing: a float64 primary image that keeps the source header (less
  BSCALE, BZERO and BLANK), a HISTORY card, and a uint8 ``CRMASK`` extension,
  each HDU with CHECKSUM and DATASUM cards (the FITS checksum convention);
  an existing output is never overwritten unless asked, and a name that
  starts with '!' (a CFITSIO overwrite prefix) is refused.
"""

import gzip
import os
from pathlib import Path

import numpy as np

BLOCK = 2880
CARD = 80
_BITPIX = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}
# keywords that describe the data layout or its scaling: written anew, never copied
_STRUCTURAL = {"SIMPLE", "XTENSION", "BITPIX", "NAXIS", "EXTEND", "PCOUNT", "GCOUNT",
               "BSCALE", "BZERO", "BLANK", "CHECKSUM", "DATASUM", "END"}


class Header:
    """The cards of one HDU in order: (keyword, value, comment) triples, with
    commentary cards (HISTORY, COMMENT, blank) holding their text as value."""

    def __init__(self, cards=()):
        self.cards = [tuple(c) for c in cards]

    def copy(self):
        return Header(self.cards)

    def _find(self, key):
        for i, (k, _, _) in enumerate(self.cards):
            if k == key:
                return i
        return -1

    def __contains__(self, key):
        return self._find(key) >= 0

    def __getitem__(self, key):
        i = self._find(key)
        if i < 0:
            raise KeyError(key)
        return self.cards[i][1]

    def get(self, key, default=None):
        i = self._find(key)
        return default if i < 0 else self.cards[i][1]

    def __setitem__(self, key, value):
        i = self._find(key)
        if i < 0:
            self.cards.append((key, value, ""))
        else:
            self.cards[i] = (key, value, self.cards[i][2])

    def pop(self, key, default=None):
        i = self._find(key)
        if i < 0:
            return default
        return self.cards.pop(i)[1]

    def add_history(self, text):
        self.cards.append(("HISTORY", text, ""))

    def keys(self):
        return [k for k, _, _ in self.cards]


def _parse_value(text):
    """The value of a card's value field: str, bool, int, float or None."""
    text = text.strip()
    if not text:
        return None
    if text.startswith("'"):
        out, i = [], 1
        while i < len(text):
            if text[i] == "'":
                if i + 1 < len(text) and text[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(text[i])
            i += 1
        return "".join(out).rstrip()
    value = text.split("/", 1)[0].strip()
    if value in ("T", "F"):
        return value == "T"
    try:
        return int(value)
    except ValueError:
        pass
    try:
        return float(value.replace("D", "E"))
    except ValueError:
        return value


def _split_comment(text):
    """Comment of a value field ('/' outside a quoted string)."""
    quoted = False
    for i, ch in enumerate(text):
        if ch == "'":
            quoted = not quoted
        elif ch == "/" and not quoted:
            return text[i + 1:].strip()
    return ""


def _read_header(raw, pos):
    cards = []
    while True:
        block = raw[pos:pos + BLOCK]
        if len(block) < BLOCK:
            raise ValueError("truncated FITS header")
        pos += BLOCK
        for i in range(0, BLOCK, CARD):
            card = block[i:i + CARD].decode("ascii", errors="replace")
            key = card[:8].rstrip()
            if key == "END":
                return Header(cards), pos
            if card[8:10] == "= ":
                cards.append((key, _parse_value(card[10:]), _split_comment(card[10:])))
            else:
                cards.append((key, card[8:].rstrip(), ""))


def _hdus(raw):
    """(header, data offset, data bytes) of every HDU."""
    pos = 0
    while pos < len(raw):
        header, pos = _read_header(raw, pos)
        naxis = int(header.get("NAXIS", 0))
        count = 1
        for k in range(1, naxis + 1):
            count *= int(header[f"NAXIS{k}"])
        nbytes = count * abs(int(header["BITPIX"])) // 8 if naxis else 0
        yield header, pos, nbytes
        pos += -(-nbytes // BLOCK) * BLOCK


def _load(path):
    raw = Path(path).read_bytes()
    return gzip.decompress(raw) if raw[:2] == b"\x1f\x8b" else raw


def read_fits(path, dtype=None, *, ext=None):
    """Pixels (native byte order) and the Header of a 2-D FITS image.

    ``ext`` names an IMAGE extension (EXTNAME); without it the primary HDU
    is read, or the first extension when the primary has no data
    (src/io.c:29-34)."""
    raw = _load(path)
    chosen = None
    for header, offset, nbytes in _hdus(raw):
        if ext is None:
            if int(header.get("NAXIS", 0)) or "SIMPLE" not in header:
                chosen = (header, offset)
                break
        elif header.get("EXTNAME") == ext:
            chosen = (header, offset)
            break
    if chosen is None:
        raise KeyError(f"{path}: no image HDU {ext!r}" if ext else f"{path}: no image data")
    header, offset = chosen
    bitpix = int(header["BITPIX"])
    if bitpix not in _BITPIX:
        raise TypeError(f"{path}: expected real image pixels")
    naxis = int(header.get("NAXIS", 0))
    shape = tuple(int(header[f"NAXIS{k}"]) for k in range(naxis, 0, -1))
    count = int(np.prod(shape)) if naxis else 0
    data = np.frombuffer(raw, _BITPIX[bitpix], count, offset).reshape(shape)
    scale = float(header.get("BSCALE", 1.0))
    zero = float(header.get("BZERO", 0.0))
    blank = header.get("BLANK") if bitpix > 0 else None
    if scale != 1 or zero != 0 or blank is not None:
        missing = data == blank if blank is not None else None
        pixels = data.astype(np.float64)
        pixels *= scale
        pixels += zero
        if missing is not None:
            pixels[missing] = np.nan
    else:
        pixels = data.astype(data.dtype.newbyteorder("="))
    return pixels.astype(dtype or pixels.dtype, copy=False), header


# --- the FITS checksum convention (Seaman, Pence & Rots) ---------------------

def ones_complement_sum(data, start=0):
    """32-bit ones' complement sum of big-endian 32-bit words (length a
    multiple of 4), continuing from ``start``."""
    words = np.frombuffer(data, ">u4").astype(np.uint64)
    total = int(words.sum()) + start
    while total >> 32:
        total = (total & 0xFFFFFFFF) + (total >> 32)
    return total


def checksum_string(value):
    """The 16 printable characters whose HDU sum complements ``value``."""
    forbidden = set(b":;<=>?@[\\]^_`")
    out = [0] * 16
    for i in range(4):
        byte = (value >> (24 - 8 * i)) & 0xFF
        quotient, remainder = byte // 4 + 0x30, byte % 4
        ch = [quotient + remainder, quotient, quotient, quotient]
        while True:
            moved = False
            for j in (0, 2):
                if ch[j] in forbidden or ch[j + 1] in forbidden:
                    ch[j] += 1
                    ch[j + 1] -= 1
                    moved = True
            if not moved:
                break
        for j in range(4):
            out[4 * j + i] = ch[j]
    return bytes(out[(k + 15) % 16] for k in range(16)).decode("ascii")


def _format_value(value):
    if isinstance(value, bool):
        return f"{'T' if value else 'F':>20}"
    if isinstance(value, (int, np.integer)):
        return f"{int(value):>20}"
    if isinstance(value, (float, np.floating)):
        text = repr(float(value)).upper()
        if "E" not in text and "." not in text and "INF" not in text and "NAN" not in text:
            text += "."
        return f"{text:>20}"
    text = str(value).replace("'", "''")
    return f"'{text:<8}'"


def _card(key, value, comment=""):
    if key in ("HISTORY", "COMMENT", ""):
        return f"{key:<8}{value}"[:CARD].ljust(CARD)
    text = f"{key:<8}= {_format_value(value)}"
    if comment:
        text += f" / {comment}"
    return text[:CARD].ljust(CARD)


def _hdu_bytes(first, header, pixels):
    """One HDU (header blocks and padded data) with CHECKSUM and DATASUM.
    ``first`` are the structural cards."""
    data = pixels.tobytes()
    data += b"\0" * (-len(data) % BLOCK)
    datasum = ones_complement_sum(data)
    cards = list(first) + [c for c in header.cards if c[0] not in _STRUCTURAL and not c[0].startswith("NAXIS")]
    cards += [("CHECKSUM", "0" * 16, "HDU checksum"), ("DATASUM", str(datasum), "data unit checksum")]
    text = "".join(_card(*c) for c in cards) + "END".ljust(CARD)
    text += " " * (-len(text) % BLOCK)
    head = bytearray(text.encode("ascii"))
    total = ones_complement_sum(bytes(head), datasum)
    code = checksum_string(~total & 0xFFFFFFFF)
    at = head.find(b"CHECKSUM= '") + len("CHECKSUM= '")
    head[at:at + 16] = code.encode("ascii")
    return bytes(head) + data


def write_fits(path, data, mask=None, *, header=None, overwrite=False):
    """Write float64 pixels as the primary image and, if given, the mask as a
    uint8 CRMASK extension. Refuses an existing file unless ``overwrite``."""
    path = os.fspath(path)
    if not path or path.startswith("!"):
        raise ValueError("Choose a new output filename")
    pixels = np.ascontiguousarray(np.asarray(data, dtype=np.float64))
    if pixels.ndim != 2:
        raise ValueError("write_fits expects one 2D image")
    h, w = pixels.shape
    header = header.copy() if header is not None else Header()
    primary = [("SIMPLE", True, "conforms to FITS standard"), ("BITPIX", -64, "array data type"),
               ("NAXIS", 2, "number of array dimensions"), ("NAXIS1", w, ""), ("NAXIS2", h, ""),
               ("EXTEND", True, "")]
    out = _hdu_bytes(primary, header, pixels.astype(">f8"))
    if mask is not None:
        flags = np.ascontiguousarray(np.asarray(mask)).astype(np.uint8)
        if flags.shape != pixels.shape:
            raise ValueError("mask must have the image's shape")
        ext = [("XTENSION", "IMAGE", "Image extension"), ("BITPIX", 8, "array data type"),
               ("NAXIS", 2, "number of array dimensions"), ("NAXIS1", w, ""), ("NAXIS2", h, ""),
               ("PCOUNT", 0, "number of parameters"), ("GCOUNT", 1, "number of groups"),
               ("EXTNAME", "CRMASK", "extension name")]
        out += _hdu_bytes(ext, Header(), flags)
    mode = "wb" if overwrite else "xb"
    with open(path, mode) as f:
        f.write(out)


def verify_checksums(path):
    """For every HDU: (its stored DATASUM equals the data's sum, the whole HDU
    sums to ones' complement zero). HDUs without the cards give (None, None)."""
    raw = _load(path)
    result = []
    pos = 0
    for header, offset, nbytes in _hdus(raw):
        padded = -(-nbytes // BLOCK) * BLOCK
        data = raw[offset:offset + padded]
        if "CHECKSUM" not in header:
            result.append((None, None))
            pos = offset + padded
            continue
        datasum = ones_complement_sum(data)
        stored = header.get("DATASUM")
        head = raw[pos:offset]
        total = ones_complement_sum(head, datasum)
        result.append((stored is not None and int(stored) == datasum, total == 0xFFFFFFFF))
        pos = offset + padded
    return result
