"""FITS reading and writing (cusmic_amd.fits) on the CPU.

Pins: the saved reference fixtures (test/data, written with Astropy) carry
CHECKSUM and DATASUM cards, so our ones' complement sum must reproduce each
DATASUM and each stored CHECKSUM must bring its HDU to -0; the checksum
string we would write for those headers must equal the stored one. Scaled
and BLANK integer pixels follow the reference's own expectations
(test/test_io.py:11-58, test/test_io.c).
"""

import numpy as np
import pytest

from conftest import GOLDEN
from cusmic_amd import fits

FIXTURES = [GOLDEN / "fits" / f"{n}.fits.gz" for n in ("input", "error", "reference")]


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: p.name)
def test_fixture_checksums(path):
    sums = fits.verify_checksums(path)
    assert sums and all(s == (True, True) for s in sums), sums


def test_checksum_string_matches_fixture():
    """Recompute each fixture HDU's CHECKSUM from its header with the card set
    to zeros: the encoding must give the stored characters."""
    checked = 0
    for path in FIXTURES:
        raw = fits._load(path)
        pos = 0
        for header, offset, nbytes in fits._hdus(raw):
            padded = -(-nbytes // fits.BLOCK) * fits.BLOCK
            head = bytearray(raw[pos:offset])
            at = head.find(b"CHECKSUM= '") + len("CHECKSUM= '")
            stored = head[at:at + 16].decode()
            head[at:at + 16] = b"0" * 16
            datasum = fits.ones_complement_sum(raw[offset:offset + padded])
            total = fits.ones_complement_sum(bytes(head), datasum)
            assert fits.checksum_string(~total & 0xFFFFFFFF) == stored
            pos = offset + padded
            checked += 1
    assert checked == 4


def test_fixture_images():
    image, header = fits.read_fits(FIXTURES[0], dtype="float64")
    error, _ = fits.read_fits(FIXTURES[1], dtype="float64")
    crmask, _ = fits.read_fits(FIXTURES[2], ext="CRMASK")
    assert image.shape == error.shape == crmask.shape == (512, 512)
    assert image.dtype == np.float64 and image.dtype.isnative
    assert (error == 5).all() and crmask.sum() == 6518
    assert header["NAXIS1"] == 512 and header["BITPIX"] == -64


def raw_fits(hdus):
    """Bytes of a FITS file from (cards, big-endian array) pairs."""
    out = b""
    for i, (cards, data) in enumerate(hdus):
        text = "".join(fits._card(*c) for c in cards) + "END".ljust(80)
        text += " " * (-len(text) % 2880)
        body = data.tobytes()
        out += text.encode() + body + b"\0" * (-len(body) % 2880)
    return out


def image_cards(bitpix, shape, first, extra=()):
    h, w = shape
    head = [("SIMPLE", True)] if first else [("XTENSION", "IMAGE")]
    head += [("BITPIX", bitpix), ("NAXIS", 2), ("NAXIS1", w), ("NAXIS2", h)]
    if not first:
        head += [("PCOUNT", 0), ("GCOUNT", 1)]
    return head + list(extra)


def test_scaled_and_blank_pixels(tmp_path):
    """test/test_io.py:27-58: IEEE bits kept; BSCALE/BZERO and BLANK."""
    bits = np.array([0x8000000000000000, 0x7FF0000000000000, 0x7FF8000000001234, 0x401C000000000000],
                    dtype=np.uint64)
    raw = np.array([[1, -32768], [2, 3]], dtype=">i2")
    path = tmp_path / "pixels.fits"
    path.write_bytes(raw_fits([
        (image_cards(-64, (2, 2), True, [("EXTEND", True)]), bits.view(np.float64).reshape(2, 2).astype(">f8")),
        (image_cards(16, (2, 2), False, [("EXTNAME", "SCALED"), ("BSCALE", 0.1), ("BZERO", 10.2),
                                          ("BLANK", -32768)]), raw),
        (image_cards(16, (2, 2), False, [("EXTNAME", "UNSIGNED"), ("BZERO", 32768), ("BLANK", -32768)]), raw),
    ]))
    pixels, _ = fits.read_fits(path, dtype="float64")
    np.testing.assert_array_equal(pixels.view(np.uint64).ravel(), bits)
    pixels, header = fits.read_fits(path, dtype="float64", ext="SCALED")
    want = np.array([0x4024999999999999, 0x7FF8000000000000, 0x4024CCCCCCCCCCCC, 0x4025000000000000],
                    dtype=np.uint64)
    np.testing.assert_array_equal(pixels.view(np.uint64).ravel(), want)
    assert header["BLANK"] == -32768
    pixels, _ = fits.read_fits(path, dtype="float64", ext="UNSIGNED")
    assert np.isnan(pixels[0, 1])
    np.testing.assert_array_equal(pixels[[0, 1, 1], [0, 0, 1]], [32769, 32770, 32771])


def test_integer_blank_round_trip(tmp_path):
    """test/test_io.py:11-24: the written image drops BLANK and keeps NaN."""
    data = np.arange(9, dtype=">i2").reshape(3, 3)
    data[1, 1] = -32768
    source, output = tmp_path / "input.fits", tmp_path / "output.fits"
    source.write_bytes(raw_fits([(image_cards(16, (3, 3), True, [("BLANK", -32768)]), data)]))
    pixels, header = fits.read_fits(source, dtype="float64")
    fits.write_fits(output, pixels, np.zeros(data.shape, bool), header=header)
    back, out_header = fits.read_fits(output)
    np.testing.assert_array_equal(back, pixels)
    assert np.isnan(pixels[1, 1]) and header["BLANK"] == -32768
    assert "BLANK" not in out_header
    assert not fits.read_fits(output, ext="CRMASK")[0].any()
    assert fits.verify_checksums(output) == [(True, True), (True, True)]


def test_write_keeps_bits_and_header(tmp_path):
    rng = np.random.default_rng(0)
    data = rng.normal(size=(7, 11))
    data.view(np.uint64)[0, :3] = [0x8000000000000000, 0x7FF8000000001234, 0xFFF0000000000000]
    mask = rng.random(data.shape) < 0.3
    header = fits.Header([("OBJECT", "M31 'core'", "target"), ("EXPTIME", 30.5, "s"), ("NCOMBINE", 3, ""),
                          ("BSCALE", 2.0, ""), ("HISTORY", "raw frame", "")])
    header.add_history("Cosmic rays removed with cusmic")
    path = tmp_path / "clean.fits"
    fits.write_fits(path, data, mask, header=header)
    back, h = fits.read_fits(path)
    np.testing.assert_array_equal(back.view(np.uint64), data.view(np.uint64))
    np.testing.assert_array_equal(fits.read_fits(path, ext="CRMASK")[0], mask.astype(np.uint8))
    assert h["OBJECT"] == "M31 'core'" and h["EXPTIME"] == 30.5 and h["NCOMBINE"] == 3
    assert "BSCALE" not in h and h.get("EXTNAME") is None
    assert [v for k, v, _ in h.cards if k == "HISTORY"] == ["raw frame", "Cosmic rays removed with cusmic"]
    assert fits.verify_checksums(path) == [(True, True), (True, True)]
    assert path.stat().st_size % 2880 == 0


def test_refuses_overwrite(tmp_path):
    path = tmp_path / "out.fits"
    fits.write_fits(path, np.zeros((2, 2)))
    with pytest.raises(FileExistsError):
        fits.write_fits(path, np.ones((2, 2)))
    fits.write_fits(path, np.ones((2, 2)), overwrite=True)
    assert fits.read_fits(path)[0].sum() == 4
    with pytest.raises(ValueError, match="new output"):
        fits.write_fits("!" + str(tmp_path / "x.fits"), np.zeros((2, 2)))
