"""The command-line cleaner (python -m cusmic_amd).

CPU: invalid images are rejected before any cleaning (test/test_io.py:61-68).
GPU: the reference's saved L.A.Cosmic scene through FITS files gives the saved
cleaned pixels bit for bit and its 6,518-pixel CRMASK; the validation cases
of test/test_reference.py:65-82; a scaled 16-bit source (as
test/test_reference.py:85-118 builds) against the oracle.
"""

import numpy as np
import pytest
from click.testing import CliRunner

import oracle
from conftest import GOLDEN, assert_same_pixels
from cusmic_amd import fits
from cusmic_amd.__main__ import main

FITS = GOLDEN / "fits"


@pytest.mark.parametrize("kind", ["no data", "3D"])
def test_cli_rejects_invalid_images(tmp_path, kind):
    source, output = tmp_path / "invalid.fits", tmp_path / "output.fits"
    if kind == "no data":
        cards = [("SIMPLE", True), ("BITPIX", 8), ("NAXIS", 0)]
        text = "".join(fits._card(*c) for c in cards) + "END".ljust(80)
        source.write_bytes((text + " " * (-len(text) % 2880)).encode())
    else:
        data = np.zeros((2, 3, 3), ">f8")
        cards = [("SIMPLE", True), ("BITPIX", -64), ("NAXIS", 3), ("NAXIS1", 3), ("NAXIS2", 3), ("NAXIS3", 2)]
        text = "".join(fits._card(*c) for c in cards) + "END".ljust(80)
        body = data.tobytes()
        source.write_bytes((text + " " * (-len(text) % 2880)).encode() + body + b"\0" * (-len(body) % 2880))
    result = CliRunner().invoke(main, [str(source), str(output), "--maxiter", "0"])
    assert result.exit_code == 1 and "Error:" in result.output, result.output
    assert not output.exists()


def test_cli_option_checks(tmp_path):
    source, output = FITS / "input.fits.gz", tmp_path / "out.fits"
    result = CliRunner().invoke(main, [str(source), str(output), "--readnoise", "1"])
    assert result.exit_code == 2 and "Provide --error or --gain" in result.output
    for options in (["--gain", "0"], ["--maxiter", "-1"], ["--contrast", "-1"]):
        result = CliRunner().invoke(main, [str(source), str(output), *options])
        assert result.exit_code != 0 and not output.exists()
    result = CliRunner().invoke(main, [str(source), "!" + str(output), "--maxiter", "0"])
    assert result.exit_code == 1 and "new output" in result.output


@pytest.mark.gpu
def test_cli_saved_reference(cuda, tmp_path):
    out = tmp_path / "clean.fits"
    args = [str(FITS / "input.fits.gz"), str(out), "--error", str(FITS / "error.fits.gz"), "--contrast", "1",
            "--cr-threshold", "5", "--neighbor-threshold", "5"]
    result = CliRunner().invoke(main, args)
    assert result.exit_code == 0, result.output
    assert "(6,518 cosmic-ray pixels)" in result.output
    pixels, header = fits.read_fits(out)
    assert_same_pixels(pixels, fits.read_fits(FITS / "reference.fits.gz")[0])
    np.testing.assert_array_equal(fits.read_fits(out, ext="CRMASK")[0].astype(bool),
                                  fits.read_fits(FITS / "reference.fits.gz", ext="CRMASK")[0].astype(bool))
    assert fits.verify_checksums(out) == [(True, True), (True, True)]
    assert "Cosmic rays removed with cusmic" in [v for k, v, _ in header.cards if k == "HISTORY"]


@pytest.mark.gpu
def test_cli_validation(cuda, tmp_path):
    """test/test_reference.py:65-82"""
    source = FITS / "input.fits.gz"
    output = tmp_path / "copy.fits"
    args = [str(source), str(output), "--maxiter", "0"]
    result = CliRunner().invoke(main, args)
    assert result.exit_code == 0, result.output
    np.testing.assert_array_equal(fits.read_fits(output)[0].view("uint64"), fits.read_fits(source)[0].view("uint64"))
    assert not fits.read_fits(output, ext="CRMASK")[0].any()
    assert CliRunner().invoke(main, args).exit_code != 0
    output = tmp_path / "invalid.fits"
    for options in (["--gain", "1", "--contrast", "nan"], ["--readnoise", "1"]):
        result = CliRunner().invoke(main, [str(source), str(output), *options])
        assert result.exit_code != 0 and not output.exists()


@pytest.mark.gpu
def test_cli_scaled_source(cuda, tmp_path):
    """A BSCALE/BZERO 16-bit source with a hot pixel and an error map."""
    data = np.ones((9, 9), ">i2")
    data[4, 4] = 10000
    cards = [("SIMPLE", True), ("BITPIX", 16), ("NAXIS", 2), ("NAXIS1", 9), ("NAXIS2", 9), ("BSCALE", 0.1),
             ("BZERO", 10.2)]
    text = "".join(fits._card(*c) for c in cards) + "END".ljust(80)
    body = data.tobytes()
    source, error = tmp_path / "scaled.fits", tmp_path / "error.fits"
    source.write_bytes((text + " " * (-len(text) % 2880)).encode() + body + b"\0" * (-len(body) % 2880))
    fits.write_fits(error, np.ones((9, 9)))
    pixels = fits.read_fits(source, dtype="float64")[0]
    for iterations in (0, 4):
        out = tmp_path / f"out-{iterations}.fits"
        result = CliRunner().invoke(main, [str(source), str(out), "--error", str(error), "--maxiter", str(iterations)])
        assert result.exit_code == 0, result.output
        want, wmask, _, _ = oracle.clean_frames([dict(data=pixels, error=np.ones((9, 9)))], maxiter=iterations)
        assert_same_pixels(fits.read_fits(out)[0], want[0])
        np.testing.assert_array_equal(fits.read_fits(out, ext="CRMASK")[0].astype(bool), wmask[0])
        assert wmask[0].sum() == (1 if iterations else 0)
