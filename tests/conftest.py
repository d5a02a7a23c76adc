import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libcusmic.so")


@pytest.fixture(scope="session")
def fixture_scene():
    """The reference's saved L.A.Cosmic scene (test/data, Apache-2.0): input,
    error map (== 5), cleaned pixels and the 6,518-pixel CRMASK."""
    from oracle.fits import read_fits

    d = GOLDEN / "fits"
    return (read_fits(d / "input.fits.gz")[0].astype(np.float64),
            read_fits(d / "error.fits.gz")[0].astype(np.float64),
            read_fits(d / "reference.fits.gz")[0],
            read_fits(d / "reference.fits.gz", ext="CRMASK")[0].astype(bool))


@pytest.fixture(scope="session")
def cuda():
    """The GPU: fails (does not skip) when the marker selected a GPU test
    without a device, like the reference's cp fixture (test/conftest.py:22-31)."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device available", pytrace=False)
    import cusmic_amd._lib as L

    L.lib()
    return torch.device("cuda", 0)


def bits(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.view(np.uint64)


def assert_same_pixels(actual, expected):
    np.testing.assert_array_equal(bits(actual), bits(expected))
