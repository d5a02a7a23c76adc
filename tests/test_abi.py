"""The C ABI library: loads without a GPU, exports what the headers declare,
and keeps the host-only semantics of test/test_api.c (maxiter = 0 runs on the
CPU; invalid calls leave host outputs untouched)."""

import ctypes
import math
import re

import numpy as np

from conftest import ROOT

from cusmic_amd import _lib


def declared_functions():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names |= set(re.findall(r"^(?:const\s+)?(?:char|int|void)\s*\*?\s*(\w+)\s*\(", text, flags=re.M))
    return names


def test_exports_match_headers():
    h = _lib.lib()
    names = declared_functions()
    assert names == set(_lib.EXPORTS)
    for name in names:
        assert getattr(h, name) is not None


def test_versions_and_defaults():
    h = _lib.lib()
    assert h.cusmic_version().decode().startswith("cusmic-amd")
    assert h.cusmic_cuda_version() > 0
    o = _lib.default_options()
    assert (o.contrast, o.cr_threshold, o.neighbor_threshold) == (3, 5, 3)
    assert math.isnan(o.gain) and math.isnan(o.background) and o.readnoise == 0
    assert (o.maxiter, o.border) == (4, _lib.MIRROR)
    h.cusmic_default_options(None)  # NULL-safe


def call_host(data, o, width, height, clean, mask):
    img = _lib.CusmicImage(data.ctypes.data, None, None, None, None, None, width, height)
    msg = ctypes.create_string_buffer(128)
    status = _lib.lib().remove_cosmics(ctypes.byref(img), ctypes.byref(o), clean.ctypes.data,
                                       mask.ctypes.data, msg, 128)
    return status, msg.value.decode()


def test_zero_iteration_host_copy():
    """test/test_api.c:35-43: -0.0 and a NaN payload survive bit for bit."""
    data = np.array([1, -0.0, 3, 4])
    data.view(np.uint64)[2] = 0x7FF8000000001234
    before = data.copy()
    clean = np.full(4, 9.0)
    mask = np.full(4, 9, np.uint8)
    o = _lib.default_options()
    o.maxiter = 0
    status, msg = call_host(data, o, 2, 2, clean, mask)
    assert status == _lib.OK, msg
    np.testing.assert_array_equal(clean.view(np.uint64), before.view(np.uint64))
    np.testing.assert_array_equal(data.view(np.uint64), before.view(np.uint64))
    assert not mask.any()


def test_invalid_call_keeps_outputs():
    """test/test_api.c:44-54"""
    data = np.array([1.0, 2, 3, 4])
    clean = np.full(4, 9.0)
    mask = np.full(4, 9, np.uint8)
    o = _lib.default_options()
    o.maxiter = 0
    o.cr_threshold = math.nan
    status, msg = call_host(data, o, 2, 2, clean, mask)
    assert status == _lib.INVALID and msg == "invalid detection settings"
    assert (clean == 9).all() and (mask == 9).all()


def test_validation_messages():
    data = np.ones(4)
    clean = np.zeros(4)
    mask = np.zeros(4, np.uint8)
    o = _lib.default_options()
    status, msg = call_host(data, o, 2, 2, clean, mask)
    assert status == _lib.INVALID and msg == "positive gain or error is required"
    o.gain = 2.0
    o.readnoise = -1
    status, msg = call_host(data, o, 2, 2, clean, mask)
    assert status == _lib.INVALID and msg == "read noise must be finite and nonnegative"
    o.readnoise = 0
    status, msg = call_host(data, o, 0, 2, clean, mask)
    assert status == _lib.INVALID and msg == "invalid image dimensions"
    o.maxiter = -1
    status, msg = call_host(data, o, 2, 2, clean, mask)
    assert status == _lib.INVALID and msg == "invalid detection settings"
    o.maxiter = 0
    o.background = math.inf
    status, msg = call_host(data, o, 2, 2, clean, mask)
    assert status == _lib.INVALID and msg == "background must be finite"


def test_host_map_contents_checked():
    data = np.ones(4)
    error = np.array([1.0, 0.0, 1.0, 1.0])
    clean = np.zeros(4)
    mask = np.zeros(4, np.uint8)
    o = _lib.default_options()
    img = _lib.CusmicImage(data.ctypes.data, error.ctypes.data, None, None, None, None, 2, 2)
    msg = ctypes.create_string_buffer(128)
    status = _lib.lib().remove_cosmics(ctypes.byref(img), ctypes.byref(o), clean.ctypes.data,
                                       mask.ctypes.data, msg, 128)
    assert status == _lib.INVALID and msg.value == b"error must be finite and positive"
