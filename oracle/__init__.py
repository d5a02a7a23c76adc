"""CPU oracle for the L.A.Cosmic cleaning loop -- TEST INFRASTRUCTURE ONLY.

ctypes bindings of ``oracle/oracle.c``, a plain-C restatement of the reference
algorithm (src/cleaner.cuh, src/filters.cuh, src/replace.cuh, src/image.cuh).
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker: the product
library (``cusmic_amd``) never calls it.
"""

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

MODES = {"reflect": 0, "constant": 1, "nearest": 2, "mirror": 3, "wrap": 4}
LAPLACIAN = {"exact": 0, "close": 1}


class Image(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("error", ctypes.c_void_p),
        ("gain", ctypes.c_void_p),
        ("readnoise", ctypes.c_void_p),
        ("background", ctypes.c_void_p),
        ("mask", ctypes.c_void_p),
        ("width", ctypes.c_size_t),
        ("height", ctypes.c_size_t),
    ]


class Options(ctypes.Structure):
    _fields_ = [
        ("contrast", ctypes.c_double),
        ("cr_threshold", ctypes.c_double),
        ("neighbor_threshold", ctypes.c_double),
        ("gain", ctypes.c_double),
        ("readnoise", ctypes.c_double),
        ("background", ctypes.c_double),
        ("maxiter", ctypes.c_int),
        ("border", ctypes.c_int),
    ]


def build(force=False):
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        p = ctypes.c_void_p
        _lib.oracle_clean.argtypes = [
            ctypes.POINTER(Image), ctypes.c_size_t, ctypes.POINTER(Options), p, p,
            ctypes.c_int, ctypes.c_int, p, p, p, p, ctypes.c_char_p, ctypes.c_size_t]
        _lib.oracle_clean.restype = ctypes.c_int
        _lib.oracle_median.argtypes = [p, p, ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.c_int]
        _lib.oracle_laplacian.argtypes = [p, p, ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.c_int]
        _lib.oracle_dilate.argtypes = [p, p, ctypes.c_long, ctypes.c_long]
        _lib.oracle_border_index.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_int]
        _lib.oracle_border_index.restype = ctypes.c_long
        _lib.oracle_replace.argtypes = [p, p, p, ctypes.c_long, ctypes.c_long, p]
        _lib.oracle_replace.restype = ctypes.c_int
        _lib.oracle_bitonic32.argtypes = [p]
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def median(image, size, mode="mirror"):
    a = np.ascontiguousarray(image, dtype=np.float64)
    out = np.empty_like(a)
    lib().oracle_median(_ptr(a), _ptr(out), a.shape[1], a.shape[0], size, MODES[mode])
    return out


def laplacian(image, mode="mirror", order="exact"):
    a = np.ascontiguousarray(image, dtype=np.float64)
    out = np.empty_like(a)
    lib().oracle_laplacian(_ptr(a), _ptr(out), a.shape[1], a.shape[0], MODES[mode], LAPLACIAN[order])
    return out


def dilate(mask):
    a = np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.empty_like(a)
    lib().oracle_dilate(_ptr(a), _ptr(out), a.shape[1], a.shape[0])
    return out


def border_index(x, n, mode):
    return lib().oracle_border_index(x, n, MODES[mode])


def replace(clean, targets, donors, record=False):
    """Replace targets in place from donors; returns donor indices if asked."""
    assert clean.flags.c_contiguous and clean.dtype == np.float64
    t = np.ascontiguousarray(targets, dtype=np.uint8)
    d = np.ascontiguousarray(donors, dtype=np.uint8)
    idx = np.full(clean.shape + (2,), -1, np.int32) if record else None
    status = lib().oracle_replace(_ptr(clean), _ptr(t), _ptr(d), clean.shape[1], clean.shape[0], _ptr(idx))
    if status == -1:
        raise ValueError("no finite replacement donors")
    return idx


def bitonic32(values):
    """The reference's 32-lane value network applied to 32 doubles."""
    v = np.array(values, dtype=np.float64)
    assert v.shape == (32,)
    lib().oracle_bitonic32(_ptr(v))
    return v


class OracleError(Exception):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def clean_frames(frames, contrast=3.0, cr_threshold=5.0, neighbor_threshold=3.0,
                 gain=float("nan"), readnoise=0.0, background=float("nan"), maxiter=4,
                 border_mode="mirror", laplacian_order="exact", frame_gain=None,
                 frame_readnoise=None, record_donors=False, threads=1):
    """Clean a list of frames (dicts with 'data' and optional maps) like the C ABI.

    Returns (list of cleaned frames, list of masks, donors or None, passes).
    Raises OracleError with the cusmic status on failure.
    """
    keep = []
    ims = (Image * len(frames))()
    sizes = []
    for f, fr in enumerate(frames):
        data = _f64(fr["data"])
        h, w = data.shape
        maps = {}
        for name in ("error", "gain", "readnoise", "background"):
            m = fr.get(name)
            maps[name] = None if m is None else _f64(np.broadcast_to(m, data.shape))
        m = fr.get("mask")
        mask = None if m is None else np.ascontiguousarray(np.broadcast_to(m, data.shape), dtype=np.uint8)
        keep += [data, mask, *maps.values()]
        ims[f] = Image(_ptr(data), _ptr(maps["error"]), _ptr(maps["gain"]), _ptr(maps["readnoise"]),
                       _ptr(maps["background"]), _ptr(mask), w, h)
        sizes.append((h, w))
    opts = Options(contrast, cr_threshold, neighbor_threshold, gain, readnoise, background,
                   maxiter, MODES[border_mode])
    total = sum(h * w for h, w in sizes)
    clean = np.empty(total, np.float64)
    mask = np.empty(total, np.uint8)
    donors = np.empty(2 * total, np.int32) if record_donors else None
    passes = np.zeros(len(frames), np.int32)
    fg = _f64(frame_gain)
    frn = _f64(frame_readnoise)
    msg = ctypes.create_string_buffer(256)
    status = lib().oracle_clean(ims, len(frames), ctypes.byref(opts), _ptr(fg), _ptr(frn),
                                LAPLACIAN[laplacian_order], threads, _ptr(clean), _ptr(mask),
                                _ptr(donors), _ptr(passes), msg, 256)
    if status:
        raise OracleError(status, msg.value.decode())
    outs, masks, dons = [], [], []
    off = 0
    for h, w in sizes:
        outs.append(clean[off:off + h * w].reshape(h, w))
        masks.append(mask[off:off + h * w].reshape(h, w).astype(bool))
        if record_donors:
            dons.append(donors[2 * off:2 * (off + h * w)].reshape(h, w, 2))
        off += h * w
    return outs, masks, (dons if record_donors else None), passes
