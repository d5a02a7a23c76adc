"""ctypes binding of libcusmic.so (include/cusmic.h, include/cusmic_amd.h).

The library must exist: the package has no fallback path. Build it with
``python -m cusmic_amd.build`` (or ``__graft_entry__.build()``).
"""

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libcusmic.so"

REFLECT, CONSTANT, NEAREST, MIRROR, WRAP = range(5)
OK, INVALID, CUDA_ERROR, NO_MEMORY = range(4)
LAPLACIAN_EXACT, LAPLACIAN_CLOSE = 0, 1
MODES = {"reflect": REFLECT, "constant": CONSTANT, "nearest": NEAREST, "mirror": MIRROR, "wrap": WRAP}

# Every symbol the headers declare.
EXPORTS = (
    "cusmic_version", "cusmic_cuda_version", "cusmic_default_options", "remove_cosmics",
    "remove_cosmics_device", "remove_cosmics_batch_device", "remove_cosmics_ragged_device",
    "remove_cosmics_batch_device_noalloc", "cusmic_adjoint_device", "cusmic_selftest_arith",
)


class CusmicImage(ctypes.Structure):
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


class CusmicOptions(ctypes.Structure):
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


class RaggedOptions(ctypes.Structure):
    _fields_ = [
        ("frame_gain", ctypes.c_void_p),
        ("frame_readnoise", ctypes.c_void_p),
        ("donors", ctypes.c_void_p),
        ("passes", ctypes.c_void_p),
        ("laplacian", ctypes.c_int),
        ("profile", ctypes.c_void_p),
        ("new_pixels", ctypes.c_void_p),
        ("new_pixels_size", ctypes.c_size_t),
    ]


PROFILE_FIELDS = ("detect_ms", "replace_ms", "prologue_ms", "epilogue_ms", "detect_launches",
                  "detect_pixels", "flagged", "seeds", "candidates", "grown", "widened",
                  "significance_ms", "detect_pass_pixels", "replace_pass_pixels", "first_pass_ms")


class CusmicError(RuntimeError):
    """A nonzero cusmic status; .status holds the code."""

    def __init__(self, status, message):
        super().__init__(message or f"cusmic status {status}")
        self.status = status


_lib = None


def lib():
    """Load libcusmic.so, failing loudly when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m cusmic_amd.build`")
    try:
        import torch  # noqa: F401  (share torch's HIP runtime when torch is in use)
    except ImportError:
        pass
    h = ctypes.CDLL(str(LIB_PATH))
    p, sz, c = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p
    img = ctypes.POINTER(CusmicImage)
    opt = ctypes.POINTER(CusmicOptions)
    h.cusmic_version.restype = ctypes.c_char_p
    h.cusmic_version.argtypes = []
    h.cusmic_cuda_version.restype = ctypes.c_int
    h.cusmic_cuda_version.argtypes = []
    h.cusmic_default_options.argtypes = [opt]
    h.cusmic_default_options.restype = None
    h.remove_cosmics.argtypes = [img, opt, p, p, c, sz]
    h.remove_cosmics_device.argtypes = [img, opt, p, p, p, c, sz]
    h.remove_cosmics_batch_device.argtypes = [img, sz, opt, p, p, p, c, sz]
    h.remove_cosmics_batch_device_noalloc.argtypes = [img, sz, opt, p, p, p, c, sz]
    h.remove_cosmics_ragged_device.argtypes = [img, sz, opt, ctypes.POINTER(RaggedOptions), p, p,
                                               p, c, sz]
    h.cusmic_adjoint_device.argtypes = [img, sz, p, p, p, p, p, c, sz]
    h.cusmic_selftest_arith.argtypes = [p, p, p, p, sz, p]
    for name in ("remove_cosmics", "remove_cosmics_device", "remove_cosmics_batch_device",
                 "remove_cosmics_batch_device_noalloc", "remove_cosmics_ragged_device",
                 "cusmic_adjoint_device", "cusmic_selftest_arith"):
        getattr(h, name).restype = ctypes.c_int
    _lib = h
    return h


def default_options():
    o = CusmicOptions()
    lib().cusmic_default_options(ctypes.byref(o))
    return o


def check(status, msg):
    if status:
        raise CusmicError(status, msg.value.decode(errors="replace") if msg is not None else "")
