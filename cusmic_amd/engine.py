"""Calls into libcusmic.so for torch tensors on a ROCm device.

Builds cusmic_image descriptors from tensors (shared or per-frame maps,
scalar calibration), runs remove_cosmics_ragged_device on the tensors' device
and current stream, and maps status codes to Python exceptions (ValueError for
invalid input and missing donors, as mod/cusmic/replace.py:53-55 raises).
"""

import ctypes
import math

import torch

from . import _lib


def _raise(status, msg):
    text = msg.value.decode(errors="replace")
    if status == _lib.INVALID:
        raise ValueError(text)
    if status == _lib.NO_MEMORY:
        raise MemoryError(text)
    raise RuntimeError(text or f"cusmic status {status}")


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Frames:
    """Contiguous per-frame device planes kept alive for the duration of a call."""

    def __init__(self):
        self.keep = []
        self.images = []
        self.shapes = []
        self.gain = []
        self.readnoise = []

    def add(self, data, error=None, gain=None, readnoise=None, background=None, mask=None):
        """data: (h, w) float64 device tensor; maps: (h, w) tensors or None."""
        data = data.contiguous()
        h, w = data.shape
        planes = [data]
        for m in (error, gain, readnoise, background):
            planes.append(None if m is None else m.to(torch.float64).contiguous())
        if mask is not None:
            mask = mask.to(torch.uint8).contiguous()
        self.keep += planes + [mask]
        d, e, g, r, b = planes
        self.images.append(_lib.CusmicImage(d.data_ptr(), _ptr(e), _ptr(g), _ptr(r), _ptr(b),
                                            _ptr(mask), w, h))
        self.shapes.append((h, w))

    @property
    def total(self):
        return sum(h * w for h, w in self.shapes)


def run(frames, device, contrast, cr_threshold, neighbor_threshold, maxiter, border,
        gain=math.nan, readnoise=0.0, background=math.nan, frame_gain=None,
        frame_readnoise=None, record_donors=False, laplacian=_lib.LAPLACIAN_EXACT,
        clean=None, mask=None, profile=None, new_pixels=None):
    """Clean the frames in one call. Returns (clean, mask_u8, donors, passes) as
    flat device tensors (donors None unless recorded) and per-frame passes.
    A dict passed as ``profile`` receives the library's kernel timings; a list
    passed as ``new_pixels`` receives the new pixels of each pass (over all
    frames), one per element."""
    h = _lib.lib()
    nf = len(frames.images)
    total = frames.total
    if clean is None:
        clean = torch.empty(total, dtype=torch.float64, device=device)
    if mask is None:
        mask = torch.empty(total, dtype=torch.uint8, device=device)
    donors = torch.empty(2 * total, dtype=torch.int32, device=device) if record_donors else None
    images = (_lib.CusmicImage * nf)(*frames.images)
    o = _lib.CusmicOptions(contrast, cr_threshold, neighbor_threshold, gain, readnoise,
                           background, maxiter, border)
    passes = (ctypes.c_int32 * nf)()
    fg = (ctypes.c_double * nf)(*frame_gain) if frame_gain is not None else None
    frn = (ctypes.c_double * nf)(*frame_readnoise) if frame_readnoise is not None else None
    prof = (ctypes.c_double * len(_lib.PROFILE_FIELDS))() if profile is not None else None
    fresh = (ctypes.c_int64 * len(new_pixels))() if new_pixels else None
    ext = _lib.RaggedOptions(
        ctypes.cast(fg, ctypes.c_void_p) if fg is not None else None,
        ctypes.cast(frn, ctypes.c_void_p) if frn is not None else None,
        _ptr(donors), ctypes.cast(passes, ctypes.c_void_p), laplacian,
        ctypes.cast(prof, ctypes.c_void_p) if prof is not None else None,
        ctypes.cast(fresh, ctypes.c_void_p) if fresh is not None else None,
        len(new_pixels) if fresh is not None else 0)
    msg = ctypes.create_string_buffer(256)
    with torch.cuda.device(device):
        stream = torch.cuda.current_stream(device).cuda_stream
        status = h.remove_cosmics_ragged_device(images, nf, ctypes.byref(o), ctypes.byref(ext),
                                                _ptr(clean), _ptr(mask), ctypes.c_void_p(stream),
                                                msg, 256)
    if status:
        _raise(status, msg)
    if profile is not None:
        profile.update(zip(_lib.PROFILE_FIELDS, list(prof)))
    if fresh is not None:
        new_pixels[:] = list(fresh)
    return clean, mask, donors, list(passes)


def adjoint(shapes, device, mask_u8, donors, grad_clean):
    """grad_data from the recorded mask and donors (flat device tensors)."""
    h = _lib.lib()
    nf = len(shapes)
    images = (_lib.CusmicImage * nf)(*[_lib.CusmicImage(None, None, None, None, None, None, w, hh)
                                       for hh, w in shapes])
    grad_clean = grad_clean.to(torch.float64).contiguous()
    out = torch.empty_like(grad_clean)
    msg = ctypes.create_string_buffer(256)
    with torch.cuda.device(device):
        stream = torch.cuda.current_stream(device).cuda_stream
        status = h.cusmic_adjoint_device(images, nf, _ptr(mask_u8), _ptr(donors), _ptr(grad_clean),
                                         _ptr(out), ctypes.c_void_p(stream), msg, 256)
    if status:
        _raise(status, msg)
    return out
