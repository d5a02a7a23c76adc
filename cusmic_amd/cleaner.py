"""Reusable detection settings (mirrors mod/cusmic/cleaner.py:26-118).

``Cleaner(...)(image)`` validates its settings exactly as the reference, then
cleans the frame or stack on the input device's current stream through
``remove_cosmics_ragged_device``; the whole maxiter loop runs on the GPU.
"""

import logging
import math
from dataclasses import dataclass
from math import isfinite
from numbers import Integral, Real

import torch

from . import _lib, engine
from .image import Array, Image

log = logging.getLogger(__name__)

MODES = ("mirror", "reflect", "nearest", "wrap", "constant")


def frames_of(image: Image):
    """Per-frame descriptors of an Image plus the scalar calibration."""
    data = image.data
    stack = data.dim() == 3
    nf = data.shape[0] if stack else 1
    frames = engine.Frames()
    scalars = {"gain": math.nan, "readnoise": 0.0, "background": math.nan}

    def plane(value, f, name):
        if value is None:
            return None
        if value.dim() == 0:
            if name in scalars:
                scalars[name] = float(value)
            return None
        if value.dim() == 3:
            return value[f]
        return value

    for f in range(nf):
        d = data[f] if stack else data
        error = plane(image.error, f, "error")
        gain = readnoise = None
        if image.error is None:
            gain = plane(image.effective_gain, f, "gain")
            readnoise = plane(image.readnoise, f, "readnoise")
        background = plane(image.background, f, "background")
        mask = None if image.mask is None else (image.mask[f] if image.mask.dim() == 3 else image.mask)
        frames.add(d, error, gain, readnoise, background, mask)
    return frames, scalars


@dataclass(frozen=True)
class Cleaner:
    """Reusable detection settings; call with an Image to get pixels and a mask."""

    contrast:           float = 3
    cr_threshold:       float = 5
    neighbor_threshold: float = 3
    maxiter:            int   = 4
    border_mode:        str   = "mirror"

    def __post_init__(self):
        for name in ("contrast", "cr_threshold", "neighbor_threshold"):
            value = getattr(self, name)
            if (isinstance(value, bool) or not isinstance(value, Real) or not isfinite(value)
                    or value < 0):
                raise ValueError(f"{name} must be finite and nonnegative")
        if isinstance(self.maxiter, bool) or not isinstance(self.maxiter, Integral) or self.maxiter < 0:
            raise ValueError("maxiter must be a nonnegative integer")
        if not isinstance(self.border_mode, str) or self.border_mode not in MODES:
            raise ValueError(f"border_mode must be one of {', '.join(MODES)}")

    def run(self, image: Image, record_donors=False, laplacian=_lib.LAPLACIAN_EXACT):
        """Clean and return (clean, mask, donors, passes); donors is (..., 2) int32 or None."""
        if self.maxiter and image.error is None and (
                image.effective_gain is None or image.readnoise is None):
            raise ValueError("Provide error, or both effective_gain and readnoise")
        frames, scalars = frames_of(image)
        clean, mask, donors, passes = engine.run(
            frames, image.device, float(self.contrast), float(self.cr_threshold),
            float(self.neighbor_threshold), int(self.maxiter), _lib.MODES[self.border_mode],
            gain=scalars["gain"], readnoise=scalars["readnoise"], background=scalars["background"],
            record_donors=record_donors, laplacian=laplacian)
        shape = image.data.shape
        clean = clean.view(shape)
        mask = mask.view(torch.bool).view(shape)
        if donors is not None:
            donors = donors.view(*shape, 2)
        for f, n in enumerate(passes):
            log.info("Frame %d: %d detection passes", f, n)
        return clean, mask, donors, passes

    def __call__(self, image: Image):
        """Return independent arrays on the input device's current stream."""
        clean, mask, _, _ = self.run(image)
        return clean, mask
