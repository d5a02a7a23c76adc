"""Detection settings and the call that cleans an Image.

The settings, their defaults and the messages of their checks are the
reference's (mod/cusmic/cleaner.py:26-58), as is the rule that the noise model
needs both gain and read noise (:71-74) and the per-iteration log line
(:108). The cleaning itself is one call of remove_cosmics_ragged_device on
the data's device and current stream: the whole loop runs on the GPU.
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


def _is_setting(value):
    """A finite, nonnegative real number (bool is not a number here)."""
    return not isinstance(value, bool) and isinstance(value, Real) and isfinite(value) and value >= 0


def _is_count(value):
    return not isinstance(value, bool) and isinstance(value, Integral) and value >= 0


def frames_of(image: Image):
    """The library's frame descriptors for an Image, and its scalar
    calibration (gain, read noise, background) for the call's options."""
    stack = image.data.dim() == 3
    count = image.data.shape[0] if stack else 1
    scalars = {"gain": math.nan, "readnoise": 0.0, "background": math.nan}
    key = {"effective_gain": "gain", "readnoise": "readnoise", "background": "background"}

    def per_frame(name, f):
        t = getattr(image, name)
        if t is None:
            return None
        if t.dim() == 0:
            scalars[key[name]] = float(t)
            return None
        return t[f] if t.dim() == 3 else t

    noise = ("effective_gain", "readnoise") if image.error is None else ()
    frames = engine.Frames()
    for f in range(count):
        maps = {name: per_frame(name, f) for name in ("error", *noise, "background")}
        mask = image.mask
        if mask is not None and mask.dim() == 3:
            mask = mask[f]
        frames.add(image.data[f] if stack else image.data, maps["error"], maps.get("effective_gain"),
                   maps.get("readnoise"), maps["background"], mask)
    return frames, scalars


@dataclass(frozen=True)
class Cleaner:
    """Detection settings, reusable across images; calling it cleans one."""

    contrast: float = 3
    cr_threshold: float = 5
    neighbor_threshold: float = 3
    maxiter: int = 4
    border_mode: str = "mirror"

    def __post_init__(self):
        for name in ("contrast", "cr_threshold", "neighbor_threshold"):
            if not _is_setting(getattr(self, name)):
                raise ValueError(f"{name} must be finite and nonnegative")
        if not _is_count(self.maxiter):
            raise ValueError("maxiter must be a nonnegative integer")
        if not (isinstance(self.border_mode, str) and self.border_mode in MODES):
            raise ValueError(f"border_mode must be one of {', '.join(MODES)}")

    def run(self, image: Image, record_donors=False, laplacian=_lib.LAPLACIAN_EXACT):
        """Clean the image; returns (clean, mask, donors, passes), with donors
        of shape (..., 2) int32 when recorded, else None."""
        needs_model = self.maxiter and image.error is None
        if needs_model and (image.effective_gain is None or image.readnoise is None):
            raise ValueError("Provide error, or both effective_gain and readnoise")
        frames, scalars = frames_of(image)
        new_pixels = [0] * min(int(self.maxiter), 1 << 16)
        clean, mask, donors, passes = engine.run(
            frames, image.device, float(self.contrast), float(self.cr_threshold),
            float(self.neighbor_threshold), int(self.maxiter), _lib.MODES[self.border_mode],
            gain=scalars["gain"], readnoise=scalars["readnoise"], background=scalars["background"],
            record_donors=record_donors, laplacian=laplacian, new_pixels=new_pixels)
        # the stack runs as long as its slowest frame, as the reference's
        # loop over the whole stack does
        for i in range(max(passes, default=0)):
            log.info("Iteration %d: %d new cosmic-ray pixels", i + 1, new_pixels[i] if i < len(new_pixels) else 0)
        shape = image.data.shape
        if donors is not None:
            donors = donors.view(*shape, 2)
        return clean.view(shape), mask.view(torch.bool).view(shape), donors, passes

    def __call__(self, image: Image) -> tuple[Array, Array]:
        """Cleaned pixels and the cosmic-ray mask, new tensors on the data's
        device, computed on its current stream."""
        clean, mask, _, _ = self.run(image)
        return clean, mask
