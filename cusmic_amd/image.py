"""Validated frames and calibration on a ROCm device (torch tensors).

Mirrors the reference's ``Image`` (mod/cusmic/image.py:22-79): the same
fields, shape rules and error messages, with torch tensors in place of CuPy
arrays. NumPy or CPU inputs are uploaded to the current GPU; GPU tensors are
borrowed, including non-contiguous views.
"""

from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

Array = Any


def _device_for(value):
    if isinstance(value, torch.Tensor) and value.is_cuda:
        return value.device
    return torch.device("cuda", torch.cuda.current_device())


def _as_tensor(value, device):
    if isinstance(value, torch.Tensor):
        return value.to(device) if value.device != device else value
    return torch.as_tensor(np.asarray(value), device=device)


def _real_numeric(t):
    return t.dtype != torch.bool and not t.is_complex()


@dataclass
class Image:
    """Borrowed float64 pixels with calibration validated on the input device.

    Maps match the data or one shared frame. Gain, read noise, and background
    may also be scalars. An error map overrides gain and read noise. See
    :func:`cusmic_amd.remove_cosmics` for units, exclusions, and streams.
    """

    data:           Array
    error:          Array | None = None
    mask:           Array | None = None
    background:     Array | float | None = None
    effective_gain: Array | float | None = None
    readnoise:      Array | float | None = None

    def __post_init__(self):
        device = _device_for(self.data)
        if not (isinstance(self.data, torch.Tensor) and self.data.is_cuda):
            self.data = _as_tensor(self.data, device)
        if self.data.dtype != torch.float64:
            raise TypeError("data must contain float64 pixels")
        if self.data.dim() not in (2, 3) or not self.data.numel():
            raise ValueError("data must be a nonempty 2D frame or 3D stack")

        shape, frame = tuple(self.data.shape), tuple(self.data.shape[-2:])
        fields = ("error",) if self.error is not None else ("effective_gain", "readnoise")
        with torch.cuda.device(device):
            for name in (*fields, "background"):
                value = getattr(self, name)
                if value is None:
                    continue
                value = _as_tensor(value, device)
                if not _real_numeric(value):
                    raise TypeError(f"{name} must be real numeric data")
                if tuple(value.shape) not in (shape, frame) and (name == "error" or value.dim() != 0):
                    raise ValueError(f"{name} has an incompatible shape")
                if not bool(torch.isfinite(value).all()):
                    raise ValueError(f"{name} must be finite")
                if name in ("error", "effective_gain") and bool((value <= 0).any()):
                    raise ValueError(f"{name} must be positive")
                if name == "readnoise" and bool((value < 0).any()):
                    raise ValueError("readnoise must be nonnegative")
                setattr(self, name, value.to(torch.float64))

            if self.mask is not None:
                mask = _as_tensor(self.mask, device)
                self.mask = mask if mask.dtype == torch.bool else mask != 0
                if tuple(self.mask.shape) not in (shape, frame):
                    raise ValueError("mask must match the data or one frame")

    @property
    def device(self):
        return self.data.device
