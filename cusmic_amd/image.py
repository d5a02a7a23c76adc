"""Frames and their calibration as torch tensors on a ROCm device.

The fields, the accepted shapes and the error messages are those of the
reference's ``Image`` (mod/cusmic/image.py:22-79), which its tests match on.
Host data (NumPy arrays, lists, CPU tensors) is copied to the GPU that holds
the data, or to the current GPU; GPU tensors are used in place, views
included.
"""

from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

Array = Any

# Per-field value rules, checked in this order after shape and finiteness:
# the error map and the gain must be positive, the read noise nonnegative.
_POSITIVE = ("error", "effective_gain")


def _target_device(data):
    if isinstance(data, torch.Tensor) and data.is_cuda:
        return data.device
    return torch.device("cuda", torch.cuda.current_device())


def _on_device(value, device):
    """value as a tensor on device (uploaded if it lives on the host)."""
    if not isinstance(value, torch.Tensor):
        value = torch.as_tensor(np.asarray(value))
    return value if value.device == device else value.to(device)


def _checked_calibration(name, value, device, shapes):
    """Validate one calibration field; returns it as a float64 tensor.

    A map must have the data's shape or one frame's (``shapes``); apart from
    the error map a field may also be a 0-d scalar. Maps live on the device;
    scalars stay on the host, so that neither their checks nor their use by
    the cleaner waits for the device.
    """
    t = value if isinstance(value, torch.Tensor) else torch.as_tensor(np.asarray(value))
    t = t.detach().cpu() if t.dim() == 0 else _on_device(t, device)
    if t.dtype == torch.bool or t.is_complex():
        raise TypeError(f"{name} must be real numeric data")
    scalar_ok = name != "error" and t.dim() == 0
    if tuple(t.shape) not in shapes and not scalar_ok:
        raise ValueError(f"{name} has an incompatible shape")
    if not bool(torch.isfinite(t).all()):
        raise ValueError(f"{name} must be finite")
    if name in _POSITIVE and bool((t <= 0).any()):
        raise ValueError(f"{name} must be positive")
    if name == "readnoise" and bool((t < 0).any()):
        raise ValueError("readnoise must be nonnegative")
    return t.to(torch.float64)


@dataclass
class Image:
    """float64 pixels (one frame or a stack) plus what the cleaner needs.

    ``error`` is a 1-sigma map; without it the noise model uses
    ``effective_gain`` and ``readnoise``, maps or scalars. ``background`` is
    a map or a scalar, ``mask`` a boolean map of pixels that never seed a
    detection nor donate a replacement. Every map has the data's shape or the
    shape of one frame, which all frames of a stack then share. Validation
    runs once, at construction, on the data's GPU.
    """

    data: Array
    error: Array | None = None
    mask: Array | None = None
    background: Array | float | None = None
    effective_gain: Array | float | None = None
    readnoise: Array | float | None = None

    def __post_init__(self):
        device = _target_device(self.data)
        if not (isinstance(self.data, torch.Tensor) and self.data.is_cuda):
            self.data = _on_device(self.data, device)
        if self.data.dtype != torch.float64:
            raise TypeError("data must contain float64 pixels")
        if self.data.dim() not in (2, 3) or self.data.numel() == 0:
            raise ValueError("data must be a nonempty 2D frame or 3D stack")
        shapes = (tuple(self.data.shape), tuple(self.data.shape[-2:]))
        # an error map replaces the noise model, whose inputs are then not
        # examined
        noise_fields = ["error"] if self.error is not None else ["effective_gain", "readnoise"]
        with torch.cuda.device(device):
            for name in noise_fields + ["background"]:
                value = getattr(self, name)
                if value is not None:
                    setattr(self, name, _checked_calibration(name, value, device, shapes))
            if self.mask is not None:
                m = _on_device(self.mask, device)
                m = m if m.dtype == torch.bool else m != 0
                if tuple(m.shape) not in shapes:
                    raise ValueError("mask must match the data or one frame")
                self.mask = m

    @property
    def device(self):
        return self.data.device
