"""The L.A.Cosmic call surface (mirrors mod/cusmic/api.py:20-63)."""

import math

import torch

from . import _lib, engine
from .cleaner import Cleaner
from .image import Image


def remove_cosmics(
    data,
    contrast,
    cr_threshold,
    neighbor_threshold,
    error=None,
    mask=None,
    background=None,
    effective_gain=None,
    readnoise=None,
    maxiter=4,
    border_mode="mirror",
):
    """Return cleaned float64 pixels and a boolean cosmic-ray mask on the GPU.

    ``data`` is a nonempty float64 frame (height, width) or stack
    (frames, height, width). NumPy inputs are uploaded; GPU tensors are
    borrowed. Frames are cleaned independently and outputs do not alias inputs.

    Supply a finite, positive 1-sigma ``error`` map in image units, or both
    ``effective_gain`` (electrons per image unit, positive) and ``readnoise``
    (electrons, nonnegative). Error overrides gain and read noise. Maps may
    match data or one shared frame; gain and read noise also accept scalars.

    ``contrast`` rejects fine structure; ``cr_threshold`` and
    ``neighbor_threshold`` set detection and growth thresholds in sigma.
    All three must be finite and nonnegative. ``mask`` excludes seed and
    donor pixels, but neighboring detections may grow into masked pixels.
    Nonfinite input pixels are preserved and never flagged.

    ``background`` is a finite scalar or map in image units, added before
    cleaning and subtracted afterward. ``maxiter`` is nonnegative; zero
    disables detection and needs no noise model, but still applies background
    arithmetic. Border modes are mirror, reflect, nearest, wrap, and constant
    (zero padding).

    Work uses the input device's current stream. A replacement with no finite,
    unmasked donors raises ValueError.
    """
    image = Image(data, error, mask, background, effective_gain, readnoise)
    clean = Cleaner(contrast, cr_threshold, neighbor_threshold, maxiter, border_mode)
    return clean(image)


def remove_cosmics_ragged(frames, contrast=3, cr_threshold=5, neighbor_threshold=3, maxiter=4,
                          border_mode="mirror", frame_gain=None, frame_readnoise=None,
                          record_donors=False, laplacian="exact"):
    """Clean a list of frames of individual shapes in one call.

    ``frames`` holds dicts with a float64 (h, w) GPU tensor ``data`` and
    optional (h, w) ``error``, ``gain``, ``readnoise``, ``background``,
    ``mask``. ``frame_gain``/``frame_readnoise`` give per-frame scalars.
    Returns (list of clean frames, list of bool masks, donors or None, passes).
    """
    Cleaner(contrast, cr_threshold, neighbor_threshold, maxiter, border_mode)  # validation
    fr = engine.Frames()
    device = frames[0]["data"].device
    for item in frames:
        fr.add(item["data"], item.get("error"), item.get("gain"), item.get("readnoise"),
               item.get("background"), item.get("mask"))
    clean, mask, donors, passes = engine.run(
        fr, device, float(contrast), float(cr_threshold), float(neighbor_threshold), int(maxiter),
        _lib.MODES[border_mode], frame_gain=frame_gain, frame_readnoise=frame_readnoise,
        readnoise=0.0, gain=math.nan, record_donors=record_donors,
        laplacian=_lib.LAPLACIAN_EXACT if laplacian == "exact" else _lib.LAPLACIAN_CLOSE)
    outs, masks, dons = [], [], []
    off = 0
    for h, w in fr.shapes:
        outs.append(clean[off:off + h * w].view(h, w))
        masks.append(mask[off:off + h * w].view(torch.bool).view(h, w))
        if donors is not None:
            dons.append(donors[2 * off:2 * (off + h * w)].view(h, w, 2))
        off += h * w
    return outs, masks, (dons if record_donors else None), passes
