"""Module-level entry points: the L.A.Cosmic call and a ragged-batch call.

``remove_cosmics`` keeps L.A.Cosmic 1.4.0's signature, as the reference's
does (mod/cusmic/api.py:20-63, checked by test/test_reference.py:46), so that
code written for the CPU package runs unchanged on tensors.
"""

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
    """Detect cosmic rays with L.A.Cosmic and replace them; returns
    ``(cleaned, crmask)``: new float64 and bool tensors on the data's GPU.

    Input: one frame of shape (height, width) or a stack (frames, height,
    width) of float64 pixels; the frames of a stack are independent. Host
    arrays are copied to the GPU, GPU tensors (views too) are read in place and
    never written.

    Noise: either ``error``, the 1-sigma uncertainty per pixel in data units
    (finite, > 0), or the noise model from ``effective_gain`` (electrons per
    data unit, > 0) and ``readnoise`` (electrons, >= 0); with ``error`` the
    other two are ignored. Maps have the data's shape or one frame's (shared
    by the stack); gain and read noise may be scalars.

    Detection: a pixel is a candidate when its significance exceeds
    ``cr_threshold`` and its ratio to the fine structure exceeds
    ``contrast``; neighbours of candidates join at ``cr_threshold`` and then
    at ``neighbor_threshold`` (all three in units of sigma, finite, >= 0).
    Up to ``maxiter`` detection and replacement passes run, fewer if a pass
    flags nothing new; ``maxiter = 0`` only applies the background
    arithmetic and needs no noise inputs.

    ``mask`` marks pixels that may neither seed a detection nor serve as a
    replacement donor (growth may still reach them). ``background`` (a
    finite scalar or map) is added before detection and subtracted from the
    result. Nonfinite pixels come back unchanged and are never flagged.
    ``border_mode`` extends the frame for the filters: "mirror", "reflect",
    "nearest", "wrap" or "constant" (zeros).

    The work is ordered on the data device's current stream; inputs written
    on another stream must be made ready with an event first. ValueError is
    raised for invalid input and when a replacement finds no finite,
    unmasked donor.
    """
    image = Image(data, error, mask, background, effective_gain, readnoise)
    return Cleaner(contrast, cr_threshold, neighbor_threshold, maxiter, border_mode)(image)


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
