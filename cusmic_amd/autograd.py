"""Differentiable cleaning: d(cleaned)/d(image) and d/d(background).

The cleaning map with the detection mask held fixed is
  out = (D + bg) - bg                     on unflagged pixels,
  out = median of donor values (+bg) - bg on flagged pixels,
so its adjoint is the identity on unflagged pixels plus, for each flagged
pixel, its upstream gradient sent to the one (odd donor count) or two (even
count, weight 1/2 each) median donors recorded by the forward pass. The
background cancels exactly everywhere, so its gradient is exactly zero.
The backward is the deterministic gather kernel of cusmic_adjoint_device.
"""

import torch

from . import engine
from .cleaner import Cleaner
from .image import Image


class _Clean(torch.autograd.Function):
    @staticmethod
    def forward(ctx, data, background, cleaner, kwargs):
        image = Image(data.detach(), background=None if background is None else background.detach(),
                      **kwargs)
        clean, mask, donors, _ = cleaner.run(image, record_donors=True)
        ctx.save_for_backward(mask, donors)
        ctx.stack_shape = data.shape
        ctx.bg_shape = None if background is None else background.shape
        ctx.mark_non_differentiable(mask)
        return clean, mask

    @staticmethod
    def backward(ctx, grad_clean, grad_mask):
        mask, donors = ctx.saved_tensors
        shape = ctx.stack_shape
        grad_data = grad_bg = None
        if ctx.needs_input_grad[0]:
            h, w = shape[-2:]
            nf = shape[0] if len(shape) == 3 else 1
            grad_data = engine.adjoint([(h, w)] * nf, grad_clean.device,
                                       mask.reshape(-1).view(torch.uint8), donors.reshape(-1),
                                       grad_clean.reshape(-1)).view(shape)
        if ctx.needs_input_grad[1]:
            grad_bg = torch.zeros(ctx.bg_shape, dtype=torch.float64, device=grad_clean.device)
        return grad_data, grad_bg, None, None


def remove_cosmics(data, contrast=3, cr_threshold=5, neighbor_threshold=3, error=None, mask=None,
                   background=None, effective_gain=None, readnoise=None, maxiter=4,
                   border_mode="mirror"):
    """Like :func:`cusmic_amd.remove_cosmics`, differentiable in ``data`` and a
    tensor ``background`` (scalar or map) with the detection mask held fixed."""
    cleaner = Cleaner(contrast, cr_threshold, neighbor_threshold, maxiter, border_mode)
    if background is not None and not isinstance(background, torch.Tensor):
        background = torch.as_tensor(background, dtype=torch.float64, device=data.device)
    kwargs = dict(error=error, mask=mask, effective_gain=effective_gain, readnoise=readnoise)
    return _Clean.apply(data, background, cleaner, kwargs)


__all__ = ["remove_cosmics"]
