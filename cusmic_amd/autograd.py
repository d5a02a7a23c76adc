"""Differentiable cleaning: d(cleaned)/d(image) and d/d(background).

With the detection mask held fixed, the cleaning map is

  out_p = (D_p + bg_p) - bg_p                         at an unflagged pixel p,
  out_p = median over donors j of (D_j + bg_j) - bg_p at a flagged pixel p,

where the median is the one middle donor (odd count) or the mean of the two
middle donors (even count) that the forward pass's last replacement chose.

* Data: the identity on unflagged pixels, plus each flagged pixel's upstream
  gradient sent to its donors with weight 1 or 1/2. This is the
  deterministic gather kernel behind cusmic_adjoint_device.
* Background map: bg_j enters through the donors exactly as D_j does, and
  -bg_p cancels the flagged pixel itself, so
  d loss / d bg = grad_data - grad_clean (0 at unflagged non-donors and at
  nonfinite pixels, whose output is the input), summed over frames when one
  map is shared by a stack.
* Scalar background: the sum of that map over all pixels, which is exactly 0
  (each flagged pixel's donor weights sum to 1), so it is returned as 0.
"""

import torch

from . import engine
from .cleaner import Cleaner
from .image import Image


def _sum_frames(g, shape):
    """Reduce a per-pixel gradient (nf, h, w) to a shared map of shape (h, w)
    by adding the frames in order (a fixed, reproducible order)."""
    if tuple(g.shape) == tuple(shape):
        return g
    acc = g[0].clone()
    for f in range(1, g.shape[0]):
        acc = acc + g[f]
    return acc.view(shape)


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
        ctx.set_materialize_grads(False)
        return clean, mask

    @staticmethod
    def backward(ctx, grad_clean, grad_mask):
        mask, donors = ctx.saved_tensors
        shape = ctx.stack_shape
        grad_data = grad_bg = None
        if grad_clean is None:  # the cleaned pixels do not reach the loss
            return None, None, None, None
        map_bg = ctx.needs_input_grad[1] and len(ctx.bg_shape) > 0
        if ctx.needs_input_grad[0] or map_bg:
            h, w = shape[-2:]
            nf = shape[0] if len(shape) == 3 else 1
            grad_clean = grad_clean.to(torch.float64).contiguous()
            grad_data = engine.adjoint([(h, w)] * nf, grad_clean.device,
                                       mask.reshape(-1).view(torch.uint8), donors.reshape(-1),
                                       grad_clean.reshape(-1)).view(shape)
        if ctx.needs_input_grad[1]:
            if map_bg:
                g = (grad_data - grad_clean).view(-1, *shape[-2:])
                grad_bg = _sum_frames(g, ctx.bg_shape)
            else:
                grad_bg = torch.zeros(ctx.bg_shape, dtype=torch.float64, device=grad_clean.device)
        if not ctx.needs_input_grad[0]:
            grad_data = None
        return grad_data, grad_bg, None, None


def remove_cosmics(data, contrast=3, cr_threshold=5, neighbor_threshold=3, error=None, mask=None,
                   background=None, effective_gain=None, readnoise=None, maxiter=4,
                   border_mode="mirror"):
    """:func:`cusmic_amd.remove_cosmics` as an autograd function of ``data``
    and of a tensor ``background`` (a scalar, a shared (h, w) map or a map
    per frame), with the detection mask held fixed at the forward pass's."""
    cleaner = Cleaner(contrast, cr_threshold, neighbor_threshold, maxiter, border_mode)
    if background is not None and not isinstance(background, torch.Tensor):
        background = torch.as_tensor(background, dtype=torch.float64, device=data.device)
    kwargs = dict(error=error, mask=mask, effective_gain=effective_gain, readnoise=readnoise)
    return _Clean.apply(data, background, cleaner, kwargs)


__all__ = ["remove_cosmics"]
