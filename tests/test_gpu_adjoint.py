"""Adjoint of the cleaning map with the mask held fixed (GPU).

Checks, on noise-model frames from the single-frame generator:
  * the recorded donors equal the oracle's;
  * grad_data equals a CPU gradient built from the oracle's donor choices,
    summed in the same order (bit for bit);
  * central finite differences (h = 1e-6 |x|) on random tie-free pixels whose
    mask is unchanged by the perturbation agree to 1e-8 relative;
  * the background gradient is exactly zero, and finite differences of the
    loss in the background are rounding noise only.
"""

import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def say(*args):
    print("[case]", *args, flush=True)


def cpu_gradient(masks, donors, weights):
    grads = []
    for m, d, w in zip(masks, donors, weights):
        flat_m, flat_w = m.ravel(), w.ravel()
        g = np.where(flat_m, 0.0, flat_w)
        dd = d.reshape(-1, 2)
        for p in np.flatnonzero(flat_m):
            lo, hi = dd[p]
            if lo == hi:
                g[lo] += flat_w[p]
            else:
                g[lo] += 0.5 * flat_w[p]
                g[hi] += 0.5 * flat_w[p]
        grads.append(g.reshape(m.shape))
    return np.stack(grads)


def batch(n=3, shape=(160, 192), seed=40):
    from cusmic_amd import scenes

    frames = [scenes.frame(shape, seed + 7 * f, "cpu", n_stars=8, n_tracks=40, max_len=5)[0].numpy()
              for f in range(n)]
    w = np.random.default_rng(seed).normal(size=(n, *shape))
    return np.stack(frames), w


def test_gradient_matches_cpu_donors(cuda):
    from cusmic_amd import autograd

    data, w = batch()
    say("adjoint vs oracle donors")
    x = torch.tensor(data, device=cuda, requires_grad=True)
    out, mask = autograd.remove_cosmics(x, 3, 5, 3, effective_gain=2.0, readnoise=5.0)
    (torch.tensor(w, device=cuda) * out).sum().backward()
    want_o, want_m, want_d, _ = oracle.clean_frames([dict(data=f) for f in data], gain=2.0, readnoise=5.0,
                                                    record_donors=True)
    np.testing.assert_array_equal(mask.cpu().numpy(), np.stack(want_m))
    np.testing.assert_array_equal(out.detach().cpu().numpy().view(np.uint64),
                                  np.stack(want_o).view(np.uint64))
    grad = x.grad.cpu().numpy()
    want = cpu_gradient(want_m, want_d, w)
    np.testing.assert_array_equal(grad.view(np.uint64), want.view(np.uint64))
    assert np.stack(want_m).sum() > 50


def loss_parts(x, w, cuda, background=None):
    from cusmic_amd import autograd

    out, mask = autograd.remove_cosmics(torch.tensor(x, device=cuda), 3, 5, 3, effective_gain=2.0,
                                        readnoise=5.0, background=background)
    return out.detach().cpu().numpy(), mask.cpu().numpy()


def test_finite_differences(cuda):
    from cusmic_amd import autograd

    data, w = batch(n=2, shape=(96, 128), seed=77)
    x = torch.tensor(data, device=cuda, requires_grad=True)
    out, mask = autograd.remove_cosmics(x, 3, 5, 3, effective_gain=2.0, readnoise=5.0)
    (torch.tensor(w, device=cuda) * out).sum().backward()
    grad = x.grad.cpu().numpy()
    base_out = out.detach().cpu().numpy()
    base_mask = mask.cpu().numpy()
    rng = np.random.default_rng(5)
    flat = data.ravel()
    # candidate pixels: half donors of flagged pixels, half arbitrary
    _, _, donors, _ = oracle.clean_frames([dict(data=f) for f in data], gain=2.0, readnoise=5.0,
                                          record_donors=True)
    donor_idx = []
    for f, d in enumerate(donors):
        dd = d.reshape(-1, 2)
        sel = dd[base_mask[f].ravel()].ravel()
        donor_idx += list(f * data[0].size + sel[sel >= 0])
    picks = list(rng.choice(donor_idx, 128)) + list(rng.integers(0, flat.size, 128))
    checked = 0
    for q in picks:
        say("fd pixel", int(q))
        v = flat[q]
        h = 1e-6 * abs(v)
        up, dn = flat.copy(), flat.copy()
        up[q], dn[q] = v + h, v - h
        ou, mu = loss_parts(up.reshape(data.shape), w, cuda)
        od, md = loss_parts(dn.reshape(data.shape), w, cuda)
        if not (np.array_equal(mu, base_mask) and np.array_equal(md, base_mask)):
            continue  # the perturbation moved a detection: not a fixed-mask point
        # tie-free: no other donor value within the perturbation of this pixel
        diff = (ou - od).ravel()
        changed = np.flatnonzero(diff)
        fd = float(np.sum(w.ravel()[changed] * diff[changed])) / (up[q] - dn[q])
        # skip pixels where the median switched donors (a tie within +-h)
        ratios = diff[changed] / (up[q] - dn[q])
        if not np.all(np.isclose(ratios, np.round(ratios * 2) / 2, rtol=0, atol=1e-9)):
            continue
        g = grad.ravel()[q]
        assert abs(fd - g) <= 1e-8 * max(1.0, abs(g)), (q, fd, g)
        checked += 1
    assert checked >= 200


def test_background_gradient_zero(cuda):
    from cusmic_amd import autograd

    data, w = batch(n=2, shape=(64, 64), seed=9)
    x = torch.tensor(data, device=cuda, requires_grad=True)
    bg = torch.tensor(150.0, dtype=torch.float64, device=cuda, requires_grad=True)
    out, _ = autograd.remove_cosmics(x, 3, 5, 3, effective_gain=2.0, readnoise=5.0, background=bg)
    (torch.tensor(w, device=cuda) * out).sum().backward()
    assert float(bg.grad) == 0.0
    h = 1e-6 * 150
    ou, _ = loss_parts(data, w, cuda, background=150.0 + h)
    od, _ = loss_parts(data, w, cuda, background=150.0 - h)
    fd = float(np.sum(w * (ou - od))) / (2 * h)
    # pure rounding: (D + bg) - bg moves by at most an ulp of D + bg per pixel
    bound = np.sum(np.abs(w) * np.spacing(np.abs(data) + 150)) / (2 * h)
    assert abs(fd) <= bound
