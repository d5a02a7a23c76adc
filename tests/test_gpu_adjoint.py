"""Adjoint of the cleaning map with the mask held fixed (GPU).

Checks, on noise-model frames from the single-frame generator:
  * the recorded donors equal the oracle's;
  * grad_data equals a CPU gradient built from the oracle's donor choices,
    summed in the same order (bit for bit);
  * central finite differences (h = 1e-6 |x|) on random tie-free pixels whose
    mask is unchanged by the perturbation agree to 1e-8 relative;
  * a scalar background's gradient is exactly zero, and finite differences of
    the loss in it are rounding noise only;
  * a background map's gradient (shared by the frames or one per frame)
    equals the CPU gradient from the oracle's donors bit for bit, and central
    differences on map pixels: -w at a flagged pixel, +w/2 or +w at a donor.
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


def cpu_background_gradient(masks, donors, weights, shared):
    """d loss / d bg from the oracle's donors: a flagged pixel p gives -w_p at
    p and its donor weights times w_p at the donors; frames are added in order
    when one map is shared."""
    grads = cpu_gradient(masks, donors, weights) - np.stack(weights)
    if not shared:
        return grads
    acc = grads[0].copy()
    for g in grads[1:]:
        acc = acc + g
    return acc


@pytest.mark.parametrize("shared", [True, False])
def test_background_map_gradient(cuda, shared):
    from cusmic_amd import autograd

    data, w = batch(n=2, shape=(80, 96), seed=21)
    rng = np.random.default_rng(21)
    bg_np = rng.uniform(50, 250, data.shape[1:] if shared else data.shape)
    say("background map gradient", "shared" if shared else "per frame")
    x = torch.tensor(data, device=cuda, requires_grad=True)
    bg = torch.tensor(bg_np, device=cuda, requires_grad=True)
    wt = torch.tensor(w, device=cuda)
    out, mask = autograd.remove_cosmics(x, 3, 5, 3, effective_gain=2.0, readnoise=5.0, background=bg)
    (wt * out).sum().backward()
    frames = [dict(data=data[f], background=bg_np if shared else bg_np[f]) for f in range(2)]
    wo, wm, wd, _ = oracle.clean_frames(frames, gain=2.0, readnoise=5.0, record_donors=True)
    np.testing.assert_array_equal(mask.cpu().numpy(), np.stack(wm))
    want_data = cpu_gradient(wm, wd, w)
    np.testing.assert_array_equal(x.grad.cpu().numpy().view(np.uint64), want_data.view(np.uint64))
    want = cpu_background_gradient(wm, wd, list(w), shared)
    got = bg.grad.cpu().numpy()
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
    flagged = np.stack(wm)
    assert flagged.sum() > 20
    # central differences on map pixels: flagged pixels, donors, others
    flat_bg = bg_np.ravel()
    per_frame = data[0].size
    picks = []
    for f in range(2):
        fl = np.flatnonzero(wm[f].ravel())
        dd = wd[f].reshape(-1, 2)[fl].ravel()
        base = 0 if shared else f * per_frame
        picks += [base + i for i in rng.choice(fl, min(20, fl.size), replace=False)]
        picks += [base + i for i in rng.choice(dd[dd >= 0], 20)]
    picks += list(rng.integers(0, flat_bg.size, 10))
    checked = kinds = 0
    for j in picks:
        hstep = 1e-6 * abs(flat_bg[j])
        up, dn = flat_bg.copy(), flat_bg.copy()
        up[j] += hstep
        dn[j] -= hstep
        ou, mu = loss_parts(data, w, cuda, background=up.reshape(bg_np.shape))
        od, md = loss_parts(data, w, cuda, background=dn.reshape(bg_np.shape))
        if not (np.array_equal(mu, flagged) and np.array_equal(md, flagged)):
            continue
        diff = (ou - od).ravel()
        changed = np.flatnonzero(diff)
        ratios = diff[changed] / (up[j] - dn[j])
        if not np.all(np.isclose(ratios, np.round(ratios * 2) / 2, rtol=0, atol=1e-6)):
            continue  # a donor order switched within the step: not a fixed point
        fd = float(np.sum(w.ravel()[changed] * diff[changed])) / (up[j] - dn[j])
        g = float(got.ravel()[j])
        # the rounding of (D + bg) - bg at the changed pixels is the only noise
        noise = 4 * float(np.sum(np.abs(w.ravel()[changed]) *
                                 np.spacing(np.abs(data.ravel()[changed]) + 300))) / (up[j] - dn[j])
        assert abs(fd - g) <= 1e-8 * max(1.0, abs(g)) + noise, (j, fd, g)
        checked += 1
        kinds += g != 0
    assert checked >= 60 and kinds >= 30


def wide_scene(kind, h=1024, w=1024, seed=3):
    """kind "hole": one hot pixel in a masked 41 x 41 square (its replacement
    window widens to r = 21); kind "grid": hot pixels every 9 px, each in its
    own masked 7 x 7 square (r = 4), about 25k widened donor targets;
    kind "plain": the hot pixel of "hole" without the square."""
    rng = np.random.default_rng(seed)
    d = rng.normal(1000, 10, size=(h, w))
    m = np.zeros((h, w), bool)
    if kind in ("hole", "plain"):
        cy, cx = 500, 600
        if kind == "hole":
            m[cy - 20:cy + 21, cx - 20:cx + 21] = True
            m[cy, cx] = False
        d[cy, cx] += 8000
    else:
        for cy in range(6, h - 6, 9):
            for cx in range(6, w - 6, 9):
                m[cy - 3:cy + 4, cx - 3:cx + 4] = True
                m[cy, cx] = False
                d[cy, cx] += 8000
    return d, m


@pytest.mark.parametrize("kind", ["hole", "grid"])
def test_widened_windows_gradient(cuda, kind):
    """Donors of widened windows (r = 21; 25k targets, more than the
    adjoint's first list holds) get their gradient bit for bit as the CPU
    gradient from the oracle's donors."""
    from cusmic_amd import autograd

    d, m = wide_scene(kind)
    say("widened windows", kind)
    w = np.random.default_rng(8).normal(size=d.shape)
    x = torch.tensor(d, device=cuda, requires_grad=True)
    out, mask = autograd.remove_cosmics(x, 3, 5, 3, effective_gain=2.0, readnoise=5.0,
                                        mask=torch.tensor(m, device=cuda))
    (torch.tensor(w, device=cuda) * out).sum().backward()
    wo, wm, wd, _ = oracle.clean_frames([dict(data=d, mask=m)], gain=2.0, readnoise=5.0,
                                        record_donors=True, threads=8)
    np.testing.assert_array_equal(mask.cpu().numpy(), wm[0])
    np.testing.assert_array_equal(out.detach().cpu().numpy().view(np.uint64), wo[0].view(np.uint64))
    want = cpu_gradient(wm, wd, w[None])[0]
    np.testing.assert_array_equal(x.grad.cpu().numpy().view(np.uint64), want.view(np.uint64))
    dd = wd[0].reshape(-1, 2)
    fl = np.flatnonzero(wm[0].ravel())
    r = np.maximum(abs(dd[fl, 0] % d.shape[1] - fl % d.shape[1]), abs(dd[fl, 0] // d.shape[1] - fl // d.shape[1]))
    assert r.max() == (21 if kind == "hole" else 4)
    if kind == "grid":
        assert (r > 2).sum() > 8192


def test_widened_window_cost(cuda):
    """The backward of a frame with one r = 21 window costs about what the
    same frame without it does."""
    import time

    from cusmic_amd import _lib, engine

    times = {}
    for kind in ("plain", "hole"):
        d, m = wide_scene(kind)
        fr = engine.Frames()
        fr.add(torch.tensor(d, device=cuda), mask=torch.tensor(m, device=cuda))
        _, mask, donors, _ = engine.run(fr, cuda, 3.0, 5.0, 3.0, 4, _lib.MIRROR, gain=2.0, readnoise=5.0,
                                        record_donors=True)
        g = torch.randn(d.size, dtype=torch.float64, device=cuda)
        ts = []
        for _ in range(15):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            engine.adjoint([d.shape], cuda, mask, donors, g)
            ts.append(time.perf_counter() - t0)
        times[kind] = float(np.median(ts))
    say("adjoint times", times)
    assert times["hole"] < 1.2 * times["plain"] + 3e-5, times
