"""The benchmark configurations at their full sizes, HIP path against the oracle (GPU).

  C0  one 512x512 frame, noise model (gain 2, read noise 5), settings (3, 5, 3),
      through remove_cosmics (the Python surface);
  C1  a stack of 16 512x512 frames with one shared error map, a mask of two bad
      columns per frame and 0.01% NaN pixels, through Image / Cleaner;
  C2  one 4096x4096 frame with a per-pixel error map, scalar background 150
      and 0.5% masked pixels, through remove_cosmics_device (the C ABI);
  C3  the 64-frame ragged batch: one remove_cosmics_ragged_device call equals
      64 plain calls bit for bit, and one frame of each shape equals the
      oracle;
  C4  8 frames of 1024x1024 (noise model): forward with donors, then the
      adjoint of sum(w * cleaned), against the oracle's donors and by central
      differences on 256 tie-free pixels.

Masks must be identical and pixels bit-identical. The scenes are generated
with torch on the GPU from fixed seeds and copied to the host for the oracle.
"""

import ctypes
import os

import numpy as np
import pytest
import torch

import oracle
from conftest import assert_same_pixels

pytestmark = pytest.mark.gpu


def say(*args):
    print("[case]", *args, flush=True)


def threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def host(t):
    return None if t is None else t.detach().cpu().numpy()


def test_c0_single_frame_noise_model(cuda):
    from cusmic_amd import remove_cosmics, scenes

    say("C0: 512x512 noise model")
    sc = scenes.single_frame_512(cuda)
    clean, mask = remove_cosmics(sc["data"], 3, 5, 3, effective_gain=sc["gain"], readnoise=sc["readnoise"])
    d = host(sc["data"])
    want, wmask, _, passes = oracle.clean_frames([dict(data=d)], gain=2.0, readnoise=5.0,
                                                 threads=threads())
    np.testing.assert_array_equal(host(mask), wmask[0])
    assert_same_pixels(host(clean), want[0])
    assert wmask[0].sum() > 300, wmask[0].sum()


def test_c1_stack_shared_error_mask_nan(cuda):
    from cusmic_amd import Cleaner, Image, scenes

    say("C1: 16 x 512x512 stack")
    sc = scenes.stack_512(cuda)
    image = Image(sc["data"], error=sc["error"], mask=sc["mask"])
    clean, mask = Cleaner(3, 5, 3)(image)
    d, e, m = host(sc["data"]), host(sc["error"]), host(sc["mask"])
    assert np.isnan(d).sum() > 16 * 10
    frames = [dict(data=d[f], error=e, mask=m[f]) for f in range(d.shape[0])]
    want, wmask, _, _ = oracle.clean_frames(frames, threads=threads())
    got_c, got_m = host(clean), host(mask)
    for f in range(d.shape[0]):
        np.testing.assert_array_equal(got_m[f], wmask[f], err_msg=f"mask of frame {f}")
        assert_same_pixels(got_c[f], want[f])
    assert sum(w.sum() for w in wmask) > 16 * 300


def test_c2_detector_frame_device_abi(cuda):
    from cusmic_amd import _lib, scenes

    say("C2: 4096x4096, error map, background 150, 0.5% mask, remove_cosmics_device")
    sc = scenes.detector_frame(cuda)
    data, error = sc["data"].contiguous(), sc["error"].contiguous()
    mask_in = sc["mask"].to(torch.uint8).contiguous()
    h, w = data.shape
    clean = torch.empty(h * w, dtype=torch.float64, device=cuda)
    mask = torch.empty(h * w, dtype=torch.uint8, device=cuda)
    img = _lib.CusmicImage(data.data_ptr(), error.data_ptr(), None, None, None, mask_in.data_ptr(), w, h)
    o = _lib.default_options()
    o.background = sc["background"]
    msg = ctypes.create_string_buffer(256)
    stream = torch.cuda.current_stream().cuda_stream
    st = _lib.lib().remove_cosmics_device(ctypes.byref(img), ctypes.byref(o), ctypes.c_void_p(clean.data_ptr()),
                                          ctypes.c_void_p(mask.data_ptr()), ctypes.c_void_p(stream), msg, 256)
    assert st == _lib.OK, msg.value
    say("C2: oracle")
    want, wmask, _, passes = oracle.clean_frames(
        [dict(data=host(data), error=host(error), mask=host(mask_in))], background=150.0,
        threads=threads())
    np.testing.assert_array_equal(host(mask).reshape(h, w).astype(bool), wmask[0])
    assert_same_pixels(host(clean).reshape(h, w), want[0])
    assert wmask[0].sum() > 20000, wmask[0].sum()


def test_c3_ragged_batch(cuda):
    from cusmic_amd import _lib, engine, scenes

    say("C3: ragged batch of 64 frames")
    batch = scenes.ragged_batch(cuda)
    fr = engine.Frames()
    for item in batch:
        fr.add(item["data"], error=item["error"])
    clean, mask, _, passes = engine.run(fr, cuda, 3.0, 5.0, 3.0, 4, _lib.MIRROR)
    off = 0
    firsts = {}
    for f, item in enumerate(batch):
        h, w = item["data"].shape
        f1 = engine.Frames()
        f1.add(item["data"], error=item["error"])
        c1, m1, _, p1 = engine.run(f1, cuda, 3.0, 5.0, 3.0, 4, _lib.MIRROR)
        assert torch.equal(c1.view(torch.int64), clean[off:off + h * w].view(torch.int64)), f"pixels of frame {f}"
        assert torch.equal(m1, mask[off:off + h * w]), f"mask of frame {f}"
        assert p1[0] == passes[f]
        firsts.setdefault((h, w), (f, off))
        off += h * w
    assert len(firsts) == len(scenes.RAGGED_SHAPES)
    for (h, w), (f, o) in sorted(firsts.items(), key=lambda kv: kv[0][0] * kv[0][1]):
        say("C3: oracle on frame", f, f"{h}x{w}")
        item = batch[f]
        want, wmask, _, wp = oracle.clean_frames([dict(data=host(item["data"]), error=host(item["error"]))],
                                                 threads=threads())
        np.testing.assert_array_equal(host(mask[o:o + h * w]).reshape(h, w).astype(bool), wmask[0])
        assert_same_pixels(host(clean[o:o + h * w]).reshape(h, w), want[0])
        assert wp[0] == passes[f]


def cpu_gradient(masks, donors, weights):
    """grad_data from the oracle's donors: the identity on unflagged pixels,
    then each flagged pixel's weight added to its donors in raster order."""
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


def test_c4_forward_backward(cuda):
    from cusmic_amd import autograd, scenes

    say("C4: 8 x 1024x1024 forward + backward")
    sc = scenes.adjoint_batch(cuda)
    data, wt = sc["data"], sc["weights"]
    x = data.clone().requires_grad_(True)
    out, mask = autograd.remove_cosmics(x, 3, 5, 3, effective_gain=2.0, readnoise=5.0)
    (wt * out).sum().backward()
    d, w = host(data), host(wt)
    say("C4: oracle with donors")
    want, wmask, wdon, _ = oracle.clean_frames([dict(data=f) for f in d], gain=2.0, readnoise=5.0,
                                               record_donors=True, threads=threads())
    base_mask = host(mask)
    np.testing.assert_array_equal(base_mask, np.stack(wmask))
    assert_same_pixels(host(out), np.stack(want))
    grad = host(x.grad)
    np.testing.assert_array_equal(grad.view(np.uint64), cpu_gradient(wmask, wdon, w).view(np.uint64))
    assert base_mask.sum() > 8 * 250

    say("C4: central differences")
    rng = np.random.default_rng(4)
    n_frame = d[0].size
    donor_idx = []
    for f, dn in enumerate(wdon):
        sel = dn.reshape(-1, 2)[wmask[f].ravel()].ravel()
        donor_idx += list(f * n_frame + sel[sel >= 0])
    picks = list(rng.choice(donor_idx, 256, replace=False)) + list(rng.integers(0, d.size, 256))
    flat = data.reshape(-1)
    base_m = mask.reshape(-1)
    wf = wt.reshape(-1)
    checked = 0
    for q in picks:
        v = float(flat[q])
        hstep = 1e-6 * abs(v)
        up, dn = flat.clone(), flat.clone()
        up[q] = v + hstep
        dn[q] = v - hstep
        ou, mu = autograd.remove_cosmics(up.view(data.shape), 3, 5, 3, effective_gain=2.0, readnoise=5.0)
        od, md = autograd.remove_cosmics(dn.view(data.shape), 3, 5, 3, effective_gain=2.0, readnoise=5.0)
        if not (torch.equal(mu.reshape(-1), base_m) and torch.equal(md.reshape(-1), base_m)):
            continue  # the step moved a detection: not a fixed-mask point
        diff = (ou - od).reshape(-1)
        span = float(up[q] - dn[q])
        changed = torch.nonzero(diff).view(-1)
        ratios = (diff[changed] / span).cpu().numpy()
        if not np.all(np.isclose(ratios, np.round(ratios * 2) / 2, rtol=0, atol=1e-9)):
            continue  # a donor order switched within the step (a tie)
        fd = float((wf[changed] * diff[changed]).sum()) / span
        g = float(grad.reshape(-1)[q])
        assert abs(fd - g) <= 1e-8 * max(1.0, abs(g)), (int(q), fd, g)
        checked += 1
        if checked == 256:
            break
    assert checked == 256, checked
