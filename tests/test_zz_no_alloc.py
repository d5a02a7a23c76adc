"""remove_cosmics_batch_device_noalloc against remove_cosmics_batch_device (GPU).

The entry point without device allocation must give the same masks and
pixels, bit for bit, as the allocating batch entry point, on consecutive
calls that share its static scratch buffer (the second with a smaller
batch, so stale contents of the first would show), and must reject a batch
that needs more scratch than the buffer holds before doing any work.
"""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*args):
    print("[case]", *args, flush=True)


def run(name, frames, errors, o, cuda):
    """Call the batch entry point `name` on device frames; returns (status,
    message, clean, mask) with the outputs as host arrays."""
    from cusmic_amd import _lib

    nf = len(frames)
    h, w = frames[0].shape
    clean = torch.full((nf * h * w,), 7.0, dtype=torch.float64, device=cuda)
    mask = torch.full((nf * h * w,), 9, dtype=torch.uint8, device=cuda)
    ims = (_lib.CusmicImage * nf)(*[
        _lib.CusmicImage(frames[f].data_ptr(), errors[f].data_ptr() if errors[f] is not None else None,
                         None, None, None, None, w, h) for f in range(nf)])
    msg = ctypes.create_string_buffer(256)
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    st = getattr(_lib.lib(), name)(ims, nf, ctypes.byref(o), ctypes.c_void_p(clean.data_ptr()),
                                   ctypes.c_void_p(mask.data_ptr()), ctypes.c_void_p(stream), msg, 256)
    return st, msg.value.decode(), clean.cpu().numpy(), mask.cpu().numpy()


def scene_batch(cuda, nf, shape, seed, error=True):
    from cusmic_amd import scenes

    frames, errors = [], []
    for f in range(nf):
        d, sig = scenes.frame(shape, seed + 10 * f, cuda, n_stars=20, n_tracks=80, max_len=8)
        frames.append(d.contiguous())
        errors.append(scenes.error_map(sig, 2.0, 5.0).contiguous() if error else None)
    return frames, errors


def test_noalloc_equals_batch_device(cuda):
    from cusmic_amd import _lib

    o = _lib.default_options()
    calls = [
        ("3 x 512x512, error maps", scene_batch(cuda, 3, (512, 512), 1), None),
        ("2 x 200x300, noise model", scene_batch(cuda, 2, (200, 300), 2, error=False), (2.0, 5.0)),
        ("1 x 64x96, error map, maxiter 9", scene_batch(cuda, 1, (64, 96), 3), None),
    ]
    for label, (frames, errors), model in calls:
        say("no-alloc call:", label)
        o = _lib.default_options()
        if model:
            o.gain, o.readnoise = model
        if "maxiter 9" in label:
            o.maxiter, o.contrast, o.cr_threshold, o.neighbor_threshold = 9, 1.0, 3.0, 2.0
        st, msg, clean, mask = run("remove_cosmics_batch_device_noalloc", frames, errors, o, cuda)
        assert st == _lib.OK, msg
        st2, msg2, want, wmask = run("remove_cosmics_batch_device", frames, errors, o, cuda)
        assert st2 == _lib.OK, msg2
        np.testing.assert_array_equal(mask, wmask)
        np.testing.assert_array_equal(clean.view(np.uint64), want.view(np.uint64))
        assert mask.sum() > 10


def test_noalloc_rejects_oversized_batch(cuda):
    from cusmic_amd import _lib

    say("no-alloc call beyond the buffer")
    d = torch.full((2048, 2048), 10.0, dtype=torch.float64, device=cuda)
    e = torch.ones_like(d)
    o = _lib.default_options()
    st, msg, clean, mask = run("remove_cosmics_batch_device_noalloc", [d], [e], o, cuda)
    assert st == _lib.INVALID
    assert "more than the" in msg and "static buffer" in msg, msg
    assert (clean == 7.0).all() and (mask == 9).all()
