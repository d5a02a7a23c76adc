"""The reference's device entry points through the C ABI (GPU).

Restates test/test_batch.cu:80-111: three same-shape frames that share one
error pointer, cleaned by remove_cosmics_batch_device on a caller stream after
an event wait on a delayed producer, then frame by frame through
remove_cosmics_device; the two must agree bit for bit and flag the marked
rays. Also: the batch's same-shape and data rules (src/api.cu:47-81), the
NULL stream, and the closed-form noise-model scenes on the GPU.
"""

import ctypes
import math

import numpy as np
import pytest
import torch

import oracle
from conftest import assert_same_pixels
from noise_scenes import plateau, plateau_medians, noise, raised_pixels

pytestmark = pytest.mark.gpu


def say(*args):
    print("[case]", *args, flush=True)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def image(data, error=None, w=None, h=None):
    from cusmic_amd import _lib

    return _lib.CusmicImage(data.data_ptr() if data is not None else None,
                            error.data_ptr() if error is not None else None, None, None, None, None,
                            w, h)


def batch_call(ims, o, clean, mask, stream):
    from cusmic_amd import _lib

    arr = (_lib.CusmicImage * len(ims))(*ims)
    msg = ctypes.create_string_buffer(256)
    st = _lib.lib().remove_cosmics_batch_device(arr, len(ims), ctypes.byref(o), ptr(clean), ptr(mask),
                                                ctypes.c_void_p(stream), msg, 256)
    return st, msg.value.decode()


def single_call(im, o, clean, mask, stream):
    from cusmic_amd import _lib

    msg = ctypes.create_string_buffer(256)
    st = _lib.lib().remove_cosmics_device(ctypes.byref(im), ctypes.byref(o), ptr(clean), ptr(mask),
                                          ctypes.c_void_p(stream), msg, 256)
    return st, msg.value.decode()


def test_batch_equals_single_frames_on_caller_stream(cuda):
    from cusmic_amd import _lib

    say("batch device vs single device, shared error pointer, caller stream")
    w = h = 9
    n, nf = w * h, 3
    producer, consumer = torch.cuda.Stream(), torch.cuda.Stream()
    ready = torch.cuda.Event()
    data = torch.empty(nf * n, dtype=torch.float64, device=cuda)
    error = torch.ones(n, dtype=torch.float64, device=cuda)
    batch = torch.full((nf * n,), -1.0, dtype=torch.float64, device=cuda)
    flags = torch.full((nf * n,), 7, dtype=torch.uint8, device=cuda)
    single = torch.empty(n, dtype=torch.float64, device=cuda)
    one_mask = torch.empty(n, dtype=torch.uint8, device=cuda)
    torch.cuda.synchronize()
    with torch.cuda.stream(producer):
        # the upload is late: the cleaner must wait for the event, not the data
        torch.cuda._sleep(5_000_000)
        for f in range(nf):
            data[f * n:(f + 1) * n] = 10 + f
            data[f * n + (4 + f % 2) * w + 4 + f // 2] = 1000 + 100 * f
        ready.record(producer)
    consumer.wait_event(ready)
    ims = [image(data[f * n:], error, w, h) for f in range(nf)]
    o = _lib.default_options()
    st, msg = batch_call(ims, o, batch, flags, consumer.cuda_stream)
    assert st == _lib.OK, msg
    for f in range(nf):
        st, msg = single_call(ims[f], o, single, one_mask, consumer.cuda_stream)
        assert st == _lib.OK, msg
        got, got_m = batch[f * n:(f + 1) * n].cpu().numpy(), flags[f * n:(f + 1) * n].cpu().numpy()
        assert_same_pixels(got, single.cpu().numpy())
        np.testing.assert_array_equal(got_m, one_mask.cpu().numpy())
        ray = (4 + f % 2) * w + 4 + f // 2
        assert one_mask[ray] == 1 and int(one_mask.sum()) == 1
        np.testing.assert_array_equal(got, 10 + f)
    # the oracle agrees on the same frames
    host = data.cpu().numpy().reshape(nf, h, w)
    want, wmask, _, _ = oracle.clean_frames([dict(data=host[f], error=np.ones((h, w)))
                                             for f in range(nf)])
    assert_same_pixels(batch.cpu().numpy(), np.concatenate([x.ravel() for x in want]))
    np.testing.assert_array_equal(flags.cpu().numpy().astype(bool),
                                  np.concatenate([m.ravel() for m in wmask]))


def test_batch_rules(cuda):
    """src/api.cu:47-81: frames of a batch need data and one shape."""
    from cusmic_amd import _lib

    say("batch shape rules")
    d = torch.full((2, 9, 9), 10.0, dtype=torch.float64, device=cuda)
    err = torch.ones((9, 9), dtype=torch.float64, device=cuda)
    clean = torch.full((2 * 81,), 3.0, dtype=torch.float64, device=cuda)
    mask = torch.zeros(2 * 81, dtype=torch.uint8, device=cuda)
    o = _lib.default_options()
    stream = torch.cuda.current_stream().cuda_stream
    st, msg = batch_call([image(d[0], err, 9, 9), image(d[1], err, 9, 8)], o, clean, mask, stream)
    assert st == _lib.INVALID and msg == "frames must have data and the same shape"
    st, msg = batch_call([image(d[0], err, 9, 9), image(None, err, 9, 9)], o, clean, mask, stream)
    assert st == _lib.INVALID and msg == "frames must have data and the same shape"
    st, msg = batch_call([], o, clean, mask, stream)
    assert st == _lib.INVALID and msg == "invalid images, batch size or output buffers"
    o.border = 5
    st, msg = batch_call([image(d[0], err, 9, 9)], o, clean, mask, stream)
    assert st == _lib.INVALID and msg == "invalid detection settings"
    # rejected before any work: the device outputs are untouched
    torch.cuda.synchronize()
    assert bool((clean == 3.0).all()) and not bool(mask.any())


def test_null_stream(cuda):
    """stream = NULL: the default stream, and still complete on return."""
    from cusmic_amd import _lib

    say("NULL stream")
    rng = np.random.default_rng(4)
    host = rng.normal(100, 3, size=(40, 37))
    host[10, 10] += 900
    host[25, 30] += 1500
    d = torch.tensor(host, device=cuda)
    err = torch.full_like(d, 3.0)
    clean = torch.empty(host.size, dtype=torch.float64, device=cuda)
    mask = torch.empty(host.size, dtype=torch.uint8, device=cuda)
    torch.cuda.synchronize()
    o = _lib.default_options()
    st, msg = single_call(image(d, err, 37, 40), o, clean, mask, 0)
    assert st == _lib.OK, msg
    want, wmask, _, _ = oracle.clean_frames([dict(data=host, error=np.full_like(host, 3.0))])
    assert_same_pixels(clean.cpu().numpy().reshape(40, 37), want[0])
    np.testing.assert_array_equal(mask.cpu().numpy().reshape(40, 37).astype(bool), wmask[0])
    assert wmask[0].sum() >= 2


def test_noise_model_closed_form(cuda):
    """The closed-form noise-model scenes: the GPU flags what the published
    formula predicts (and what the oracle flags)."""
    from cusmic_amd import remove_cosmics_ragged

    T = 5.0
    rng = np.random.default_rng(5)
    shape = (49, 53)
    gmap = rng.uniform(1.0, 4.0, shape)
    rnmap = rng.uniform(2.0, 10.0, shape)
    cases = []
    for g, rn, c in ((2.0, 5.0, 1000.0), (1.0, 0.0, 0.0), (4.0, 10.0, -30.0), (0.5, 3.0, 1e-6)):
        data, want = raised_pixels((40, 44), c, g, rn, T, seed=3)
        cases.append((dict(data=data), g, rn, want))
    data, want = raised_pixels(shape, 800.0, gmap, rnmap, T, seed=7)
    cases.append((dict(data=data, gain=gmap, readnoise=rnmap), float("nan"), 0.0, want))
    for kind in "XY":
        d, cy, cx = plateau(kind)
        n5 = noise(plateau_medians(d, cy, cx)[5], 1e-4, 0.0)
        for fac in (0.95, 1.05):
            dd = d.copy()
            dd[cy, cx] += T * n5 * fac
            want = np.zeros(d.shape, bool)
            want[cy, cx] = fac > 1
            cases.append((dict(data=dd), 1e-4, 0.0, want))
    for k, (frame, g, rn, want) in enumerate(cases):
        say("closed-form noise scene", k)
        contrast = 0.25 if k >= 5 else 3.0
        tn = T if k >= 5 else 3.0
        tf = {key: torch.tensor(v, device=cuda) for key, v in frame.items()}
        outs, masks, _, _ = remove_cosmics_ragged([tf], contrast=contrast, cr_threshold=T,
                                                  neighbor_threshold=tn, maxiter=1,
                                                  frame_gain=[g], frame_readnoise=[rn])
        np.testing.assert_array_equal(masks[0].cpu().numpy(), want)
        wo, wm, _, _ = oracle.clean_frames([frame], contrast=contrast, cr_threshold=T,
                                           neighbor_threshold=tn, maxiter=1, frame_gain=[g],
                                           frame_readnoise=[rn])
        np.testing.assert_array_equal(wm[0], want)
        assert_same_pixels(outs[0].cpu().numpy(), wo[0])


@pytest.mark.parametrize("shape", [(2900, 3001), (1536, 1365)])
def test_host_entry_point_staged_equals_device(cuda, shape):
    """remove_cosmics on host arrays large enough to go through the pinned
    staging buffers (planes of several 32 MiB chunks, the last one partial;
    a mask plane of more than 4 MiB) equals remove_cosmics_device on the same
    inputs, bit for bit; a bad map value is still reported before any
    output is written."""
    from cusmic_amd import _lib, scenes

    say("host entry point, staged copies", shape)
    h, w = shape
    data, signal = scenes.frame(shape, 31, "cpu", n_stars=60, n_tracks=400, max_len=8)
    d = np.ascontiguousarray(data.numpy())
    d[::97, ::89] = np.nan
    err = np.ascontiguousarray(scenes.error_map(signal, 2.0, 5.0).numpy())
    bg = np.ascontiguousarray(np.linspace(-30.0, 30.0, d.size).reshape(shape))
    m = np.ascontiguousarray((np.arange(d.size).reshape(shape) % 211 == 0).astype(np.uint8))
    o = _lib.default_options()
    o.border = _lib.WRAP
    img = _lib.CusmicImage(d.ctypes.data, err.ctypes.data, None, None, bg.ctypes.data, m.ctypes.data, w, h)
    out, flags = np.full(d.shape, 7.0), np.full(d.shape, 7, np.uint8)
    msg = ctypes.create_string_buffer(256)
    st = _lib.lib().remove_cosmics(ctypes.byref(img), ctypes.byref(o), out.ctypes.data, flags.ctypes.data, msg, 256)
    assert st == _lib.OK, msg.value

    dev = torch.device("cuda", 0)
    td, te, tb, tm = (torch.from_numpy(a).to(dev) for a in (d, err, bg, m))
    dimg = _lib.CusmicImage(td.data_ptr(), te.data_ptr(), None, None, tb.data_ptr(), tm.data_ptr(), w, h)
    clean = torch.empty(d.size, dtype=torch.float64, device=dev)
    mask = torch.empty(d.size, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    st, text = single_call(dimg, o, clean, mask, 0)
    assert st == _lib.OK, text
    torch.cuda.synchronize()
    assert_same_pixels(out.ravel(), clean.cpu().numpy())
    np.testing.assert_array_equal(flags.ravel(), mask.cpu().numpy())
    assert flags.sum() > 1000

    say("host entry point, staged copies, bad map value", shape)
    err2 = err.copy()
    err2[-1, -1] = 0.0
    img = _lib.CusmicImage(d.ctypes.data, err2.ctypes.data, None, None, bg.ctypes.data, m.ctypes.data, w, h)
    out2, flags2 = np.full(d.shape, 7.0), np.full(d.shape, 7, np.uint8)
    st = _lib.lib().remove_cosmics(ctypes.byref(img), ctypes.byref(o), out2.ctypes.data, flags2.ctypes.data,
                                   msg, 256)
    assert st == _lib.INVALID and msg.value == b"error must be finite and positive"
    assert (out2 == 7.0).all() and (flags2 == 7).all()


@pytest.mark.parametrize("shape", [(2900, 3001), (48, 64)])
def test_host_entry_point_checks_maps_while_uploading(cuda, shape):
    """With a device, remove_cosmics checks each map as it copies it to the
    staging buffers (planes of several 32 MiB chunks here, or under 4 MiB,
    copied directly). Every bad value is reported with the reference's
    message (src/api.cu:107-114), in its order (error, gain, read noise,
    background), before any kernel runs and with the outputs untouched."""
    from cusmic_amd import _lib, scenes

    h, w = shape
    data, signal = scenes.frame(shape, 7, "cpu", n_stars=4, n_tracks=20, max_len=4)
    d = np.ascontiguousarray(data.numpy())
    err = np.ascontiguousarray(scenes.error_map(signal, 2.0, 5.0).numpy())
    gain = np.full(shape, 2.0)
    rn = np.full(shape, 5.0)
    bg = np.zeros(shape)
    o = _lib.default_options()
    msg = ctypes.create_string_buffer(256)

    def call(error, g, r, b):
        img = _lib.CusmicImage(d.ctypes.data, None if error is None else error.ctypes.data,
                               None if g is None else g.ctypes.data, None if r is None else r.ctypes.data,
                               None if b is None else b.ctypes.data, None, w, h)
        out, flags = np.full(shape, 7.0), np.full(shape, 7, np.uint8)
        st = _lib.lib().remove_cosmics(ctypes.byref(img), ctypes.byref(o), out.ctypes.data,
                                       flags.ctypes.data, msg, 256)
        if st != _lib.OK:
            assert (out == 7.0).all() and (flags == 7).all()
        return st, msg.value.decode()

    def bad(a, where, value):
        a = a.copy()
        a.flat[where] = value
        return a

    first, mid = 0, d.size // 2
    say("map checks while uploading, error first", shape)
    assert call(bad(err, first, 0.0), None, None, bg) == (_lib.INVALID, "error must be finite and positive")
    say("map checks while uploading, background after a good error map", shape)
    assert call(err, None, None, bad(bg, mid, math.inf)) == (_lib.INVALID, "background must be finite")
    say("map checks while uploading, error before background", shape)
    assert call(bad(err, mid, math.nan), None, None, bad(bg, first, math.inf)) == (
        _lib.INVALID, "error must be finite and positive")
    say("map checks while uploading, noise-model maps", shape)
    assert call(None, bad(gain, mid, 0.0), rn, None) == (_lib.INVALID, "gain must be finite and positive")
    assert call(None, gain, bad(rn, d.size - 1, -1.0), None) == (
        _lib.INVALID, "read noise must be finite and nonnegative")
    assert call(None, bad(gain, first, -1.0), bad(rn, first, -1.0), None) == (
        _lib.INVALID, "gain must be finite and positive")
    say("map checks while uploading, good maps", shape)
    assert call(err, None, None, bg)[0] == _lib.OK


@pytest.mark.parametrize("background", [False, True])
def test_output_mask_in_place_and_unaligned(cuda, background):
    """The output mask may be the input mask plane (it is read before the
    output is written), and may start at any byte address: bytes just
    outside it stay untouched. Both equal the oracle, with and without a
    background (two ways of writing the outputs)."""
    from cusmic_amd import _lib, scenes

    h, w = 61, 203
    data, signal = scenes.frame((h, w), 41, "cpu", n_stars=4, n_tracks=14, max_len=5)
    host = data.numpy()
    err = scenes.error_map(signal, 2.0, 5.0).numpy()
    rng = np.random.default_rng(5)
    in_mask = (rng.random((h, w)) < 0.02).astype(np.uint8)
    bg = 7.25
    o = _lib.default_options()
    if background:
        o.background = bg
    want, wmask, _, _ = oracle.clean_frames([dict(data=host, error=err, mask=in_mask)],
                                            background=bg if background else float("nan"))
    assert wmask[0].sum() > 0
    d, e = torch.from_numpy(host).to(cuda), torch.from_numpy(err).to(cuda)

    def run(mask_in, mask_out):
        im = _lib.CusmicImage(d.data_ptr(), e.data_ptr(), None, None, None, mask_in.data_ptr(), w, h)
        clean = torch.empty(h * w, dtype=torch.float64, device=cuda)
        st, msg = single_call(im, o, clean, mask_out, 0)
        assert st == 0, msg
        torch.cuda.synchronize()
        return clean.cpu().numpy().reshape(h, w)

    say("in place, background", background)
    shared = torch.from_numpy(in_mask.ravel().copy()).to(cuda)
    got = run(shared, shared)
    assert_same_pixels(got, want[0])
    np.testing.assert_array_equal(shared.cpu().numpy().reshape(h, w).astype(bool), wmask[0])

    for shift in (1, 3, 8, 13):
        say("output mask at byte offset", shift)
        guard = torch.full((h * w + 32,), 0xAB, dtype=torch.uint8, device=cuda)
        view = guard[shift:shift + h * w]
        got = run(torch.from_numpy(in_mask.ravel().copy()).to(cuda), view)
        assert_same_pixels(got, want[0])
        g = guard.cpu().numpy()
        np.testing.assert_array_equal(g[shift:shift + h * w].reshape(h, w).astype(bool), wmask[0])
        assert (g[:shift] == 0xAB).all() and (g[shift + h * w:] == 0xAB).all()
