"""HIP path against the oracle and the reference's fixtures (GPU).

Bar: masks identical and pixels bit-identical (integer/selection arithmetic
and IEEE-exact float stages), on the same seeded inputs.
"""

import ctypes
import math

import numpy as np
import pytest
import torch

import oracle
from conftest import assert_same_pixels, bits

pytestmark = pytest.mark.gpu

MODES = ("mirror", "reflect", "nearest", "wrap", "constant")


def say(*args):
    print("[case]", *args, flush=True)


def np_of(t):
    return t.detach().cpu().numpy()


def test_arith_selftest(cuda):
    """IEEE division and square root on the GPU equal the host bit for bit."""
    from cusmic_amd import _lib

    say("arith selftest")
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.uniform(0, 1e6, 50000), rng.lognormal(0, 30, 50000),
                        np.array([0.0, 1e-310, 5e-324, 1.7e308, 1.0, 2.0])])
    b = np.concatenate([rng.uniform(1e-3, 100, 50000), rng.lognormal(0, 30, 50000),
                        np.array([3.0, 7.0, 3.0, 1e-300, 3.0, 3.0])])
    ta, tb = torch.tensor(a, device=cuda), torch.tensor(b, device=cuda)
    q, r = torch.empty_like(ta), torch.empty_like(ta)
    stream = torch.cuda.current_stream().cuda_stream
    st = _lib.lib().cusmic_selftest_arith(ctypes.c_void_p(ta.data_ptr()), ctypes.c_void_p(tb.data_ptr()),
                                          ctypes.c_void_p(q.data_ptr()), ctypes.c_void_p(r.data_ptr()),
                                          a.size, ctypes.c_void_p(stream))
    assert st == 0
    with np.errstate(over="ignore"):
        assert_same_pixels(np_of(q), a / b)
    assert_same_pixels(np_of(r), np.sqrt(a))


def test_saved_reference_torch(cuda, fixture_scene):
    from cusmic_amd import remove_cosmics

    say("saved reference, python API")
    image, error, expected, expected_mask = fixture_scene
    cleaned, mask = remove_cosmics(torch.tensor(image, device=cuda), error=torch.tensor(error, device=cuda),
                                   contrast=1, cr_threshold=5, neighbor_threshold=5, maxiter=4)
    np.testing.assert_array_equal(np_of(mask), expected_mask)
    assert_same_pixels(np_of(cleaned), expected)


def test_saved_reference_host_abi(cuda, fixture_scene):
    """test/test_reference.c: the host entry point on the saved scene."""
    from cusmic_amd import _lib

    say("saved reference, host ABI")
    image, error, expected, expected_mask = fixture_scene
    image, error = np.ascontiguousarray(image), np.ascontiguousarray(error)
    clean = np.empty_like(image)
    mask = np.empty(image.shape, np.uint8)
    img = _lib.CusmicImage(image.ctypes.data, error.ctypes.data, None, None, None, None, 512, 512)
    o = _lib.default_options()
    o.contrast, o.cr_threshold, o.neighbor_threshold = 1, 5, 5
    msg = ctypes.create_string_buffer(256)
    st = _lib.lib().remove_cosmics(ctypes.byref(img), ctypes.byref(o), clean.ctypes.data,
                                   mask.ctypes.data, msg, 256)
    assert st == 0, msg.value
    np.testing.assert_array_equal(mask.astype(bool), expected_mask)
    assert_same_pixels(clean, expected)


def scene(shape, seed, nan=0.0, mask_frac=0.0, stars=12, tracks=25, max_len=6):
    from cusmic_amd import scenes

    data, signal = scenes.frame(shape, seed, "cpu", n_stars=stars, n_tracks=tracks, max_len=max_len)
    rng = np.random.default_rng(seed)
    d = data.numpy().copy()
    if nan:
        d[rng.random(shape) < nan] = np.nan
    m = rng.random(shape) < mask_frac if mask_frac else None
    err = scenes.error_map(signal, 2.0, 5.0).numpy()
    return d, err, m


def gpu_clean(frames, **kw):
    from cusmic_amd import remove_cosmics_ragged

    dev = torch.device("cuda", 0)
    tf = []
    for fr in frames:
        tf.append({k: (torch.tensor(v, device=dev) if isinstance(v, np.ndarray) else v)
                   for k, v in fr.items()})
    outs, masks, donors, passes = remove_cosmics_ragged(tf, **kw)
    return ([np_of(o) for o in outs], [np_of(m) for m in masks],
            None if donors is None else [np_of(d) for d in donors], passes)


ORACLE_KEYS = dict(contrast="contrast", cr_threshold="cr_threshold",
                   neighbor_threshold="neighbor_threshold", maxiter="maxiter",
                   border_mode="border_mode", frame_gain="frame_gain",
                   frame_readnoise="frame_readnoise", record_donors="record_donors")


def compare(frames, laplacian="exact", **kw):
    try:
        want = oracle.clean_frames(frames, laplacian_order=laplacian, **kw)
    except oracle.OracleError as e:
        with pytest.raises(ValueError, match=str(e)):
            gpu_clean(frames, laplacian=laplacian, **kw)
        return None, None
    got = gpu_clean(frames, laplacian=laplacian, **kw)
    for f, (o, m, wo, wm) in enumerate(zip(got[0], got[1], want[0], want[1])):
        np.testing.assert_array_equal(m, wm, err_msg=f"mask of frame {f}")
        assert_same_pixels(o, wo)
    np.testing.assert_array_equal(got[3], want[3])
    if kw.get("record_donors"):
        for d, wd in zip(got[2], want[2]):
            np.testing.assert_array_equal(d, wd)
    return got, want


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("laplacian", ["exact", "close"])
def test_error_map_scene(cuda, mode, laplacian):
    d, err, m = scene((150, 211), 5, nan=2e-3, mask_frac=5e-3)
    say("error-map scene", mode, laplacian)
    got, want = compare([dict(data=d, error=err, mask=m)], laplacian=laplacian, border_mode=mode,
                        record_donors=True)
    assert want[1][0].sum() > 20


@pytest.mark.parametrize("mode", MODES)
def test_noise_model_scene(cuda, mode):
    d, _, m = scene((133, 97), 9, nan=1e-3, mask_frac=2e-3)
    say("noise-model scene", mode)
    compare([dict(data=d, mask=m)], border_mode=mode, frame_gain=[2.0], frame_readnoise=[5.0])


def test_maps_and_background(cuda):
    d, err, m = scene((128, 128), 13, nan=1e-3, mask_frac=1e-2)
    rng = np.random.default_rng(1)
    gain = rng.uniform(1.5, 2.5, d.shape)
    rn = rng.uniform(3, 7, d.shape)
    bg = rng.uniform(-50, 50, d.shape)
    say("gain/readnoise/background maps")
    compare([dict(data=d, gain=gain, readnoise=rn, background=bg, mask=m)])
    say("scalar background via the python API")
    from cusmic_amd import remove_cosmics

    c, k = remove_cosmics(torch.tensor(d, device=cuda), 3, 5, 3, error=torch.tensor(err, device=cuda),
                          background=150.0)
    wo, wm, _, _ = oracle.clean_frames([dict(data=d, error=err)], background=150.0)
    np.testing.assert_array_equal(np_of(k), wm[0])
    assert_same_pixels(np_of(c), wo[0])


@pytest.mark.parametrize("mode", MODES)
def test_invalid_noise_values(cuda, mode):
    """Device path without content checks: negative error values make S
    negative, so the S' <= S screen must give way around them (first pass
    and the dirty-block passes)."""
    d, err, m = scene((120, 150), 21, nan=1e-3, mask_frac=2e-3)
    rng = np.random.default_rng(21)
    bad = rng.random(err.shape) < 5e-3
    err = err.copy()
    err[bad] = -err[bad]
    say("invalid noise values", mode)
    got, want = compare([dict(data=d, error=err, mask=m)], border_mode=mode, record_donors=True)
    assert want[1][0].sum() > 20


def test_many_passes(cuda):
    """Low thresholds keep frames growing for many passes: the incremental
    replacement, dirty blocks and stamps must track the dense loop."""
    frames = []
    for i, shape in enumerate(((96, 130), (65, 200), (150, 64))):
        d, err, m = scene(shape, 30 + i, nan=1e-3, mask_frac=2e-3, stars=20, tracks=60, max_len=10)
        frames.append(dict(data=d, error=err, mask=m))
    for mode in ("mirror", "wrap", "constant"):
        say("many passes", mode)
        got, want = compare(frames, contrast=1.0, cr_threshold=3.0, neighbor_threshold=1.5,
                            maxiter=12, border_mode=mode, record_donors=True)
        assert max(want[3]) > 4, want[3]


def test_many_passes_noise_model(cuda):
    """Under the noise model S' reaches one pixel farther (N is a median of C),
    so its dirty blocks reach farther than an error-map batch's: frames of both
    kinds in one batch, many passes, dense tracks."""
    frames = []
    for i, shape in enumerate(((120, 140), (70, 190))):
        d, _, m = scene(shape, 40 + i, nan=1e-3, mask_frac=2e-3, stars=20, tracks=70, max_len=10)
        frames.append(dict(data=d, mask=m))
    d, err, m = scene((90, 110), 42, stars=20, tracks=50, max_len=10)
    frames.append(dict(data=d, error=err, mask=m))
    for mode in ("reflect", "nearest", "wrap"):
        say("many passes, noise model", mode)
        got, want = compare(frames, contrast=1.0, cr_threshold=3.0, neighbor_threshold=1.5,
                            maxiter=12, border_mode=mode, frame_gain=[2.0, 2.5, 1.0],
                            frame_readnoise=[5.0, 4.0, 1.0])
        assert max(want[3]) > 4, want[3]


def test_many_passes_interior(cuda):
    """A frame with interior tiles and blocks on every side of the changes."""
    d, err, m = scene((256, 300), 50, nan=5e-4, mask_frac=1e-3, stars=40, tracks=300, max_len=10)
    for mode in ("mirror", "constant"):
        say("many passes, interior", mode)
        got, want = compare([dict(data=d, error=err, mask=m)], contrast=1.0, cr_threshold=3.0,
                            neighbor_threshold=1.5, maxiter=10, border_mode=mode)
        assert want[3][0] > 4, want[3]


def test_tiny_and_odd_shapes(cuda):
    rng = np.random.default_rng(2)
    for shape in ((1, 9), (9, 1), (1, 1), (2, 2), (3, 70), (70, 3), (7, 7), (33, 65), (65, 33)):
        for mode in MODES:
            d = rng.normal(100, 2, size=shape)
            d.flat[rng.integers(0, d.size)] += 2000
            say("shape", shape, mode)
            compare([dict(data=d, error=np.full(shape, 2.0))], border_mode=mode)


@pytest.mark.parametrize("mode", MODES)
def test_rays_on_frame_edges(cuda, mode):
    """Cosmic-ray pixels on the four edges, at the corners and one to four
    pixels in: their seeds take the fine-structure path for windows that cross
    the frame edge (the border mode's extension), over several passes."""
    rng = np.random.default_rng(11)
    h, w = 96, 80
    d = rng.normal(100, 3, size=(h, w))
    hits = [(0, 0), (0, w - 1), (h - 1, 0), (h - 1, w - 1), (0, 37), (h - 1, 21), (45, 0),
            (61, w - 1), (1, 1), (2, w - 3), (h - 3, 2), (3, 40), (50, 3), (h - 4, w - 4), (4, 4)]
    for y, x in hits:
        d[y, x] += 1500
        d[min(y + 1, h - 1), x] += 900
    say("edge rays", mode)
    compare([dict(data=d, error=np.full((h, w), 3.0))], border_mode=mode, maxiter=6,
            record_donors=True)
    compare([dict(data=d)], border_mode=mode, frame_gain=[2.0], frame_readnoise=[5.0], maxiter=6)


def test_ragged_batch_equals_frames(cuda):
    shapes = ((64, 64), (31, 130), (200, 77), (1, 50), (96, 96))
    frames = []
    for i, s in enumerate(shapes):
        d, err, m = scene(s, 20 + i, nan=1e-3, mask_frac=5e-3, stars=3, tracks=8)
        frames.append(dict(data=d, error=err, mask=m))
    say("ragged batch")
    batch, _ = compare(frames, record_donors=True)
    for i, fr in enumerate(frames):
        say("ragged batch frame", i)
        single = gpu_clean([fr], record_donors=True)
        assert_same_pixels(batch[0][i], single[0][0])
        np.testing.assert_array_equal(batch[1][i], single[1][0])
        np.testing.assert_array_equal(batch[2][i], single[2][0])


def test_stack_api_and_sliced_view(cuda):
    """test/test_batch.py:6-26: a sliced stack equals independent calls and the
    input is untouched."""
    from cusmic_amd import Cleaner, Image

    say("sliced stack")
    movie = torch.full((4, 9, 18), 10.0, dtype=torch.float64, device=cuda)
    view = movie[::2, :, ::2]
    view[0, 4, 4] = 1000
    view[1] = float("nan")
    original = movie.clone()
    image = Image(view, error=torch.ones((9, 9), dtype=torch.float64, device=cuda))
    assert image.data is view
    cleaner = Cleaner()
    clean, mask = cleaner(image)
    for i in range(2):
        want, flags = cleaner(Image(view[i], error=image.error))
        assert_same_pixels(np_of(clean[i]), np_of(want))
        np.testing.assert_array_equal(np_of(mask[i]), np_of(flags))
    assert torch.equal(movie.view(torch.int64), original.view(torch.int64))
    movie.fill_(0)
    np.testing.assert_array_equal(np_of(clean[0]), 10)
    assert bool(torch.isnan(clean[1]).all()) and not bool(mask[1].any())


def test_reference_kats(cuda):
    from cusmic_amd import Cleaner, Image, remove_cosmics

    say("dense patch")
    image = torch.full((15, 15), 10.0, dtype=torch.float64, device=cuda)
    image[5:10, 5:10] = 1000
    cleaned, mask = remove_cosmics(image, 3, 5, 3, error=torch.ones_like(image))
    np.testing.assert_array_equal(np_of(cleaned), 10)
    np.testing.assert_array_equal(np_of(mask), np_of(image == 1000))

    say("zero iterations background")
    image = torch.full((3, 3), 0.1, dtype=torch.float64, device=cuda)
    cleaned, mask = remove_cosmics(image, 3, 5, 3, background=1e12, maxiter=0)
    assert_same_pixels(np_of(cleaned), (np.full((3, 3), 0.1) + 1e12) - 1e12)
    assert not bool(mask.any())

    say("nonfinite pixels")
    data = np.full((9, 9), 10.0)
    data[2:7, 2:7] = np.nan
    data[3, 3], data[3, 5], data[4, 4] = np.inf, -np.inf, 1000
    data.view("uint64")[2, 2] = 0x7FF8000000000001
    original = data.copy()
    for maxiter in (0, 4):
        clean, mask = remove_cosmics(data, 3, 5, 3, error=np.ones_like(data), background=100, maxiter=maxiter)
        expected = data.copy()
        if maxiter:
            expected[4, 4] = 10
        assert_same_pixels(np_of(clean), expected)
        assert int(mask.sum()) == bool(maxiter)
        assert_same_pixels(data, original)
    data[:] = np.nan
    clean, mask = remove_cosmics(data, 3, 5, 3, error=np.ones_like(data))
    assert_same_pixels(np_of(clean), data)
    assert not bool(mask.any())

    say("independent calls")
    cleaner = Cleaner()
    for shape in ((9, 9), (7, 11)):
        data = torch.full(shape, 10.0, dtype=torch.float64, device=cuda)
        data[3, 3] = 1000
        cleaned, mask = cleaner(Image(data, error=torch.ones_like(data)))
        np.testing.assert_array_equal(np_of(cleaned), 10)
        assert int(mask.sum()) == 1
        assert float(data[3, 3]) == 1000


def test_singleton_borders_host(cuda):
    """test/test_batch.cu:113-141 through the host entry point."""
    from cusmic_amd import _lib

    probe = np.full(9, 10.0)
    probe[4] = 1000
    error = np.ones(9)
    o = _lib.default_options()
    for width, height in ((9, 1), (1, 9)):
        for mode in range(5):
            say("singleton", width, height, mode)
            o.border = mode
            clean = np.empty(9)
            det = np.empty(9, np.uint8)
            img = _lib.CusmicImage(probe.ctypes.data, error.ctypes.data, None, None, None, None, width, height)
            msg = ctypes.create_string_buffer(128)
            st = _lib.lib().remove_cosmics(ctypes.byref(img), ctypes.byref(o), clean.ctypes.data,
                                           det.ctypes.data, msg, 128)
            assert st == 0, msg.value
            want = np.zeros(9, bool)
            want[4] = True
            if mode == _lib.CONSTANT:
                want[[0, 1, 7, 8]] = True
            np.testing.assert_array_equal(clean, 10)
            np.testing.assert_array_equal(det.astype(bool), want)


def test_no_donors_raises(cuda):
    from cusmic_amd import remove_cosmics

    say("no donors")
    data = np.full((5, 5), 10.0)
    data[2, 2] = 1e6
    mask = np.ones((5, 5), bool)
    mask[2, 2] = False
    with pytest.raises(ValueError, match="no finite replacement donors"):
        remove_cosmics(data, 3, 5, 3, error=np.ones_like(data), mask=mask)


def test_stream_handoff(cuda):
    """test/test_streams.py:7-33 with torch streams and events."""
    from cusmic_amd import Cleaner, Image

    say("stream handoff")
    producer, consumer = torch.cuda.Stream(), torch.cuda.Stream()
    ready, done = torch.cuda.Event(), torch.cuda.Event()
    data = torch.empty((9, 9), dtype=torch.float64, device=cuda)
    error = torch.empty((9, 9), dtype=torch.float64, device=cuda)
    with torch.cuda.stream(producer):
        torch.cuda._sleep(2_000_000)
        data.fill_(10)
        data[4, 4] = 1000
        error.fill_(1)
        ready.record()
    with torch.cuda.stream(consumer):
        consumer.wait_event(ready)
        clean, mask = Cleaner()(Image(data, error=error))
        done.record()
    with torch.cuda.stream(producer):
        producer.wait_event(done)
        data.fill_(0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(np_of(clean), 10)
    assert int(mask.sum()) == 1


def test_validation_python(cuda):
    from cusmic_amd import Cleaner, Image

    for data in (torch.ones(3, device=cuda, dtype=torch.float64),
                 torch.ones((0, 3), device=cuda, dtype=torch.float64),
                 torch.ones((1, 2, 3, 4), device=cuda, dtype=torch.float64)):
        with pytest.raises(ValueError, match="nonempty 2D"):
            Image(data)
    for dtype in (torch.bool, torch.complex128, torch.float32):
        with pytest.raises(TypeError, match="float64"):
            Image(torch.ones((3, 3), dtype=dtype, device=cuda))
    data = torch.ones((3, 3), dtype=torch.float64, device=cuda)
    for name in ("error", "mask", "background", "effective_gain", "readnoise"):
        with pytest.raises(ValueError, match=name):
            Image(data, **{name: torch.ones((1, 3), device=cuda)})
    for name, value in (("error", 0), ("effective_gain", -1), ("readnoise", -1), ("background", math.inf)):
        with pytest.raises(ValueError, match=name):
            Image(data, **{name: torch.full_like(data, value)})
    with pytest.raises(ValueError, match="Provide error"):
        Cleaner()(Image(data))


def test_large_maxiter_stops_at_convergence(cuda):
    """maxiter far beyond convergence (src/cleaner.cuh:175-181 breaks out of
    the loop): a scene that needs 14 passes runs through four chunks of the
    pass-counter ring and equals the oracle; a tiny frame at maxiter = 70000
    does too."""
    from cusmic_amd import scenes

    d, sig = scenes.frame((128, 160), 54, "cpu", n_stars=30, n_tracks=90, max_len=12)
    d = d.numpy()
    err = scenes.error_map(sig, 2.0, 5.0).numpy()
    say("14-pass scene, maxiter 100000")
    got, want = compare([dict(data=d, error=err)], contrast=1.0, cr_threshold=3.0, neighbor_threshold=2.0,
                        maxiter=100000, record_donors=True)
    assert want[3][0] == 14, want[3]
    say("14-pass scene with a converged frame beside it, maxiter 9")
    d2 = np.full((40, 40), 100.0)
    d2[20, 20] = 5000
    got, want = compare([dict(data=d, error=err), dict(data=d2, error=np.full_like(d2, 2.0))],
                        contrast=1.0, cr_threshold=3.0, neighbor_threshold=2.0, maxiter=9)
    assert list(want[3]) == [9, 2], want[3]
    say("tiny frame, maxiter 70000")
    small = np.full((9, 11), 10.0)
    small[4, 5] = 1000
    small[2, 2] = 700
    got, want = compare([dict(data=small, error=np.ones_like(small))], maxiter=70000)
    assert want[1][0].sum() == 2 and want[3][0] == 2


def test_large_maxiter_costs_no_more(cuda):
    """A scene that converges within the first chunk costs about the same at
    maxiter = 1000 as at 4: passes after convergence are not enqueued."""
    import time

    from cusmic_amd import _lib, engine

    d, err, _ = scene((512, 512), 7, stars=20, tracks=100)
    fr = engine.Frames()
    fr.add(torch.tensor(d, device=cuda), error=torch.tensor(err, device=cuda))

    def timed(maxiter):
        engine.run(fr, cuda, 3.0, 5.0, 3.0, maxiter, _lib.MIRROR)
        ts = []
        for _ in range(7):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, _, _, passes = engine.run(fr, cuda, 3.0, 5.0, 3.0, maxiter, _lib.MIRROR)
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), passes[0]

    t4, p4 = timed(4)
    t1000, p1000 = timed(1000)
    say("maxiter 4:", t4, p4, "maxiter 1000:", t1000, p1000)
    assert p4 == p1000 <= 4
    assert t1000 < 1.5 * t4 + 2e-4, (t4, t1000)


def test_iteration_log(cuda, caplog):
    """mod/cusmic/cleaner.py:108 logs each pass's new pixels over the stack,
    up to the pass that finds none."""
    import logging

    from cusmic_amd import remove_cosmics

    say("iteration log")
    stack = torch.full((2, 15, 15), 10.0, dtype=torch.float64, device=cuda)
    stack[0, 5:10, 5:10] = 1000
    stack[1, 3, 11] = 900
    with caplog.at_level(logging.INFO, logger="cusmic_amd.cleaner"):
        clean, mask = remove_cosmics(stack, 3, 5, 3, error=torch.ones_like(stack))
    lines = [r.getMessage() for r in caplog.records if r.name == "cusmic_amd.cleaner"]
    # pass k's new pixels from the oracle: the first k passes are a run with
    # maxiter = k, so its mask grows by exactly those pixels
    host = np_of(stack)
    frames = [dict(data=host[f], error=np.ones((15, 15))) for f in range(2)]
    want, flagged, k = [], 0, 0
    while True:
        k += 1
        _, masks, _, _ = oracle.clean_frames(frames, maxiter=k)
        now = int(sum(m.sum() for m in masks))
        want.append(f"Iteration {k}: {now - flagged} new cosmic-ray pixels")
        if now == flagged:
            break
        flagged = now
    assert lines == want, (lines, want)
    assert int(mask.sum()) == flagged and len(want) > 2


def test_background_overflow(cuda):
    """src/cleaner.cuh:87-97: a background that makes every donor infinite
    while holes need filling raises; one that overflows only some pixels
    (infinite samples of C) gives the oracle's masks and pixels."""
    say("background overflow: no finite donor")
    d = np.full((20, 24), 1.5e308)
    d[5, 7] = np.nan
    compare([dict(data=d, error=np.ones_like(d), background=np.full_like(d, 1e308))])
    with pytest.raises(ValueError, match="no finite replacement donors"):
        gpu_clean([dict(data=d, error=np.ones_like(d), background=np.full_like(d, 1e308))])
    say("background overflow at some pixels")
    rng = np.random.default_rng(17)
    d, err, m = scene((96, 112), 23, nan=2e-3, mask_frac=2e-3)
    big = rng.random(d.shape) < 0.01
    d = d.copy()
    d[big] = 1.5e308
    bg = np.where(big, 1e308, rng.uniform(-20, 20, d.shape))
    for mode in ("mirror", "constant"):
        got, want = compare([dict(data=d, error=err, mask=m, background=bg)], border_mode=mode,
                            record_donors=True)
        assert want[1][0].sum() > 10
    say("background overflow, noise model")
    compare([dict(data=d, mask=m, background=bg)], frame_gain=[2.0], frame_readnoise=[5.0])
    say("background overflow, scalar background")
    from cusmic_amd import Cleaner, Image

    # scaled so that a scalar background overflows the DBL_MAX pixels only
    ds, es = d * 1e290, err * 1e290
    ds[big] = np.finfo(np.float64).max
    dev = torch.device("cuda", 0)
    for mode in ("mirror", "wrap"):
        outs, masks, _, passes = oracle.clean_frames([dict(data=ds, error=es, mask=m)], background=3e292,
                                                    border_mode=mode)
        clean, mask, _, gpasses = Cleaner(border_mode=mode).run(
            Image(torch.tensor(ds, device=dev), error=torch.tensor(es, device=dev),
                  mask=torch.tensor(m, device=dev), background=3e292))
        np.testing.assert_array_equal(np_of(mask), masks[0])
        assert_same_pixels(np_of(clean), outs[0])
        assert list(gpasses) == list(passes) and masks[0].sum() > 10


def zero_scene(shape, rng):
    """+0 and -0 everywhere, a few +-1, rays of 40 and runs of up to 7 pixels
    of 400 along rows (widened windows), holes and masked pixels."""
    h, w = shape
    d = rng.choice(np.array([-0.0, 0.0]), shape)
    d[rng.random(shape) < 0.08] = rng.choice([-1.0, 1.0])
    d[rng.random(shape) < 0.04] = 40.0
    for _ in range(max(2, h * w // 200)):
        y, x = rng.integers(h), rng.integers(w)
        d[y, x:x + rng.integers(2, 8)] = 400.0
    d[rng.random(shape) < 0.01] = np.nan
    return d, rng.random(shape) < 0.03


def test_signed_zero_ties(cuda):
    """Replacement windows whose middle donors are zeros of both signs: the
    sign is the one the reference's 32-lane value network leaves for windows
    of at most 32 pixels (also widened windows of narrow frames), the key
    order for larger ones, as the oracle computes them."""
    rng = np.random.default_rng(29)
    negative = positive = 0
    for shape in ((64, 96), (1, 80), (2, 70), (7, 5), (40, 3)):
        say("signed zero ties", shape)
        d, m = zero_scene(shape, rng)
        for mode in ("mirror", "constant"):
            got, want = compare([dict(data=d, error=np.ones(shape), mask=m)], border_mode=mode,
                                contrast=1, record_donors=True)
            out, flagged = want[0][0], want[1][0] | np.isnan(d)
            zero = flagged & (out == 0)
            negative += int((zero & np.signbit(out)).sum())
            positive += int((zero & ~np.signbit(out)).sum())
    assert negative > 20 and positive > 20
