"""The CPU oracle pinned against the reference's own fixtures and numpy/scipy.

Pins (each chosen so that a plausible mistake in the restatement fails):
  * test/data fixture: whole loop, mask and every pixel bit-exact;
  * scipy golden vectors: 3/5/7 medians and the L.A.Cosmic Laplacian for all
    five border modes, including 1-pixel axes; 3x3 dilation;
  * reference KATs: test/test_cleaner.py and test/test_batch.cu cases;
  * brute force: replacement against a direct numpy restatement.
"""

import numpy as np
import pytest
from scipy import ndimage

import oracle
from conftest import GOLDEN, assert_same_pixels, bits

SETTINGS = dict(contrast=3, cr_threshold=5, neighbor_threshold=3)


@pytest.mark.parametrize("order", ["exact", "close"])
def test_saved_reference(fixture_scene, order):
    image, error, expected, expected_mask = fixture_scene
    outs, masks, _, passes = oracle.clean_frames(
        [dict(data=image, error=error)], contrast=1, cr_threshold=5, neighbor_threshold=5,
        maxiter=4, laplacian_order=order)
    assert masks[0].sum() == 6518
    np.testing.assert_array_equal(masks[0], expected_mask)
    assert_same_pixels(outs[0], expected)
    assert passes[0] == 3


def test_stage_golden_vectors():
    z = np.load(GOLDEN / "stages.npz")
    checked = 0
    for key in [k[3:] for k in z.files if k.startswith("in_")]:
        x = z["in_" + key]
        for mode in oracle.MODES:
            for size in (3, 5, 7):
                got = oracle.median(x, size, mode)
                np.testing.assert_array_equal(got, z[f"med{size}_{mode}_{key}"], err_msg=f"{size} {mode} {key}")
            lap = oracle.laplacian(x, mode, "exact")
            np.testing.assert_array_equal(bits(lap), bits(z[f"lap_{mode}_{key}"]), err_msg=f"lap {mode} {key}")
            checked += 4
        np.testing.assert_array_equal(oracle.dilate(z["bin_" + key]).astype(bool), z["dil_" + key])
    assert checked > 500


def transposition_sorted(values):
    """The reference's window sort (src/sort.cuh:18-36), element by element:
    odd-even transposition passes that swap a pair only when b < a."""
    v = list(values)
    for _ in range((len(v) + 1) // 2):
        for start in (0, 1):
            for i in range(start, len(v) - 1, 2):
                if v[i + 1] < v[i]:
                    v[i], v[i + 1] = v[i + 1], v[i]
    return v


def segment_median(values):
    """The same result in closed form: no compare with a NaN swaps, so a NaN
    stays where it is, and each run between NaNs (at most 25 long, 26 phases)
    ends sorted in place; the median is the middle slot."""
    v = np.asarray(values)
    mid = len(v) // 2
    if np.isnan(v[mid]):
        return np.nan
    nan = np.flatnonzero(np.isnan(v))
    lo = nan[nan < mid].max() + 1 if (nan < mid).any() else 0
    hi = nan[nan > mid].min() if (nan > mid).any() else len(v)
    return np.sort(v[lo:hi])[mid - lo]


def test_median_with_nan_follows_the_reference_sort():
    """5x5 windows holding NaN (S of infinite samples): the median is the
    middle slot of the reference's sorted window, which depends on where the
    NaN lies, and not the median of the finite values."""
    rng = np.random.default_rng(5)
    checked = differs = 0
    for trial in range(400):
        window = rng.integers(0, 6, 25).astype(np.float64)
        window[rng.random(25) < rng.choice([0.04, 0.15, 0.4])] = np.nan
        if not np.isnan(window).any():
            window[rng.integers(25)] = np.nan
        got = oracle.median(window.reshape(5, 5), 5, "nearest")[2, 2]
        want = transposition_sorted(window)[12]
        np.testing.assert_array_equal(got, want, err_msg=str(window))
        np.testing.assert_array_equal(got, segment_median(window))
        finite = window[~np.isnan(window)]
        differs += bool(np.isnan(got) or got != np.sort(finite)[len(finite) // 2])
        checked += 1
    assert checked == 400 and differs > 100


@pytest.mark.parametrize("mode", ["mirror", "reflect", "nearest", "wrap", "constant"])
def test_laplacian_kat(mode):
    """test/test_cleaner.py:33-47 on both orders: integer frames where the two
    orders agree with scipy's convolution exactly."""
    kernel = np.array([[0, -1, 0], [-1, 4, -1], [0, -1, 0]], dtype=np.float64)
    for pixels in ([[12, -4, 20]], [[12], [-4], [20]], [[12, -4], [20, 0]]):
        image = np.asarray(pixels, dtype=np.float64)
        sampled = np.repeat(np.repeat(image / 4, 2, axis=0), 2, axis=1)
        lap2 = np.maximum(ndimage.convolve(sampled, kernel, mode=mode), 0)
        expected = (lap2[0::2, 0::2] + lap2[0::2, 1::2]) + (lap2[1::2, 0::2] + lap2[1::2, 1::2])
        for order in ("exact", "close"):
            np.testing.assert_array_equal(oracle.laplacian(image, mode, order), expected)


def test_border_index_matches_numpy_pad():
    """numpy.pad implements the same extensions (symmetric=reflect,
    reflect=mirror, edge=nearest, wrap) for pads within one period."""
    pads = {"reflect": "symmetric", "mirror": "reflect", "nearest": "edge", "wrap": "wrap"}
    for n in (2, 3, 5, 9):
        base = np.arange(n)
        for mode, npmode in pads.items():
            padded = np.pad(base, n - 1 if mode == "mirror" else n, mode=npmode)
            r = n - 1 if mode == "mirror" else n
            got = [oracle.border_index(x, n, mode) for x in range(-r, n + r)]
            np.testing.assert_array_equal(got, padded)


def warp_network(values):
    """The reference's warp sort (src/replace.cuh:68-83) lane by lane: in each
    step every lane reads its xor partner and keeps the minimum or the maximum,
    its own value on a tie."""
    v = np.array(values, dtype=np.float64)
    lane = np.arange(32)
    span = 2
    while span <= 32:
        step = span // 2
        while step:
            peer = v[lane ^ step]
            lower = ((lane & span) == 0) == ((lane & step) == 0)
            v = np.where(lower, np.where(peer < v, peer, v), np.where(v < peer, peer, v))
            step //= 2
        span *= 2
    return v


def direct_replace(clean, targets, donors):
    """Straight numpy statement of the replacement (src/replace.cuh:112-150):
    windows of at most 32 pixels sort on the warp network with the window in
    raster order over the lanes (+inf for non-donors), larger ones by key
    (numpy's sort puts -0 and +0 as equals in input order, so the key order
    -0 below +0 is made explicit)."""
    h, w = clean.shape
    out = clean.copy()
    for y, x in zip(*np.nonzero(targets)):
        for r in range(2, max(h, w) + 1):
            y0, y1, x0, x1 = max(0, y - r), min(h, y + r + 1), max(0, x - r), min(w, x + r + 1)
            win = clean[y0:y1, x0:x1]
            ok = donors[y0:y1, x0:x1] & np.isfinite(win)
            vals = win[ok]
            if vals.size:
                n = vals.size
                if win.size <= 32:
                    lanes = np.full(32, np.inf)
                    lanes[:win.size] = np.where(ok, win, np.inf).ravel()
                    vals = warp_network(lanes)
                else:
                    vals = vals[np.lexsort((~np.signbit(vals), vals))]
                lo, hi = vals[(n - 1) // 2], vals[n // 2]
                if n % 2:
                    out[y, x] = 0.0 + lo
                elif abs(lo) < 1 and abs(hi) < 1:
                    out[y, x] = (lo + hi) / 2
                else:
                    out[y, x] = lo / 2 + hi / 2
                break
    return out


def test_replacement_brute_force():
    rng = np.random.default_rng(7)
    for trial in range(40):
        h, w = rng.integers(1, 14, size=2)
        clean = rng.normal(0, 3, size=(h, w)).round(1)
        targets = rng.random((h, w)) < rng.uniform(0.05, 0.8)
        donors = ~targets & (rng.random((h, w)) < 0.9)
        if not donors.any() or not targets.any():
            continue
        want = direct_replace(clean, targets, donors)
        got = clean.copy()
        oracle.replace(got, targets, donors)
        assert_same_pixels(got, want)


def test_bitonic32_is_the_reference_warp_sort():
    """The oracle's pairwise network equals the lane-by-lane statement, sorts
    by value, and leaves -0 and +0 where the network puts them."""
    rng = np.random.default_rng(8)
    moved = 0
    for trial in range(300):
        v = rng.choice([-2.0, -1.0, -0.0, 0.0, 1.0, np.inf], 32)
        got = oracle.bitonic32(v)
        assert_same_pixels(got, warp_network(v))
        np.testing.assert_array_equal(got, np.sort(v))
        key_order = np.sort(np.signbit(v[v == 0]))[::-1]  # -0 before +0
        moved += not np.array_equal(np.signbit(got[got == 0]), key_order)
    assert moved > 100
    np.testing.assert_array_equal(oracle.bitonic32(rng.permutation(32).astype(float)), np.arange(32.0))


def test_replacement_zero_signs():
    """Replacement windows full of +0 and -0: an even-count median of two zero
    middles is -0 only when the network leaves -0 in both middle lanes, which
    the key order (-0 below +0) does not always give; windows above 32 pixels
    do use the key order."""
    rng = np.random.default_rng(9)
    differs = 0
    for trial in range(300):
        h, w = rng.integers(1, 9, size=2)
        clean = rng.choice([-1.0, -0.0, 0.0, 1.0], (h, w), p=[0.2, 0.3, 0.3, 0.2])
        targets = rng.random((h, w)) < 0.3
        donors = ~targets & (rng.random((h, w)) < 0.85)
        if not donors.any() or not targets.any():
            continue
        want = direct_replace(clean, targets, donors)
        got = clean.copy()
        oracle.replace(got, targets, donors)
        assert_same_pixels(got, want)
        by_key = clean.copy()
        for y, x in zip(*np.nonzero(targets)):
            for r in range(2, max(h, w) + 1):
                y0, y1, x0, x1 = max(0, y - r), min(h, y + r + 1), max(0, x - r), min(w, x + r + 1)
                vals = clean[y0:y1, x0:x1][donors[y0:y1, x0:x1]]
                if vals.size:
                    vals = vals[np.lexsort((~np.signbit(vals), vals))]
                    n = vals.size
                    by_key[y, x] = 0.0 + vals[n // 2] if n % 2 else (vals[(n - 1) // 2] + vals[n // 2]) / 2
                    break
        differs += int((bits(by_key) != bits(got)).sum())
    assert differs > 20


def test_replacement_kats():
    """test/test_cleaner.py:135-151: window growth and the overflow-safe midpoint."""
    data = np.asarray([[0.0, 1000, 1000, 1000, 1000, 1000, 60]])
    targets = data == 1000
    oracle.replace(data, targets, ~targets)
    np.testing.assert_array_equal(data, [[0, 0, 0, 30, 60, 60, 60]])
    high = np.finfo("float64").max
    data = np.asarray([[high, 0, high]])
    donors = np.asarray([[True, False, True]])
    oracle.replace(data, ~donors, donors)
    np.testing.assert_array_equal(data, np.full((1, 3), high))


def test_dense_patch():
    """test/test_cleaner.py:22-30"""
    image = np.full((15, 15), 10.0)
    image[5:10, 5:10] = 1000
    outs, masks, _, _ = oracle.clean_frames([dict(data=image, error=np.ones_like(image))], **SETTINGS)
    np.testing.assert_array_equal(outs[0], 10)
    np.testing.assert_array_equal(masks[0], image == 1000)


def test_zero_iterations_background():
    """test/test_cleaner.py:82-89"""
    image = np.full((3, 3), 0.1)
    outs, masks, _, _ = oracle.clean_frames([dict(data=image)], background=1e12, maxiter=0, **SETTINGS)
    assert_same_pixels(outs[0], (np.full((3, 3), 0.1) + 1e12) - 1e12)
    assert not masks[0].any()


def test_nonfinite_pixels():
    """test/test_cleaner.py:92-112"""
    data = np.full((9, 9), 10.0)
    data[2:7, 2:7] = np.nan
    data[3, 3], data[3, 5], data[4, 4] = np.inf, -np.inf, 1000
    data.view("uint64")[2, 2] = 0x7FF8000000000001
    for maxiter in (0, 4):
        outs, masks, _, _ = oracle.clean_frames([dict(data=data, error=np.ones_like(data))],
                                                background=100, maxiter=maxiter, **SETTINGS)
        expected = data.copy()
        if maxiter:
            expected[4, 4] = 10
        assert_same_pixels(outs[0], expected)
        assert masks[0].sum() == bool(maxiter)
    data[:] = np.nan
    outs, masks, _, _ = oracle.clean_frames([dict(data=data, error=np.ones_like(data))], **SETTINGS)
    assert_same_pixels(outs[0], data)
    assert not masks[0].any()


def test_singleton_borders():
    """test/test_batch.cu:113-141: 1x9 and 9x1 frames, every border mode."""
    probe = np.full(9, 10.0)
    probe[4] = 1000
    for shape in ((1, 9), (9, 1)):
        for mode in oracle.MODES:
            outs, masks, _, _ = oracle.clean_frames(
                [dict(data=probe.reshape(shape), error=np.ones(shape))], border_mode=mode, **SETTINGS)
            want = np.zeros(9, bool)
            want[4] = True
            if mode == "constant":
                want[[0, 1, 7, 8]] = True
            np.testing.assert_array_equal(outs[0].ravel(), 10)
            np.testing.assert_array_equal(masks[0].ravel(), want)


def test_no_donor_error_and_validation():
    data = np.full((5, 5), 10.0)
    data[2, 2] = 1e6
    with pytest.raises(oracle.OracleError, match="no finite replacement donors"):
        oracle.clean_frames([dict(data=data, error=np.ones_like(data), mask=np.ones((5, 5), bool) ^ (data > 100))],
                            **SETTINGS)
    with pytest.raises(oracle.OracleError, match="invalid detection settings"):
        oracle.clean_frames([dict(data=data, error=np.ones_like(data))], contrast=np.nan)
    with pytest.raises(oracle.OracleError, match="positive gain or error"):
        oracle.clean_frames([dict(data=data)], **SETTINGS)


def test_batch_equals_single_frames():
    """test/test_batch.cu:85-112 semantics: a batch equals independent calls,
    although the batch continues until every frame converges."""
    rng = np.random.default_rng(3)
    frames = []
    for f in range(3):
        d = rng.normal(100, 3, size=(20 + f, 17))
        d[rng.integers(0, 20, 5), rng.integers(0, 17, 5)] += 500 * (f + 1)
        frames.append(dict(data=d, error=np.full_like(d, 3.0)))
    outs, masks, _, _ = oracle.clean_frames(frames, **SETTINGS)
    for fr, o, m in zip(frames, outs, masks):
        o1, m1, _, _ = oracle.clean_frames([fr], **SETTINGS)
        assert_same_pixels(o, o1[0])
        np.testing.assert_array_equal(m, m1[0])


def test_donor_record_consistent():
    """Recorded donors reproduce every flagged pixel from the input values."""
    rng = np.random.default_rng(11)
    d = rng.normal(1000, 20, size=(40, 50))
    d[rng.integers(0, 40, 30), rng.integers(0, 50, 30)] += 5000
    outs, masks, donors, _ = oracle.clean_frames([dict(data=d, gain=2.0, readnoise=5.0)],
                                                 record_donors=True, **SETTINGS)
    flat = d.ravel()
    for y, x in zip(*np.nonzero(masks[0])):
        lo, hi = donors[0][y, x]
        assert lo >= 0 and hi >= 0 and not masks[0].ravel()[lo] and not masks[0].ravel()[hi]
        want = flat[lo] if lo == hi else (flat[lo] / 2 + flat[hi] / 2)
        assert outs[0][y, x] == want
    assert (donors[0][~masks[0]] == -1).all()
