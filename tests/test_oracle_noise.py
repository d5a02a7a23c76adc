"""The oracle's noise-model path pinned by closed-form flag thresholds (CPU).

Each scene's mask follows from the published noise model alone (see
noise_scenes.py), so the checks fail for a misread gain (g versus 1/g, a
missing / g), read noise (rn versus rn^2), signal floor (1e-5, applied to the
median before the gain), map index, the factor 2 of S = L / (2 N), or a median
size other than 5.
"""

import numpy as np
import pytest

import oracle
from noise_scenes import noise, plateau, plateau_medians, raised_pixels

T = 5.0

SCALARS = [
    (2.0, 5.0, 1000.0),   # the benchmark frames' calibration
    (1.0, 0.0, 0.0),      # c = 0: the floor, no read noise
    (4.0, 10.0, -30.0),   # negative sky: the floor
    (0.5, 3.0, 1e-6),     # c below the floor
    (1.7, 0.0, 250.0),    # Poisson only
    (3.0, 7.5, 2e-5),     # just above the floor
    (0.25, 12.0, 5e4),    # gain below 1, bright sky
]


@pytest.mark.parametrize("g,rn,c", SCALARS)
def test_flag_threshold_scalars(g, rn, c):
    data, want = raised_pixels((40, 44), c, g, rn, T, seed=int(abs(c)) + 1)
    _, masks, _, passes = oracle.clean_frames([dict(data=data)], gain=g, readnoise=rn,
                                              cr_threshold=T, neighbor_threshold=3.0, contrast=3.0,
                                              maxiter=1)
    np.testing.assert_array_equal(masks[0], want)
    raised = data != c
    assert want.sum() >= 3 and (raised & ~want).sum() >= 3


def test_flag_threshold_maps():
    """Per-pixel gain and read-noise maps: each raised pixel uses its own
    map values (a neighbour's would move its flip by far more than 1e-6)."""
    rng = np.random.default_rng(5)
    shape = (49, 53)
    g = rng.uniform(1.0, 4.0, shape)
    rn = rng.uniform(2.0, 10.0, shape)
    for c in (800.0, 0.0, 3e-6):
        data, want = raised_pixels(shape, c, g, rn, T, seed=7)
        _, masks, _, _ = oracle.clean_frames([dict(data=data, gain=g, readnoise=rn)], cr_threshold=T,
                                             neighbor_threshold=3.0, contrast=3.0, maxiter=1)
        np.testing.assert_array_equal(masks[0], want)
        assert want.sum() >= 5 and (~want).sum() > 0


def test_flag_threshold_mixed_map_and_scalar():
    """A gain map with a scalar read noise, and the reverse."""
    rng = np.random.default_rng(8)
    shape = (40, 40)
    g = rng.uniform(0.5, 3.0, shape)
    rn = rng.uniform(1.0, 9.0, shape)
    data, want = raised_pixels(shape, 500.0, g, 4.0, T, seed=3)
    _, masks, _, _ = oracle.clean_frames([dict(data=data, gain=g)], readnoise=4.0, cr_threshold=T,
                                         neighbor_threshold=3.0, maxiter=1)
    np.testing.assert_array_equal(masks[0], want)
    data, want = raised_pixels(shape, 500.0, 2.5, rn, T, seed=4)
    _, masks, _, _ = oracle.clean_frames([dict(data=data, readnoise=rn)], gain=2.5, cr_threshold=T,
                                         neighbor_threshold=3.0, maxiter=1)
    np.testing.assert_array_equal(masks[0], want)


@pytest.mark.parametrize("kind", ["X", "Y"])
def test_median_size_is_five(kind):
    """With gain 1e-4 and no read noise, N at the centre is 0.32 for a median
    of 1e-5 and 10 for a median of 0.01: design X separates the 5x5 median
    from the 3x3 one, design Y from the 7x7 one."""
    g, rn = 1e-4, 0.0
    data, cy, cx = plateau(kind)
    meds = plateau_medians(data, cy, cx)
    if kind == "X":
        assert meds[3] != meds[5] == meds[7]
    else:
        assert meds[3] == meds[5] != meds[7]
    n5 = noise(meds[5], g, rn)
    other = 3 if kind == "X" else 7
    opposite = 0
    for fac, flagged in ((0.95, False), (1.05, True)):
        d = data.copy()
        d[cy, cx] += T * n5 * fac
        # does the size a misreading would use predict the other outcome?
        opposite += (T * n5 * fac > T * noise(meds[other], g, rn)) != flagged
        _, masks, _, _ = oracle.clean_frames([dict(data=d)], gain=g, readnoise=rn, cr_threshold=T,
                                             neighbor_threshold=T, contrast=0.25, maxiter=1)
        want = np.zeros(d.shape, bool)
        want[cy, cx] = flagged
        np.testing.assert_array_equal(masks[0], want)
    assert opposite == 1


def test_threshold_factor_two():
    """S = L / (2 N): the Laplacian of a raised pixel is 2a, so the flip sits
    at a = t N, not at t N / 2 or 2 t N."""
    g, rn, c = 2.0, 5.0, 100.0
    n = noise(c, g, rn)
    for a, flagged in ((0.6 * T * n, False), (1.2 * T * n, True), (2.2 * T * n, True)):
        d = np.full((21, 21), c)
        d[10, 10] += a
        _, masks, _, _ = oracle.clean_frames([dict(data=d)], gain=g, readnoise=rn, cr_threshold=T,
                                             neighbor_threshold=3.0, maxiter=1)
        assert bool(masks[0][10, 10]) == flagged and masks[0].sum() == flagged
