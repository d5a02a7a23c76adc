"""Scenes whose detections follow in closed form from the noise model.

The L.A.Cosmic noise model (van Dokkum 2001, eq. 10; reference
src/filters.cuh:107-123 and mod/cusmic/filters.py:81-91) is

    N = sqrt(rn^2 + g * max(med5(C), 1e-5)) / g

with med5 the 5x5 median of the clean image. On a flat frame of value c with
one pixel raised by a > 0, the subsampled Laplacian is 2a at the raised pixel
and 0 everywhere else, every 5x5 median of C is c and every 5x5 median of S is
0, so S' = a / N at the raised pixel, F is at its floor (0.01) and the pixel
is flagged exactly when a / N > cr_threshold. A raised pixel placed just above
or just below t * N therefore flags or not as the formula says, and a
misread gain, read noise, floor, factor or map index moves the flip.

The plateau scenes pin the median size: tiny gain makes N depend strongly
on the local median while the plateau's own Laplacian stays far below the
thresholds, and the plateau shape makes the 3x3, 5x5 and 7x7 medians differ
at the raised pixel. Medians come from scipy, not from the oracle.
"""

import numpy as np
from scipy import ndimage

SIGNAL_FLOOR = 1e-5


def noise(m, g, rn):
    """The published noise model at median m."""
    return np.sqrt(rn * rn + g * np.maximum(m, SIGNAL_FLOOR)) / g


def raised_pixels(shape, c, g, rn, t, seed, spacing=9, margin=1e-6):
    """Flat frame c with raised pixels on a grid, alternately just above and
    just below the flag amplitude t * N of their own (g, rn) (scalars or
    maps). Returns (data, expected mask)."""
    h, w = shape
    rng = np.random.default_rng(seed)
    data = np.full(shape, float(c))
    want = np.zeros(shape, bool)
    gm = np.broadcast_to(g, shape)
    rm = np.broadcast_to(rn, shape)
    k = 0
    for y in range(4, h - 4, spacing):
        for x in range(4, w - 4, spacing):
            above = (k + rng.integers(0, 2)) % 2 == 0
            k += 1
            n = noise(c, gm[y, x], rm[y, x])
            data[y, x] += t * n * (1 + margin if above else 1 - margin)
            want[y, x] = above
    return data, want


def plateau(kind, shape=(31, 33), c0=1e-5, step=0.01):
    """A plateau of height step around the centre. kind "X": a 3x3 plateau, so
    that med3 = c0 + step but med5 = med7 = c0 at the centre. kind "Y": the 5x5
    square without its corners, so that med3 = med5 = c0 + step but med7 = c0.
    Returns (data, centre y, centre x)."""
    data = np.full(shape, c0)
    cy, cx = shape[0] // 2, shape[1] // 2
    if kind == "X":
        data[cy - 1:cy + 2, cx - 1:cx + 2] += step
    else:
        data[cy - 2:cy + 3, cx - 2:cx + 3] += step
        for dy in (-2, 2):
            for dx in (-2, 2):
                data[cy + dy, cx + dx] = c0
    return data, cy, cx


def plateau_medians(data, cy, cx):
    """The 3x3, 5x5 and 7x7 medians of data at (cy, cx), from scipy."""
    return {s: float(ndimage.median_filter(data, size=s, mode="mirror")[cy, cx]) for s in (3, 5, 7)}
