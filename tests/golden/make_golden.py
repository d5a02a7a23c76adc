"""Write the stage golden vectors in tests/golden/stages.npz.

Inputs are seeded random frames (plus ties and exact zeros); expected outputs
come from numpy/scipy, the libraries L.A.Cosmic 1.4.0 builds its stages on:
  * median_filter(size=3/5/7, mode=...)        -> medians, every border mode
  * convolve(block_replicate(x)/4, laplace) clipped, 2x2 np.sum block reduce
                                               -> Laplacian in scipy's order
  * binary_dilation(3x3)                      -> growth
Run: python tests/golden/make_golden.py   (numpy 2.2, scipy 1.15 here)
"""

from pathlib import Path

import numpy as np
import scipy
from scipy import ndimage

MODES = ("reflect", "constant", "nearest", "mirror", "wrap")
SHAPES = ((37, 53), (1, 9), (9, 1), (2, 2), (3, 3), (2, 7), (6, 5), (1, 1), (4, 13))
KERNEL = np.array([[0, -1, 0], [-1, 4, -1], [0, -1, 0]], dtype=np.float64)


def lacosmic_laplacian(image, mode):
    """Laplacian as L.A.Cosmic computes it: replicate 2x2 conserving the sum,
    convolve, clip at 0, and sum 2x2 blocks (astropy block_reduce -> np.sum)."""
    h, w = image.shape
    sampled = np.repeat(np.repeat(image / 4.0, 2, axis=0), 2, axis=1)
    lap2 = ndimage.convolve(sampled, KERNEL, mode=mode).clip(min=0.0)
    return np.sum(lap2.reshape(h, 2, w, 2), axis=(1, 3))


def frame(rng, shape, kind):
    if kind == "ties":
        return rng.integers(-3, 4, size=shape).astype(np.float64)
    if kind == "zeros":
        a = rng.normal(size=shape)
        a[rng.random(shape) < 0.3] = 0.0
        a[rng.random(shape) < 0.1] = -0.0
        return a
    return rng.normal(1000.0, 30.0, size=shape) + (rng.random(shape) < 0.05) * 5000.0


def main():
    rng = np.random.default_rng(20261203)
    out = {"versions": np.array([np.__version__, scipy.__version__])}
    for s, shape in enumerate(SHAPES):
        for kind in ("noise", "ties", "zeros"):
            x = frame(rng, shape, kind)
            key = f"{s}_{kind}"
            out[f"in_{key}"] = x
            for mode in MODES:
                for size in (3, 5, 7):
                    out[f"med{size}_{mode}_{key}"] = ndimage.median_filter(x, size=size, mode=mode)
                out[f"lap_{mode}_{key}"] = lacosmic_laplacian(x, mode)
            b = rng.random(shape) < 0.2
            out[f"bin_{key}"] = b
            out[f"dil_{key}"] = ndimage.binary_dilation(b, np.ones((3, 3), bool))
    np.savez_compressed(Path(__file__).with_name("stages.npz"), **out)


if __name__ == "__main__":
    main()
