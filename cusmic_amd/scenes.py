"""Seeded synthetic astronomical scenes for the benchmark configurations.

A frame is sky + Gaussian stars (sigma 1.5 px), observed with Poisson noise at
``gain`` e-/ADU plus Gaussian read noise of ``readnoise`` e-, then hit by
cosmic-ray tracks: straight runs of 1..L pixels at a uniform angle in [0, pi),
each pixel adding U(500, 30000) ADU. Everything is generated with torch on the
requested device from an explicit seed, so a scene is reproducible on that
device. lacosmic.utils (the reference's generator, test/mkref.py:12-25) is not
available here; this generator follows the same recipe.
"""

import math

import torch


def _gen(seed, device):
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def stars(shape, n, seed, device, peak=(500.0, 20000.0), sigma=1.5):
    """Sum of n Gaussian stars with uniform positions and peaks."""
    h, w = shape
    g = _gen(seed, device)
    img = torch.zeros(h * w, dtype=torch.float64, device=device)
    if n == 0:
        return img.view(h, w)
    cy = torch.rand(n, generator=g, device=device, dtype=torch.float64) * h
    cx = torch.rand(n, generator=g, device=device, dtype=torch.float64) * w
    amp = peak[0] + torch.rand(n, generator=g, device=device, dtype=torch.float64) * (peak[1] - peak[0])
    r = int(math.ceil(5 * sigma))
    off = torch.arange(-r, r + 1, device=device)
    for start in range(0, n, 4096):
        sl = slice(start, min(n, start + 4096))
        iy = cy[sl].floor().long()[:, None, None] + off[None, :, None]
        ix = cx[sl].floor().long()[:, None, None] + off[None, None, :]
        dy = iy.double() + 0.5 - cy[sl][:, None, None]
        dx = ix.double() + 0.5 - cx[sl][:, None, None]
        val = amp[sl][:, None, None] * torch.exp(-(dx * dx + dy * dy) / (2 * sigma * sigma))
        ok = (iy >= 0) & (iy < h) & (ix >= 0) & (ix < w)
        img.index_add_(0, (iy * w + ix)[ok], val[ok])
    return img.view(h, w)


def cosmic_rays(shape, n, seed, device, max_len=4, flux=(500.0, 30000.0)):
    """Additive image of n straight tracks of 1..max_len pixels."""
    h, w = shape
    g = _gen(seed, device)
    img = torch.zeros(h * w, dtype=torch.float64, device=device)
    if n == 0:
        return img.view(h, w)
    y0 = torch.rand(n, generator=g, device=device, dtype=torch.float64) * h
    x0 = torch.rand(n, generator=g, device=device, dtype=torch.float64) * w
    theta = torch.rand(n, generator=g, device=device, dtype=torch.float64) * math.pi
    length = torch.randint(1, max_len + 1, (n,), generator=g, device=device)
    step = torch.arange(max_len, device=device, dtype=torch.float64)
    py = (y0[:, None] + step[None, :] * torch.sin(theta)[:, None]).floor().long()
    px = (x0[:, None] + step[None, :] * torch.cos(theta)[:, None]).floor().long()
    val = flux[0] + torch.rand((n, max_len), generator=g, device=device, dtype=torch.float64) * (flux[1] - flux[0])
    ok = (step[None, :] < length[:, None]) & (py >= 0) & (py < h) & (px >= 0) & (px < w)
    img.index_add_(0, (py * w + px)[ok], val[ok])
    return img.view(h, w)


def observe(signal, gain, readnoise, seed, device):
    """Poisson(gain * signal) / gain + N(0, readnoise) / gain, in ADU."""
    g = _gen(seed, device)
    electrons = torch.poisson((signal.clamp(min=0) * gain), generator=g)
    read = torch.randn(signal.shape, generator=g, device=device, dtype=torch.float64) * readnoise
    return (electrons + read) / gain


def error_map(signal, gain, readnoise):
    """1-sigma noise in ADU: sqrt(max(signal, 0) / gain + (readnoise / gain)^2)."""
    return torch.sqrt(signal.clamp(min=0) / gain + (readnoise / gain) ** 2)


def frame(shape, seed, device, sky=1000.0, n_stars=40, n_tracks=150, max_len=4, gain=2.0,
          readnoise=5.0, star_seed=None):
    """One frame: returns (data, truth signal without cosmic rays)."""
    s = stars(shape, n_stars, seed if star_seed is None else star_seed, device)
    signal = (sky if torch.is_tensor(sky) else torch.full(shape, float(sky), dtype=torch.float64,
                                                          device=device)) + s
    data = observe(signal, gain, readnoise, seed + 1, device)
    data += cosmic_rays(shape, n_tracks, seed + 2, device, max_len)
    return data, signal


def single_frame_512(device, seed=0):
    """One 512x512 frame, noise-model path (gain 2, read noise 5)."""
    data, _ = frame((512, 512), seed, device)
    return dict(data=data, gain=2.0, readnoise=5.0)


def stack_512(device, n=16, seed=1, shape=(512, 512)):
    """Stack of n frames with a shared error map, 2 bad columns per frame in the
    mask and 0.01% NaN pixels."""
    g = _gen(seed + 1000, device)
    signal = torch.full(shape, 1000.0, dtype=torch.float64, device=device) + stars(shape, 40, seed, device)
    frames = []
    for f in range(n):
        d = observe(signal, 2.0, 5.0, seed + 10 * f + 1, device)
        d += cosmic_rays(shape, 150, seed + 10 * f + 2, device)
        frames.append(d)
    data = torch.stack(frames)
    mask = torch.zeros(data.shape, dtype=torch.bool, device=device)
    cols = torch.randint(0, shape[1], (n, 2), generator=g, device=device)
    for f in range(n):
        mask[f, :, cols[f]] = True
    nan = torch.rand(data.shape, generator=g, device=device, dtype=torch.float64) < 1e-4
    data[nan] = float("nan")
    return dict(data=data, error=error_map(signal, 2.0, 5.0), mask=mask)


def detector_frame(device, seed=2, shape=(4096, 4096)):
    """Detector-size frame: sky gradient 200..3000, 2000 stars, 8000 tracks of
    1..12 px, per-pixel error map, scalar background 150, 0.5% masked."""
    h, w = shape
    g = _gen(seed + 1000, device)
    ramp = torch.linspace(200.0, 3000.0, w, dtype=torch.float64, device=device)
    sky = ramp[None, :].expand(h, w).clone()
    signal = sky + stars(shape, 2000, seed, device)
    data = observe(signal, 2.0, 5.0, seed + 1, device)
    data += cosmic_rays(shape, 8000, seed + 2, device, max_len=12)
    mask = torch.rand(shape, generator=g, device=device, dtype=torch.float64) < 5e-3
    return dict(data=data, error=error_map(signal, 2.0, 5.0), mask=mask, background=150.0)


RAGGED_SHAPES = ((256, 256), (512, 2048), (1024, 1024), (2048, 4096), (4096, 4096))


def ragged_batch(device, n=64, seed=3, shapes=RAGGED_SHAPES, density=2e-4):
    """Ragged batch: shapes drawn uniformly from ``shapes``, per-frame seeds,
    per-frame error maps, 2e-4 tracks/px, gain in [1,4], read noise in [2,10]."""
    g = _gen(seed, "cpu")
    pick = torch.randint(0, len(shapes), (n,), generator=g)
    gains = 1 + 3 * torch.rand(n, generator=g, dtype=torch.float64)
    rns = 2 + 8 * torch.rand(n, generator=g, dtype=torch.float64)
    out = []
    for f in range(n):
        shape = shapes[int(pick[f])]
        npx = shape[0] * shape[1]
        gain, rn = float(gains[f]), float(rns[f])
        s = stars(shape, max(1, npx // 8000), seed + 100 * f, device)
        signal = torch.full(shape, 1000.0, dtype=torch.float64, device=device) + s
        data = observe(signal, gain, rn, seed + 100 * f + 1, device)
        data += cosmic_rays(shape, int(round(density * npx)), seed + 100 * f + 2, device, max_len=8)
        out.append(dict(data=data, error=error_map(signal, gain, rn), gain=gain, readnoise=rn))
    return out


def adjoint_batch(device, n=8, seed=4, shape=(1024, 1024)):
    """Batch of n noise-model frames (the single-frame generator) plus seeded
    weights."""
    frames = [frame(shape, seed + 10 * f, device)[0] for f in range(n)]
    g = _gen(seed + 999, device)
    data = torch.stack(frames)
    weights = torch.randn(data.shape, generator=g, device=device, dtype=torch.float64)
    return dict(data=data, gain=2.0, readnoise=5.0, weights=weights)
