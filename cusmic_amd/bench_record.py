"""Benchmark records in the reference's schema, and their A/B comparison.

``measure`` times the reference's protocol (bench/bench.cu:181-257) through
this library's C ABI: the saved 512 x 512 scene stacked ``frames`` times with
one shared error map, cleaned by ``remove_cosmics_batch_device`` with the
saved scene's settings (1, 5, 5), maxiter 4. Stages, each timed between
device synchronisations on the host clock:

* first_result_ms: the first complete call (upload, clean, download);
* upload_ms: device buffers for the stack and the error map, filled from
  pageable host memory;
* clean_ms: the clean alone, inputs resident;
* download_ms: the cleaned stack and its masks back to the host;
* total_ms: complete calls (upload, clean, download).

Every result is checked against the saved reference pixels and mask outside
the timed calls. The record carries the reference's fields
(bench/bench.cu:297-332), with backend "hip"; ``cuda_runtime`` and
``cuda_driver`` hold the HIP runtime and driver versions, as
``cusmic_cuda_version`` does. ``compare`` applies the reference's A/B rules
(bench/ab.py:38-78): same GPU, software, shape, settings, timing settings and
detected-pixel count, distinct clean Git revisions, reference-exact results.
"""

import ctypes
import json
import math
import os
import statistics
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib, fits

ROOT = Path(__file__).resolve().parent.parent
SCENE = ROOT / "tests" / "golden" / "fits"
FRAMES = (1, 4, 16)
STAGES = ("clean_ms", "total_ms")
SETTINGS = dict(contrast=1.0, cr_threshold=5.0, neighbor_threshold=5.0, maxiter=4)


def _log(*a):
    print("[record]", *a, file=sys.stderr, flush=True)


def _median(values):
    v = sorted(values)
    n = len(v)
    return (v[(n - 1) // 2] + v[n // 2]) / 2


def _stats(samples):
    return {"median": _median(samples), "minimum": min(samples), "maximum": max(samples),
            "samples": list(samples)}


def _revision():
    """(revision, dirty) of the source tree: from Git where there is a
    checkout, else from CUSMIC_REVISION (a tree copied without .git)."""
    try:
        rev = subprocess.run(["git", "rev-parse", "HEAD"], cwd=ROOT, capture_output=True, text=True,
                             check=True).stdout.strip()
        dirty = bool(subprocess.run(["git", "status", "--porcelain", "--untracked-files=no"], cwd=ROOT,
                                    capture_output=True, text=True, check=True).stdout.strip())
        return rev, dirty
    except (OSError, subprocess.CalledProcessError):
        rev = os.environ.get("CUSMIC_REVISION")
        return rev, None if rev is None else os.environ.get("CUSMIC_SOURCE_DIRTY") == "1"


def _driver_version():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        v = ctypes.c_int()
        return int(v.value) if hip.hipDriverGetVersion(ctypes.byref(v)) == 0 else None
    except OSError:
        return None


def load_scene():
    data = np.ascontiguousarray(fits.read_fits(SCENE / "input.fits.gz", dtype="float64")[0])
    error = np.ascontiguousarray(fits.read_fits(SCENE / "error.fits.gz", dtype="float64")[0])
    ref = fits.read_fits(SCENE / "reference.fits.gz", dtype="float64")[0]
    crmask = fits.read_fits(SCENE / "reference.fits.gz", ext="CRMASK")[0].astype(bool)
    return data, error, ref, crmask


def measure(frames, warmups=4, repeats=16):
    """One record of the reference's schema for a stack of ``frames``."""
    import torch

    lib = _lib.lib()
    dev = torch.device("cuda", torch.cuda.current_device())
    data, error, ref, crmask = load_scene()
    h, w = data.shape
    n = data.size
    stack = np.ascontiguousarray(np.broadcast_to(data, (frames, h, w)))
    opts = _lib.default_options()
    opts.contrast, opts.cr_threshold = SETTINGS["contrast"], SETTINGS["cr_threshold"]
    opts.neighbor_threshold, opts.maxiter = SETTINGS["neighbor_threshold"], SETTINGS["maxiter"]
    want_bits = np.ascontiguousarray(ref).view(np.uint64)

    def upload():
        return torch.from_numpy(stack).to(dev), torch.from_numpy(error).to(dev)

    def clean(inputs):
        pixels, err = inputs
        ims = (_lib.CusmicImage * frames)()
        for f in range(frames):
            ims[f] = _lib.CusmicImage(pixels.data_ptr() + 8 * f * n, err.data_ptr(), None, None, None,
                                      None, w, h)
        out = torch.empty(frames * n, dtype=torch.float64, device=dev)
        mask = torch.empty(frames * n, dtype=torch.uint8, device=dev)
        msg = ctypes.create_string_buffer(256)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check(lib.remove_cosmics_batch_device(ims, frames, ctypes.byref(opts), out.data_ptr(),
                                                   mask.data_ptr(), stream, msg, 256), msg)
        return out, mask

    def download(result):
        out, mask = result
        return out.cpu().numpy(), mask.cpu().numpy()

    def complete():
        return download(clean(upload()))

    def timed(call):
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        value = call()
        torch.cuda.synchronize(dev)
        return value, 1e3 * (time.perf_counter() - t0)

    def check(host):
        out, mask = host
        for f in range(frames):
            if not (np.array_equal(out[f * n:(f + 1) * n].view(np.uint64), want_bits.ravel()) and
                    np.array_equal(mask[f * n:(f + 1) * n].astype(bool), crmask.ravel())):
                raise RuntimeError(f"frame {f} differs from the saved reference")

    _log(f"HIP: first result ({frames} frames, {warmups} warmups, {repeats} samples)")
    first, first_ms = timed(complete)
    exact = True
    try:
        check(first)
    except RuntimeError:
        exact = False
    out0 = first[0].reshape(frames, h, w)
    max_abs_error = float(np.max(np.abs(np.where(np.isfinite(out0) & np.isfinite(ref), out0 - ref, 0.0))))
    detected = int(first[1].astype(bool).sum())
    for i in range(warmups):
        complete()
    upload_ms, clean_ms, download_ms, total_ms = [], [], [], []
    for i in range(repeats):
        _, ms = timed(upload)
        upload_ms.append(ms)
    resident = upload()
    for i in range(warmups):
        timed(lambda: clean(resident))
    for i in range(repeats):
        result, ms = timed(lambda: clean(resident))
        clean_ms.append(ms)
        host, ms = timed(lambda: download(result))
        download_ms.append(ms)
        check(host)
    for i in range(repeats):
        host, ms = timed(complete)
        total_ms.append(ms)
        check(host)
    rev, dirty = _revision()
    return {
        "backend": "hip", "dtype": "float64", "shape": ([frames] if frames > 1 else []) + [h, w],
        "warmups": warmups, "repeats": repeats, "first_result_ms": first_ms,
        "host_memory": "pageable", "allocation": "torch caching allocator",
        "detected_pixels": detected,
        "settings": dict(SETTINGS), "source_revision": rev, "source_dirty": dirty,
        "cusmic": lib.cusmic_version().decode(), "gpu": torch.cuda.get_device_name(dev),
        "cuda_runtime": int(lib.cusmic_cuda_version()), "cuda_driver": _driver_version(),
        "reference_mode": "exact", "reference_close": exact, "reference_exact": exact,
        "max_abs_error": max_abs_error, "mask_disagreements": 0 if exact else None,
        "milliseconds": {"upload_ms": _stats(upload_ms), "clean_ms": _stats(clean_ms),
                         "download_ms": _stats(download_ms), "total_ms": _stats(total_ms)},
        "frames_per_second": 1000 * frames / _median(total_ms),
    }


def record(directory, frames=FRAMES, warmups=4, repeats=16):
    """Write hip-<frames>.json for each stack size into ``directory``."""
    directory = Path(directory)
    directory.mkdir(parents=True, exist_ok=True)
    out = []
    for nf in frames:
        rec = measure(nf, warmups, repeats)
        (directory / f"hip-{nf}.json").write_text(json.dumps(rec) + "\n")
        _log(f"{nf} frame(s): clean {rec['milliseconds']['clean_ms']['median'] / nf:.3f} ms/frame, "
             f"total {rec['milliseconds']['total_ms']['median'] / nf:.3f} ms/frame, "
             f"exact {rec['reference_exact']}")
        out.append(rec)
    return out


def load(directory, frames, backend="hip"):
    path = Path(directory) / f"{backend}-{frames}.json"
    rec = json.loads(path.read_text())
    shape = rec["shape"]
    if rec["backend"] != backend or (shape[0] if len(shape) == 3 else 1) != frames:
        raise ValueError(f"wrong case in {path}")
    return rec


def samples(rec, stage):
    values = rec["milliseconds"][stage]["samples"]
    if len(values) != rec["repeats"] or not values or any(not math.isfinite(t) or t <= 0 for t in values):
        raise ValueError("stage samples must match repeats and be positive and finite")
    return values


def compare(baseline, candidate):
    """[(stage, baseline samples, candidate samples)] of two matched records,
    or ValueError naming the first rule the pair breaks."""
    same = ("backend", "dtype", "gpu", "cuda_runtime", "cuda_driver", "shape", "warmups", "repeats",
            "host_memory", "allocation")
    if any(baseline.get(k) != candidate.get(k) for k in same):
        raise ValueError("GPU, software, shape or timing settings differ")
    revs = baseline.get("source_revision"), candidate.get("source_revision")
    if revs[0] == revs[1] or any(not isinstance(r, str) or len(r) not in (40, 64) or
                                 any(c not in "0123456789abcdef" for c in r) for r in revs):
        raise ValueError("A/B comparison needs distinct Git source revisions")
    counts = baseline.get("detected_pixels"), candidate.get("detected_pixels")
    if counts[0] != counts[1] or any(not isinstance(c, int) or isinstance(c, bool) or c < 0 for c in counts):
        raise ValueError("A/B comparison needs the same detected-pixel count")
    if baseline.get("settings") != candidate.get("settings"):
        raise ValueError("algorithm settings differ")
    for rec in (baseline, candidate):
        if rec.get("source_dirty") is True:
            raise ValueError("A/B comparison requires unmodified source trees")
        if rec.get("mask_disagreements", 0) != 0 or rec.get("reference_exact") is not True:
            raise ValueError("reference pixels must match exactly and masks must agree")
    return [(stage, samples(baseline, stage), samples(candidate, stage)) for stage in STAGES]


def report(baseline_dir, candidate_dir, frames=FRAMES):
    """The reference's A/B table (bench/ab.py:78-117) as lines of text."""
    lines = ["Median warmed milliseconds per frame; speedup >1 means candidate is faster",
             "Frames Backend Stage     Baseline Candidate  Speedup"]
    for nf in frames:
        old, new = load(baseline_dir, nf), load(candidate_dir, nf)
        for stage, before, after in compare(old, new):
            a, b = statistics.median(before) / nf, statistics.median(after) / nf
            name = "resident" if stage == "clean_ms" else "ordinary"
            lines.append(f"{nf:>6} {'hip':>7} {name:>8} {a:>12.3f} {b:>8.3f} {a / b:>7.2f}x")
            ma, mb = statistics.mean(before) / nf, statistics.mean(after) / nf
            lines.append(f"       mean: baseline {ma:.3f}; candidate {mb:.3f}; speedup {ma / mb:.3f}x "
                         f"({100 * (mb / ma - 1):+.1f}% time)")

            def spread(v):
                sd = statistics.stdev(v) if len(v) > 1 else 0.0
                return f"{min(v) / nf:.3f}..{max(v) / nf:.3f} (SD {sd / nf:.3f})"

            lines.append(f"       sample range: baseline {spread(before)}; candidate {spread(after)}")
    return lines
