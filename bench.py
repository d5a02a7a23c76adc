#!/usr/bin/env python3
"""Benchmark of the L.A.Cosmic cleaning loop on one MI355X.

Workload (default): the ragged batch -- 64 float64 frames with shapes drawn
from {256x256, 512x2048, 1024x1024, 2048x4096, 4096x4096}, per-frame seeds and
error maps, 2e-4 cosmic-ray tracks/px, settings (3, 5, 3), maxiter 4, mirror
borders -- cleaned in one call of remove_cosmics_ragged_device with inputs
resident in HBM. A step is one such call. Also reported:
  * per_frame: the same 64 frames as 64 separate plain-path calls;
  * fwd_bwd: 8 x 1024^2 noise-model frames, forward with donor recording then
    the adjoint of loss = sum(w * cleaned), vs forward alone;
  * roofline: the dense significance kernel's algorithmic bytes (8 B of clean
    image + 8 B of error map per processed pixel) over its event-timed
    duration; call_roofline: the whole call's algorithmic bytes
    (17 + M + I_d (8 + E) + I_r 8 per pixel) over the step time;
  * cpu_baseline: the oracle (the C restatement, one thread) on a bounded
    sample of the same workload; cpu_baseline_all_cores: the same oracle on
    every host core (OpenMP over rows) on up to four 1 Mpx frames.
Prints one JSON line on stdout; progress goes to stderr.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "cleaned Mpix/s fp64 maxiter=4: ragged-batch vs per-frame; fwd+bwd Mpix/s"
HBM_PEAK_GBS = 8000.0
# first detection pass, per frame pixel: reads the data and the error map,
# writes the working image and the flags
FIRST_BYTES_PER_PX = 8.0 + 8.0 + 8.0 + 1.0
TRAFFIC_FILE = ROOT / "profiles" / "significance_maps_ragged64_traffic.json"


def significance_traffic(launches):
    """HBM bytes per first-pass launch from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE of the same workload, scaled by a copy kernel of
    known size), or None when there is none."""
    try:
        rec = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None, None
    if rec.get("launches") != launches:
        return None, None
    return round(rec["hbm_bytes_per_launch"]), TRAFFIC_FILE.name


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_info():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--no-per-frame", action="store_true")
    ap.add_argument("--no-backward", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="start with a 1 GiB device-to-device copy, whose hardware counter "
                         "readings scale FETCH_SIZE/WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; on several GPUs the frames of a batch would be "
              "sharded by pixel count, with no exchange between GPUs", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from cusmic_amd import _lib, engine, scenes

    dev = torch.device("cuda", 0)
    _lib.lib()
    if args.calibration_copy:
        a = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        torch.empty_like(a).copy_(a)
        torch.cuda.synchronize()
        del a
    log("generating the ragged batch on the GPU")
    batch = scenes.ragged_batch(dev, n=args.frames)
    fr = engine.Frames()
    for item in batch:
        fr.add(item["data"], error=item["error"])
    total = fr.total
    mpix = total / 1e6
    clean = torch.empty(total, dtype=torch.float64, device=dev)
    mask = torch.empty(total, dtype=torch.uint8, device=dev)
    log(f"{len(batch)} frames, {mpix:.1f} Mpx")

    def step(profile=None):
        return engine.run(fr, dev, 3.0, 5.0, 3.0, 4, _lib.MIRROR, clean=clean, mask=mask, profile=profile)

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup}")
    profiles = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        prof = {}
        step(prof)
        profiles.append(prof)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = 1e3 * elapsed / args.steps
    value = mpix / (ms_step / 1e3)
    log(f"ragged: {ms_step:.3f} ms/step, {value:.1f} Mpix/s")
    flagged = int(mask.sum())

    def mean(key):
        return float(np.mean([p[key] for p in profiles]))

    # dominant kernel: the dense first detection pass (significance_maps, one
    # launch per call) over every frame pixel, timed with hipEvents on its
    # stream
    sig_ms, launches = mean("first_pass_ms"), 1
    sig_bytes = FIRST_BYTES_PER_PX * total
    achieved = sig_bytes / (sig_ms / 1e3) / 1e9
    traffic, traffic_note = significance_traffic(launches)
    # whole call, the dense loop's algorithmic bytes per frame pixel:
    # B = 17 + M + I_d (8 + E) + I_r 8 with M = 8 (error map read once) and
    # E = 8 (its re-read per detection pass); I_d / I_r are the detection and
    # replacement passes each frame ran. The incremental passes touch only
    # the pixels near changes, so the figure can exceed 100%.
    call_bytes = 25.0 * total + 16.0 * mean("detect_pass_pixels") + 8.0 * mean("replace_pass_pixels")
    call_gbs = call_bytes / (ms_step / 1e3) / 1e9
    result = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ragged batch of {len(batch)} float64 frames (256^2..4096^2), "
                               "per-frame error maps, 2e-4 CR tracks/px, contrast 3, thresholds 5/3, "
                               "maxiter 4, mirror, one remove_cosmics_ragged_device call",
                   "frames": len(batch), "megapixels": round(mpix, 3), "maxiter": 4,
                   "flagged_pixels": flagged,
                   "detection_passes_per_pixel": round(mean("detect_pass_pixels") / total, 4),
                   "replacement_passes_per_pixel": round(mean("replace_pass_pixels") / total, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": "significance_maps",
                     "algorithmic_bytes_per_launch": round(sig_bytes / launches),
                     "avg_launch_ms": round(sig_ms / launches, 4), "launches": launches,
                     "traffic_source": traffic_note},
        "call_roofline": {"bytes_per_pixel": round(call_bytes / total, 3),
                          "achieved": round(call_gbs, 1), "unit": "GB/s",
                          "frac": round(call_gbs / HBM_PEAK_GBS, 4)},
        "stages_ms": {k: round(mean(k), 4) for k in
                      ("prologue_ms", "first_pass_ms", "significance_ms", "detect_ms", "replace_ms",
                       "epilogue_ms")},
        "lists_per_call": {k: int(profiles[0][k]) for k in ("seeds", "candidates", "grown", "widened")},
        "block_pass_pixels": int(mean("detect_pixels") - total),
    }

    if not args.no_per_frame:
        log("per-frame plain path")
        singles = []
        for item in batch:
            f1 = engine.Frames()
            f1.add(item["data"], error=item["error"])
            singles.append(f1)

        def per_frame():
            for f1 in singles:
                engine.run(f1, dev, 3.0, 5.0, 3.0, 4, _lib.MIRROR)
        per_frame()
        torch.cuda.synchronize()
        reps = max(1, args.steps // 4)
        t0 = time.perf_counter()
        for _ in range(reps):
            per_frame()
        torch.cuda.synchronize()
        pf_ms = 1e3 * (time.perf_counter() - t0) / reps
        # outside the timed region: every frame's plain-path output equals its
        # slice of the ragged call's output, bit for bit
        same, off = 0, 0
        for f1 in singles:
            n1 = f1.total
            c1, m1, _, _ = engine.run(f1, dev, 3.0, 5.0, 3.0, 4, _lib.MIRROR)
            same += bool(torch.equal(c1.view(torch.int64), clean[off:off + n1].view(torch.int64))
                         and torch.equal(m1, mask[off:off + n1]))
            off += n1
        result["per_frame"] = {"value": round(mpix / (pf_ms / 1e3), 2), "unit": "Mpix/s",
                               "ms_per_batch": round(pf_ms, 3),
                               "ragged_speedup": round(pf_ms / ms_step, 3),
                               "outputs_equal_ragged": same == len(singles),
                               "frames_equal": same}
        log(f"per-frame: {pf_ms:.3f} ms/batch; {same}/{len(singles)} frames equal the ragged output")
        del singles

    if not args.no_host:
        # the host entry point on the batch's largest frame: validation,
        # allocation and pageable copies both ways included (never `value`)
        import ctypes
        big = max(batch, key=lambda it: it["data"].numel())
        d = np.ascontiguousarray(big["data"].cpu().numpy())
        e = np.ascontiguousarray(big["error"].cpu().numpy())
        hh, ww = d.shape
        out_h = np.empty_like(d)
        mask_h = np.empty(d.shape, np.uint8)
        img = _lib.CusmicImage(d.ctypes.data, e.ctypes.data, None, None, None, None, ww, hh)
        opts = _lib.default_options()
        msg = ctypes.create_string_buffer(256)

        def host_call():
            _lib.check(_lib.lib().remove_cosmics(ctypes.byref(img), ctypes.byref(opts), out_h.ctypes.data,
                                                 mask_h.ctypes.data, msg, 256), msg)
        log(f"host entry point, {hh}x{ww}")
        host_call()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            host_call()
        host_ms = 1e3 * (time.perf_counter() - t0) / reps
        result["host_call"] = {"value": round(d.size / 1e6 / (host_ms / 1e3), 2), "unit": "Mpix/s",
                               "ms_per_call": round(host_ms, 3), "frame": f"{hh}x{ww}, error map",
                               "note": "remove_cosmics on host arrays: validation, allocation and "
                                       "pageable copies to and from the device included"}
        log(f"host call: {host_ms:.3f} ms")
        del d, e, out_h, mask_h

        # complete calls on one 512x512 frame with an error map of 5 and
        # settings (1, 5, 5), maxiter 4, mirror: the shape and settings of the
        # reference's published complete-call runs (README.md:123-140), on a
        # synthetic scene; 4 warmups, median of 16 (bench/bench.cu:204-257)
        d, _ = scenes.frame((512, 512), 11, "cpu")
        d = np.ascontiguousarray(d.numpy())
        e = np.full_like(d, 5.0)
        out_h = np.empty_like(d)
        mask_h = np.empty(d.shape, np.uint8)
        img = _lib.CusmicImage(d.ctypes.data, e.ctypes.data, None, None, None, None, 512, 512)
        opts = _lib.default_options()
        opts.contrast, opts.cr_threshold, opts.neighbor_threshold = 1.0, 5.0, 5.0
        for _ in range(4):
            host_call()
        samples = []
        for _ in range(16):
            t0 = time.perf_counter()
            host_call()
            samples.append(1e3 * (time.perf_counter() - t0))
        ms512 = float(np.median(samples))
        result["host_call_512"] = {"value": round(0.262144 / (ms512 / 1e3), 2), "unit": "Mpix/s",
                                   "ms_per_frame": round(ms512, 3),
                                   "frame": "512x512 synthetic, error map 5, settings (1, 5, 5), maxiter 4",
                                   "note": "complete host calls (validation, allocation, copies); "
                                           "median of 16 after 4 warmups"}
        log(f"host call 512x512: {ms512:.3f} ms")
        del d, e, out_h, mask_h

    if not args.no_backward:
        log("forward+backward, 8 x 1024^2 noise-model frames")
        from cusmic_amd import autograd

        adj = scenes.adjoint_batch(dev)
        data, w = adj["data"], adj["weights"]
        mp4 = data.numel() / 1e6

        def fwd():
            x = data.clone().requires_grad_(True)
            out, _ = autograd.remove_cosmics(x, 3, 5, 3, effective_gain=2.0, readnoise=5.0)
            return x, out

        def fwd_bwd():
            x, out = fwd()
            (w * out).sum().backward()
            return x.grad

        for _ in range(2):
            fwd_bwd()
        torch.cuda.synchronize()
        reps = max(2, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(reps):
            fwd()
        torch.cuda.synchronize()
        f_ms = 1e3 * (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            fwd_bwd()
        torch.cuda.synchronize()
        fb_ms = 1e3 * (time.perf_counter() - t0) / reps
        result["fwd_bwd"] = {"value": round(mp4 / (fb_ms / 1e3), 2), "unit": "Mpix/s",
                             "forward_mpix_s": round(mp4 / (f_ms / 1e3), 2),
                             "ms_forward": round(f_ms, 3), "ms_forward_backward": round(fb_ms, 3),
                             "ratio": round(fb_ms / f_ms, 3), "frames": "8 x 1024x1024"}
        log(f"fwd {f_ms:.3f} ms, fwd+bwd {fb_ms:.3f} ms")

    if not args.no_cpu:
        import oracle

        cores, model = cpu_info()
        # the first 1 Mpx frame of the batch, and where its pixels sit in the
        # ragged call's output
        offs = np.cumsum([0] + [b["data"].numel() for b in batch])
        pick = next((k for k, b in enumerate(batch) if b["data"].numel() == 1 << 20), 0)
        sample = [batch[pick]]
        frames = [dict(data=b["data"].cpu().numpy(), error=b["error"].cpu().numpy()) for b in sample]
        spx = sum(f["data"].size for f in frames) / 1e6
        log(f"cpu baseline: oracle, 1 thread, {spx:.2f} Mpx")
        t0 = time.perf_counter()
        o_clean, o_mask, _, _ = oracle.clean_frames(frames, threads=1)
        c_s = time.perf_counter() - t0
        # outside the timed regions: the ragged call's frame equals the oracle's
        lo, hi = int(offs[pick]), int(offs[pick + 1])
        g_clean = clean[lo:hi].cpu().numpy().reshape(o_clean[0].shape)
        g_mask = mask[lo:hi].cpu().numpy().reshape(o_mask[0].shape).astype(bool)
        oracle_equal = bool((g_clean.view(np.uint64) == o_clean[0].view(np.uint64)).all()
                            and (g_mask == o_mask[0]).all())
        shape = "x".join(map(str, frames[0]["data"].shape))
        result["cpu_baseline"] = {"value": round(spx / c_s, 4), "unit": "Mpix/s", "cores": 1,
                                  "kind": "port", "cpu": model, "host_cores": cores,
                                  "sample": f"frame {pick} ({shape}) of the same ragged batch "
                                            f"(seed 3), error map, maxiter 4; {c_s:.1f} s",
                                  "gpu_output_equals_oracle": oracle_equal}
        log(f"cpu: oracle equals the ragged call's frame {pick}: {oracle_equal}")
        log(f"cpu: {spx / c_s:.3f} Mpix/s")
        # the same oracle on every host core (OpenMP over rows), on up to four
        # 1 Mpx frames of the batch
        wide = [b for b in batch if b["data"].numel() == 1 << 20][:4] or batch[:1]
        wframes = [dict(data=b["data"].cpu().numpy(), error=b["error"].cpu().numpy()) for b in wide]
        wpx = sum(f["data"].size for f in wframes) / 1e6
        log(f"cpu baseline: oracle, {cores} threads, {wpx:.2f} Mpx")
        t0 = time.perf_counter()
        oracle.clean_frames(wframes, threads=cores)
        a_s = time.perf_counter() - t0
        result["cpu_baseline_all_cores"] = {
            "value": round(wpx / a_s, 4), "unit": "Mpix/s", "cores": cores, "kind": "port",
            "cpu": model,
            "sample": f"{len(wframes)} frame(s) of "
                      f"{'x'.join(map(str, wframes[0]['data'].shape))} from the same ragged batch, "
                      f"error map, maxiter 4, OpenMP over rows; {a_s:.1f} s"}
        log(f"cpu, {cores} threads: {wpx / a_s:.3f} Mpix/s")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
