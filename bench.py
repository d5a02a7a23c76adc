#!/usr/bin/env python3
"""Benchmark of the L.A.Cosmic cleaning loop on one MI355X.

Workload (the headline): the ragged batch -- 64 float64 frames with shapes
drawn from {256x256, 512x2048, 1024x1024, 2048x4096, 4096x4096}, per-frame
seeds and error maps, 2e-4 cosmic-ray tracks/px, settings (3, 5, 3), maxiter
4, mirror borders -- cleaned in one call of remove_cosmics_ragged_device with
inputs resident in HBM. A step is one such call. Also reported, each outside
the timed steps:
  * per_frame: the same 64 frames as 64 separate plain-path calls, whose
    outputs must equal the ragged call's bit for bit;
  * detector: one 4096x4096 frame with an error map, background 150 and 0.5%
    masked pixels (the detector-size configuration), with the roofline of its
    first pass;
  * host_call / host_call_512: the host entry point remove_cosmics on the
    batch's largest frame, and on the reference's saved 512x512 scene with
    its settings (1, 5, 5) -- complete calls, as the reference times them;
  * fwd_bwd: 8 x 1024^2 noise-model frames, forward with donor recording then
    the adjoint of loss = sum(w * cleaned), against the forward alone;
  * roofline: the dominant kernel (the first detection pass) -- its
    algorithmic bytes over its hipEvent time; call_roofline: the whole
    call's algorithmic bytes (17 + M + I_d (8 + E) + I_r 8 per pixel);
  * cpu_baseline: the oracle (the C restatement, one thread) on one 1 Mpx frame
    of the batch, checked against the GPU's output for that frame;
    cpu_baseline_all_cores: the same oracle on every host core on four frames.
Prints one JSON line on stdout; progress goes to stderr. ``--only LEG``
(detector or fwd-bwd) runs one leg alone, for profiling. ``--record DIR``
instead writes records in the reference's schema (bench/bench.cu:297-332) for
its saved scene stacked 1, 4 and 16 times, and ``--ab BASELINE CANDIDATE``
compares two such directories under the reference's A/B rules (bench/ab.py).
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
# first detection pass, per frame pixel: reads the data and the error map and
# writes the working image (the flags plane is zeroed by a streaming fill
# kernel; the pass writes flags and S' only at the ~1% of pixels it screens),
# plus a read of the input mask where there is one
FIRST_BYTES_PER_PX = 8.0 + 8.0 + 8.0
TRAFFIC_FILE = ROOT / "profiles" / "significance_maps_ragged64_traffic.json"
DETECTOR_TRAFFIC_FILE = ROOT / "profiles" / "significance_maps_detector4096_traffic.json"
SAVED_SCENE = ROOT / "tests" / "golden" / "fits"


def significance_traffic(launches, path=TRAFFIC_FILE):
    """HBM bytes per first-pass launch from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE of the same workload, scaled by a copy kernel of
    known size), or None when there is none."""
    try:
        rec = json.loads(path.read_text())
    except (OSError, ValueError):
        return None, None
    if rec.get("launches") != launches:
        return None, None
    return round(rec["hbm_bytes_per_launch"]), path.name


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


def roofline(bytes_per_launch, ms, launches=1, traffic=None, source=None, kernel="significance_maps"):
    achieved = bytes_per_launch / (ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
            "algorithmic_bytes_per_launch": round(bytes_per_launch), "avg_launch_ms": round(ms, 4),
            "launches": launches, "traffic_source": source}


class Bench:
    def __init__(self, args):
        import numpy as np
        import torch

        from cusmic_amd import _lib, engine, scenes

        self.args, self.np, self.torch = args, np, torch
        self.lib, self.engine, self.scenes = _lib, engine, scenes
        self.dev = torch.device("cuda", 0)
        _lib.lib()

    def run(self, fr, clean=None, mask=None, profile=None, **kw):
        return self.engine.run(fr, self.dev, 3.0, 5.0, 3.0, 4, self.lib.MIRROR, clean=clean, mask=mask,
                               profile=profile, **kw)

    # ------------------------------------------------------------ ragged
    def ragged(self):
        np, torch, args = self.np, self.torch, self.args
        log("generating the ragged batch on the GPU")
        self.batch = batch = self.scenes.ragged_batch(self.dev, n=args.frames)
        fr = self.engine.Frames()
        for item in batch:
            fr.add(item["data"], error=item["error"])
        total = self.total = fr.total
        self.mpix = mpix = total / 1e6
        self.clean = clean = torch.empty(total, dtype=torch.float64, device=self.dev)
        self.mask = mask = torch.empty(total, dtype=torch.uint8, device=self.dev)
        log(f"{len(batch)} frames, {mpix:.1f} Mpx")
        for i in range(args.warmup):
            self.run(fr, clean, mask)
            log(f"warmup {i + 1}/{args.warmup}")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            self.run(fr, clean, mask)
        torch.cuda.synchronize()
        self.ms_step = ms_step = 1e3 * (time.perf_counter() - t0) / args.steps
        value = mpix / (ms_step / 1e3)
        log(f"ragged: {ms_step:.3f} ms/step, {value:.1f} Mpix/s")
        # the stage timings and the first pass's launch time come from
        # separate calls that record events between the stages, so that the
        # timed calls above carry no event records
        profiles = []
        for _ in range(max(3, min(args.steps, 10))):
            prof = {}
            self.run(fr, clean, mask, prof)
            profiles.append(prof)
        torch.cuda.synchronize()

        def mean(key):
            return float(np.mean([p[key] for p in profiles]))

        # the dominant kernel: the first detection pass (one launch per call),
        # timed with hipEvents on its stream
        sig_ms = mean("first_pass_ms")
        traffic, source = significance_traffic(1)
        # the whole call, the dense loop's algorithmic bytes per frame pixel:
        # B = 17 + M + I_d (8 + E) + I_r 8 with M = 8 (error map read once) and
        # E = 8 (its re-read per detection pass); I_d / I_r are the passes each
        # frame ran. The incremental passes touch only the pixels near changes,
        # so the figure can exceed 100%.
        call_bytes = 25.0 * total + 16.0 * mean("detect_pass_pixels") + 8.0 * mean("replace_pass_pixels")
        call_gbs = call_bytes / (ms_step / 1e3) / 1e9
        return {
            "metric": METRIC, "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"ragged batch of {len(batch)} float64 frames (256^2..4096^2), "
                                   "per-frame error maps, 2e-4 CR tracks/px, contrast 3, thresholds 5/3, "
                                   "maxiter 4, mirror, one remove_cosmics_ragged_device call",
                       "frames": len(batch), "megapixels": round(mpix, 3), "maxiter": 4,
                       "flagged_pixels": int(mask.sum()),
                       "detection_passes_per_pixel": round(mean("detect_pass_pixels") / total, 4),
                       "replacement_passes_per_pixel": round(mean("replace_pass_pixels") / total, 4)},
            "roofline": roofline(FIRST_BYTES_PER_PX * total, sig_ms, 1, traffic, source),
            "call_roofline": {"bytes_per_pixel": round(call_bytes / total, 3), "achieved": round(call_gbs, 1),
                              "unit": "GB/s", "frac": round(call_gbs / HBM_PEAK_GBS, 4)},
            "stages_ms": {k: round(mean(k), 4) for k in
                          ("prologue_ms", "first_pass_ms", "significance_ms", "detect_ms", "replace_ms",
                           "epilogue_ms")},
            "lists_per_call": {k: int(profiles[0][k]) for k in ("seeds", "candidates", "grown", "widened")},
            "block_pass_pixels": int(mean("detect_pixels") - total),
        }

    # --------------------------------------------------------- per frame
    def per_frame(self):
        torch = self.torch
        log("per-frame plain path")
        singles = []
        for item in self.batch:
            f1 = self.engine.Frames()
            f1.add(item["data"], error=item["error"])
            singles.append(f1)

        def all_frames():
            for f1 in singles:
                self.run(f1)
        all_frames()
        torch.cuda.synchronize()
        reps = max(1, self.args.steps // 4)
        t0 = time.perf_counter()
        for _ in range(reps):
            all_frames()
        torch.cuda.synchronize()
        pf_ms = 1e3 * (time.perf_counter() - t0) / reps
        # outside the timed region: every frame's plain-path output equals its
        # slice of the ragged call's output, bit for bit
        same, off = 0, 0
        for f1 in singles:
            n1 = f1.total
            c1, m1, _, _ = self.run(f1)
            same += bool(torch.equal(c1.view(torch.int64), self.clean[off:off + n1].view(torch.int64))
                         and torch.equal(m1, self.mask[off:off + n1]))
            off += n1
        log(f"per-frame: {pf_ms:.3f} ms/batch; {same}/{len(singles)} frames equal the ragged output")
        return {"value": round(self.mpix / (pf_ms / 1e3), 2), "unit": "Mpix/s", "ms_per_batch": round(pf_ms, 3),
                "ragged_speedup": round(pf_ms / self.ms_step, 3), "outputs_equal_ragged": same == len(singles),
                "frames_equal": same}

    # ---------------------------------------------------------- detector
    def detector(self):
        np, torch = self.np, self.torch
        log("detector frame: 4096x4096, error map, background 150, 0.5% masked")
        sc = self.scenes.detector_frame(self.dev)
        fr = self.engine.Frames()
        fr.add(sc["data"], error=sc["error"], mask=sc["mask"])
        n = fr.total
        clean = torch.empty(n, dtype=torch.float64, device=self.dev)
        mask = torch.empty(n, dtype=torch.uint8, device=self.dev)
        for _ in range(3):
            self.run(fr, clean, mask, background=sc["background"])
        profiles, samples = [], []
        for _ in range(max(5, self.args.steps)):
            prof = {}
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            self.run(fr, clean, mask, prof, background=sc["background"])
            torch.cuda.synchronize()
            samples.append(1e3 * (time.perf_counter() - t0))
            profiles.append(prof)
        ms = float(np.median(samples))
        first = float(np.median([p["first_pass_ms"] for p in profiles]))
        log(f"detector: {ms:.3f} ms per call, first pass {first:.3f} ms")
        # first pass per pixel: data, error map and input mask read, working
        # image written (the scalar background costs no bytes)
        traffic, source = significance_traffic(1, DETECTOR_TRAFFIC_FILE)
        return {"value": round(n / 1e6 / (ms / 1e3), 2), "unit": "Mpix/s", "ms_per_call": round(ms, 4),
                "frame": "4096x4096, error map, background 150, 0.5% masked", "flagged_pixels": int(mask.sum()),
                "first_pass_roofline": roofline((FIRST_BYTES_PER_PX + 1.0) * n, first, 1, traffic, source)}

    # -------------------------------------------------------------- host
    def host(self):
        import ctypes

        np, lib = self.np, self.lib
        from cusmic_amd import fits

        def host_call(img, opts, out, flags):
            msg = ctypes.create_string_buffer(256)
            lib.check(lib.lib().remove_cosmics(ctypes.byref(img), ctypes.byref(opts), out.ctypes.data,
                                               flags.ctypes.data, msg, 256), msg)

        result = {}
        # the host entry point on the batch's largest frame: validation,
        # allocation and copies both ways included (never `value`)
        big = max(self.batch, key=lambda it: it["data"].numel())
        d = np.ascontiguousarray(big["data"].cpu().numpy())
        e = np.ascontiguousarray(big["error"].cpu().numpy())
        hh, ww = d.shape
        out, flags = np.empty_like(d), np.empty(d.shape, np.uint8)
        img = lib.CusmicImage(d.ctypes.data, e.ctypes.data, None, None, None, None, ww, hh)
        opts = lib.default_options()
        log(f"host entry point, {hh}x{ww}")
        host_call(img, opts, out, flags)
        t0 = time.perf_counter()
        for _ in range(3):
            host_call(img, opts, out, flags)
        ms = 1e3 * (time.perf_counter() - t0) / 3
        result["host_call"] = {"value": round(d.size / 1e6 / (ms / 1e3), 2), "unit": "Mpix/s",
                               "ms_per_call": round(ms, 3), "frame": f"{hh}x{ww}, error map",
                               "note": "remove_cosmics on host arrays: validation, allocation and copies "
                                       "to and from the device included"}
        log(f"host call: {ms:.3f} ms")
        # complete calls on the reference's saved 512x512 scene, error map 5,
        # settings (1, 5, 5), maxiter 4, mirror: the inputs of its only
        # published timings; first call apart, then 4 warmups and the median
        # of 16 (bench/bench.cu:204-257); output checked outside the timing
        d = np.ascontiguousarray(fits.read_fits(SAVED_SCENE / "input.fits.gz", dtype="float64")[0])
        e = np.ascontiguousarray(fits.read_fits(SAVED_SCENE / "error.fits.gz", dtype="float64")[0])
        ref = fits.read_fits(SAVED_SCENE / "reference.fits.gz")[0]
        crmask = fits.read_fits(SAVED_SCENE / "reference.fits.gz", ext="CRMASK")[0].astype(bool)
        out, flags = np.empty_like(d), np.empty(d.shape, np.uint8)
        img = lib.CusmicImage(d.ctypes.data, e.ctypes.data, None, None, None, None, 512, 512)
        opts = lib.default_options()
        opts.contrast, opts.cr_threshold, opts.neighbor_threshold = 1.0, 5.0, 5.0
        t0 = time.perf_counter()
        host_call(img, opts, out, flags)
        first = 1e3 * (time.perf_counter() - t0)
        for _ in range(4):
            host_call(img, opts, out, flags)
        samples = []
        for _ in range(16):
            t0 = time.perf_counter()
            host_call(img, opts, out, flags)
            samples.append(1e3 * (time.perf_counter() - t0))
        ms = float(np.median(samples))
        exact = bool((out.view(np.uint64) == ref.astype(np.float64).view(np.uint64)).all()
                     and (flags.astype(bool) == crmask).all())
        result["host_call_512"] = {"value": round(0.262144 / (ms / 1e3), 2), "unit": "Mpix/s",
                                   "ms_per_frame": round(ms, 3), "first_call_ms": round(first, 3),
                                   "frame": "the reference's saved 512x512 scene, error map 5, settings "
                                            "(1, 5, 5), maxiter 4",
                                   "matches_saved_reference": exact,
                                   "note": "complete host calls (validation, allocation, copies); median of 16 "
                                           "after 4 warmups"}
        log(f"host call, saved 512x512 scene: {ms:.3f} ms, matches the saved reference: {exact}")
        return result

    # ----------------------------------------------------------- fwd_bwd
    def fwd_bwd(self):
        torch = self.torch
        from cusmic_amd import autograd

        log("forward+backward, 8 x 1024^2 noise-model frames")
        adj = self.scenes.adjoint_batch(self.dev)
        data, w = adj["data"], adj["weights"]
        mp = data.numel() / 1e6

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
        reps = max(4, self.args.steps)
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
        log(f"fwd {f_ms:.3f} ms, fwd+bwd {fb_ms:.3f} ms")
        # Outside the timed loops, the leg's two dominant kernels, each timed
        # with hipEvents on its stream over several launches:
        # * the forward's first pass under the noise model (significance_model):
        #   reads D and writes C, 16 B per frame pixel (gain and read noise are
        #   scalars, so no noise plane is read);
        # * the adjoint (cusmic_adjoint_device, one whole call): reads the
        #   upstream gradient and writes the data gradient (8 + 8 B), reads the
        #   mask (1 B) and the two donor indices of each flagged pixel (8 B).
        nf, h, wd = data.shape
        frames = self.engine.Frames()
        for f in range(nf):
            frames.add(data[f])
        first = []
        for _ in range(reps):
            prof = {}
            clean, mask, donors, _ = self.engine.run(frames, self.dev, 3.0, 5.0, 3.0, 4, self.lib.MIRROR,
                                                     gain=2.0, readnoise=5.0, record_donors=True, profile=prof)
            first.append(prof["first_pass_ms"])
        first_ms = float(self.np.median(first))
        grad = w.reshape(-1).contiguous()
        shapes = [(h, wd)] * nf
        for _ in range(2):
            self.engine.adjoint(shapes, self.dev, mask, donors, grad)
        # the device entry point returns after its stream synchronises, so each
        # call is timed alone; the events enclose its zeroing launch and its
        # host-side setup as well, which makes the rate a lower bound
        adj = []
        for _ in range(reps):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            self.engine.adjoint(shapes, self.dev, mask, donors, grad)
            ev[1].record()
            torch.cuda.synchronize()
            adj.append(ev[0].elapsed_time(ev[1]))
        adj_ms = float(self.np.median(adj))
        n_px = data.numel()
        flagged = int(mask.sum())
        adj_bytes = 17.0 * n_px + 8.0 * flagged
        log(f"noise-model first pass {first_ms:.4f} ms, adjoint {adj_ms:.4f} ms")
        return {"value": round(mp / (fb_ms / 1e3), 2), "unit": "Mpix/s",
                "forward_mpix_s": round(mp / (f_ms / 1e3), 2), "ms_forward": round(f_ms, 3),
                "ms_forward_backward": round(fb_ms, 3), "ratio": round(fb_ms / f_ms, 3),
                "frames": "8 x 1024x1024",
                "first_pass_roofline": roofline(16.0 * n_px, first_ms, kernel="significance_model"),
                "adjoint_roofline": roofline(adj_bytes, adj_ms, kernel="cusmic_adjoint_device call"),
                "flagged_pixels": flagged}

    # --------------------------------------------------------------- cpu
    def cpu(self):
        np = self.np
        import oracle

        result = {}
        cores, model = cpu_info()
        batch = self.batch
        offs = np.cumsum([0] + [b["data"].numel() for b in batch])
        pick = next((k for k, b in enumerate(batch) if b["data"].numel() == 1 << 20), 0)
        frames = [dict(data=batch[pick]["data"].cpu().numpy(), error=batch[pick]["error"].cpu().numpy())]
        spx = frames[0]["data"].size / 1e6
        log(f"cpu baseline: oracle, 1 thread, {spx:.2f} Mpx")
        t0 = time.perf_counter()
        o_clean, o_mask, _, _ = oracle.clean_frames(frames, threads=1)
        c_s = time.perf_counter() - t0
        # outside the timed regions: the ragged call's frame equals the oracle's
        lo, hi = int(offs[pick]), int(offs[pick + 1])
        g_clean = self.clean[lo:hi].cpu().numpy().reshape(o_clean[0].shape)
        g_mask = self.mask[lo:hi].cpu().numpy().reshape(o_mask[0].shape).astype(bool)
        equal = bool((g_clean.view(np.uint64) == o_clean[0].view(np.uint64)).all() and (g_mask == o_mask[0]).all())
        shape = "x".join(map(str, frames[0]["data"].shape))
        result["cpu_baseline"] = {"value": round(spx / c_s, 4), "unit": "Mpix/s", "cores": 1, "kind": "port",
                                  "cpu": model, "host_cores": cores,
                                  "sample": f"frame {pick} ({shape}) of the same ragged batch (seed 3), "
                                            f"error map, maxiter 4; {c_s:.1f} s",
                                  "gpu_output_equals_oracle": equal}
        log(f"cpu: {spx / c_s:.3f} Mpix/s; the oracle equals the ragged call's frame {pick}: {equal}")
        wide = [b for b in batch if b["data"].numel() == 1 << 20][:4] or batch[:1]
        wframes = [dict(data=b["data"].cpu().numpy(), error=b["error"].cpu().numpy()) for b in wide]
        wpx = sum(f["data"].size for f in wframes) / 1e6
        log(f"cpu baseline: oracle, {cores} threads, {wpx:.2f} Mpx")
        t0 = time.perf_counter()
        oracle.clean_frames(wframes, threads=cores)
        a_s = time.perf_counter() - t0
        result["cpu_baseline_all_cores"] = {
            "value": round(wpx / a_s, 4), "unit": "Mpix/s", "cores": cores, "kind": "port", "cpu": model,
            "sample": f"{len(wframes)} frame(s) of {'x'.join(map(str, wframes[0]['data'].shape))} from the same "
                      f"ragged batch, error map, maxiter 4, OpenMP over rows; {a_s:.1f} s"}
        log(f"cpu, {cores} threads: {wpx / a_s:.3f} Mpix/s")
        return result


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--no-per-frame", action="store_true")
    ap.add_argument("--no-detector", action="store_true")
    ap.add_argument("--no-backward", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--only", choices=("detector", "fwd-bwd"),
                    help="run this leg alone (for profiling) and print its record")
    ap.add_argument("--record", type=Path, metavar="DIR",
                    help="write benchmark records in the reference's schema (its saved scene, stacks of "
                         "1, 4 and 16 frames) into DIR, and exit")
    ap.add_argument("--ab", nargs=2, type=Path, metavar=("BASELINE", "CANDIDATE"),
                    help="compare two record directories with the reference's A/B rules, and exit")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="start with a 1 GiB device-to-device copy, whose hardware counter "
                         "readings scale FETCH_SIZE/WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.ab:
        from cusmic_amd import bench_record
        try:
            print("\n".join(bench_record.report(*args.ab)))
        except (OSError, ValueError, KeyError, TypeError, json.JSONDecodeError) as exc:
            ap.error(str(exc))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only; on several GPUs the frames of a batch would be "
              "sharded by pixel count, with no exchange between GPUs", file=sys.stderr)
        sys.exit(2)
    if args.record:
        from cusmic_amd import bench_record
        bench_record.record(args.record)
        return
    b = Bench(args)
    if args.calibration_copy:
        a = b.torch.ones(1 << 27, dtype=b.torch.float64, device=b.dev)
        b.torch.empty_like(a).copy_(a)
        b.torch.cuda.synchronize()
        del a
    if args.only:
        leg = b.detector() if args.only == "detector" else b.fwd_bwd()
        print(json.dumps({"leg": args.only, **leg}), flush=True)
        return
    result = b.ragged()
    if not args.no_per_frame:
        result["per_frame"] = b.per_frame()
    if not args.no_detector:
        result["detector"] = b.detector()
    if not args.no_host:
        result.update(b.host())
    if not args.no_backward:
        result["fwd_bwd"] = b.fwd_bwd()
    if not args.no_cpu:
        result.update(b.cpu())
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
