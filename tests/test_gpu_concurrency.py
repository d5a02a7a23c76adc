"""Calls from several host threads at once, each on its own stream (GPU).

Each call takes its own pinned block for the frame table and the counters
(blocks come from a shared pool and return to it), its own device scratch,
and posts its counters from its own stream. Calls that overlap must give
exactly what the same calls give one after the other: masks and pixels bit
for bit, and the same pass counts. maxiter 9 makes each call read its
counters between chunks of passes as well as at the end.
"""

import threading

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*args):
    print("[case]", *args, flush=True)


def batch(seed, shapes):
    from cusmic_amd import engine, scenes

    fr = engine.Frames()
    keep = []
    for k, shape in enumerate(shapes):
        data, signal = scenes.frame(shape, seed + k, "cpu", n_stars=6, n_tracks=20, max_len=6)
        d = data.to("cuda")
        e = scenes.error_map(signal, 2.0, 5.0).to("cuda")
        keep.append((d, e))
        fr.add(d, error=e)
    return fr, keep


def clean(fr, maxiter):
    from cusmic_amd import _lib, engine

    out, mask, _, passes = engine.run(fr, torch.device("cuda", 0), 3.0, 5.0, 3.0, maxiter, _lib.MIRROR)
    return out.clone(), mask.clone(), passes


@pytest.mark.parametrize("maxiter", [4, 9])
def test_threads_on_own_streams_equal_sequential(cuda, maxiter):
    jobs = [batch(100 + 10 * j, shapes) for j, shapes in
            enumerate([((256, 300), (64, 500)), ((300, 200),), ((128, 128), (77, 301), (20, 20))])]
    say("sequential reference, maxiter", maxiter)
    want = [clean(fr, maxiter) for fr, _ in jobs]
    torch.cuda.synchronize()
    rounds = 6
    got = [[None] * rounds for _ in jobs]
    errors = []

    def worker(j):
        try:
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                for r in range(rounds):
                    got[j][r] = clean(jobs[j][0], maxiter)
            stream.synchronize()
        except Exception as e:  # reported by the main thread
            errors.append((j, e))

    say("threads", len(jobs), "x", rounds, "calls")
    threads = [threading.Thread(target=worker, args=(j,)) for j in range(len(jobs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads), "a worker did not finish"
    assert not errors, errors
    for j, (w_out, w_mask, w_passes) in enumerate(want):
        assert int(w_mask.sum()) > 0
        for r in range(rounds):
            out, mask, passes = got[j][r]
            np.testing.assert_array_equal(out.view(torch.int64).cpu().numpy(),
                                          w_out.view(torch.int64).cpu().numpy(),
                                          err_msg=f"pixels of job {j}, round {r}")
            np.testing.assert_array_equal(mask.cpu().numpy(), w_mask.cpu().numpy(),
                                          err_msg=f"mask of job {j}, round {r}")
            assert passes == w_passes
