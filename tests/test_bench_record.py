"""The record schema and A/B rules of bench_record (bench/bench.cu:297-332,
bench/ab.py:38-117), on synthetic records; no GPU."""

import copy
import json

import pytest

from cusmic_amd import bench_record


def fake(rev, clean_ms, total_ms, frames=1, **over):
    rec = {"backend": "hip", "dtype": "float64", "shape": ([frames] if frames > 1 else []) + [512, 512],
           "warmups": 4, "repeats": len(clean_ms), "first_result_ms": 5.0, "host_memory": "pageable",
           "allocation": "torch caching allocator", "detected_pixels": 6518 * frames,
           "settings": dict(bench_record.SETTINGS), "source_revision": rev, "source_dirty": False,
           "cusmic": "x", "gpu": "MI355X", "cuda_runtime": 1, "cuda_driver": 2, "reference_mode": "exact",
           "reference_close": True, "reference_exact": True, "max_abs_error": 0.0, "mask_disagreements": 0,
           "milliseconds": {k: bench_record._stats(v) for k, v in
                            (("upload_ms", clean_ms), ("clean_ms", clean_ms), ("download_ms", clean_ms),
                             ("total_ms", total_ms))},
           "frames_per_second": 1000 * frames / bench_record._median(total_ms)}
    rec.update(over)
    return rec


A, B = "a" * 40, "b" * 40


def write(tmp, name, recs):
    d = tmp / name
    d.mkdir()
    for r in recs:
        shape = r["shape"]
        (d / f"hip-{shape[0] if len(shape) == 3 else 1}.json").write_text(json.dumps(r))
    return d


def test_statistics_match_the_reference():
    s = bench_record._stats([3.0, 1.0, 2.0, 4.0])
    assert s == {"median": 2.5, "minimum": 1.0, "maximum": 4.0, "samples": [3.0, 1.0, 2.0, 4.0]}


def test_report_speedups(tmp_path):
    base = write(tmp_path, "base", [fake(A, [2.0, 2.0], [4.0, 4.0]), fake(A, [6.0, 6.0], [12.0, 12.0], 4)])
    cand = write(tmp_path, "cand", [fake(B, [1.0, 1.0], [2.0, 2.0]), fake(B, [4.0, 4.0], [8.0, 8.0], 4)])
    lines = bench_record.report(base, cand, frames=(1, 4))
    rows = [line.split() for line in lines if line.strip().startswith(("1 ", "4 "))]
    assert [r[2] for r in rows] == ["resident", "ordinary", "resident", "ordinary"]
    assert [r[-1] for r in rows] == ["2.00x", "2.00x", "1.50x", "1.50x"]
    assert rows[2][3] == "1.500" and rows[2][4] == "1.000"  # per-frame medians


@pytest.mark.parametrize("change, message", [
    (dict(gpu="other"), "GPU, software"),
    (dict(source_revision=A), "distinct Git"),
    (dict(source_revision="not-a-sha"), "distinct Git"),
    (dict(detected_pixels=1), "detected-pixel"),
    (dict(settings={"contrast": 3.0}), "settings differ"),
    (dict(source_dirty=True), "unmodified"),
    (dict(reference_exact=False), "reference pixels"),
])
def test_rules(change, message):
    base = fake(A, [1.0], [2.0])
    cand = fake(B, [1.0], [2.0])
    bench_record.compare(base, cand)
    bad = copy.deepcopy(cand)
    bad.update(change)
    with pytest.raises(ValueError, match=message):
        bench_record.compare(base, bad)


def test_samples_must_match_repeats():
    rec = fake(A, [1.0, 2.0], [2.0, 3.0])
    rec["repeats"] = 3
    with pytest.raises(ValueError, match="repeats"):
        bench_record.samples(rec, "clean_ms")
