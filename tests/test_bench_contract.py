"""bench.py's contract, on the CPU: one GPU only, the byte models behind the
roofline figures, and the committed counter summaries it reads."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert r.stdout == ""
    assert "exactly one GPU" in r.stderr


@pytest.mark.parametrize("n", [128, 512, 1024])
def test_byte_models(n):
    """The schedule's own bytes per step, (7 + 8h) s_c N^3, and the five-pass
    model the headline fraction uses, (11 + 8h) s_c N^3, h = (N/2+1)/N."""
    h = (n // 2 + 1) / n
    for sc, sr in ((16, 8), (8, 4)):
        pb = bench.pass_bytes(n, sc, sr)
        assert len(pb) == len(bench.PASS_NAMES)
        assert pb[0] == 0.0 and pb[-1] == 0.0           # drift table, reduction
        assert sum(pb) == pytest.approx((7 + 8 * h) * sc * n ** 3, rel=1e-12)
        assert bench.pencil_model_bytes(n, sc) == pytest.approx((11 + 8 * h) * sc * n ** 3, rel=1e-12)
    assert bench.pencil_model_bytes(512) == pytest.approx(32.25e9, rel=1e-3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_counter_traffic_covers_every_pass(precision):
    """Every pass of the 512^3 step has a counter-measured byte count, within
    a few percent of its algorithmic bytes (no pass re-reads)."""
    t = bench.counter_traffic(512, precision)
    sc, sr = (16, 8) if precision == "fp64" else (8, 4)
    alg = dict(zip(bench.PASS_NAMES, bench.pass_bytes(512, sc, sr)))
    passes = bench.PASS_NAMES[1:-1]
    assert set(t) == set(passes)
    for p in passes:
        assert 0.99 < t[p] / alg[p] < 1.05, (p, t[p] / alg[p])


def test_roofline_object():
    """roofline() prices the dominant pass's algorithmic bytes over its live
    duration in the timed steps against 8 TB/s, and reports the average of
    the pass-by-pass profile beside it."""
    avg = [0.006, 1.1, 0.9, 1.25, 0.4, 0.46, 0.4, 0.38, 0.01]
    r = bench.roofline(512, 3, 1.2, avg, 16, 8, "fp64")
    assert r["kernel"] == "C_zdrift_src_r2c" and r["bound"] == "hbm" and r["peak"] == 8000.0
    assert r["kernel_avg_ms"] == 1.2 and r["kernel_avg_ms_every_pass_profiled"] == 1.25
    alg = bench.pass_bytes(512)[3]
    assert r["achieved"] == pytest.approx(alg / 1.2e-3 / 1e9, rel=1e-3)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0, rel=1e-3)
    assert r["traffic"] == bench.counter_traffic(512, "fp64")["C_zdrift_src_r2c"]
    json.dumps(r)
