"""Host helpers of bench.py (no GPU): the HBM traffic figure comes from the
committed rocprofv3 counter summary of the same workload, the roofline line
names the kernel with the largest total time, and p50/p99 are percentiles of
the step latencies."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402


def test_traffic_comes_from_the_committed_counter_summary():
    want = None
    for line in open(bench.PMC_SUMMARY):
        f = line.split()
        if len(f) == 5 and f[:4] == ["traffic_bytes_per_launch", "body_apply", "instances=2",
                                     "schedule=solo"]:
            want = float(f[4])
    assert want is not None and want > 0
    assert bench.traffic_from_profile("body_apply", 2, "solo") == want
    # another configuration has no committed summary: no figure rather than a wrong one
    assert bench.traffic_from_profile("body_apply", 3, "solo") is None
    assert bench.traffic_from_profile("body_apply", 2, "lockstep") is None
    assert bench.traffic_from_profile("no_such_kernel", 2, "solo") is None


def test_traced_average_comes_from_the_committed_kernel_trace():
    body = bench.traced_from_profile("body_apply",
                                     {"body_chunk": 512, "body_threads": 256, "prefetch": False})
    constr = bench.traced_from_profile("constraint_apply", {})
    assert body and constr and 0 < constr < body
    # a variant the trace did not run has no figure
    assert bench.traced_from_profile(
        "body_apply", {"body_chunk": 2048, "body_threads": 128, "prefetch": False}) is None


def test_schedule_follows_the_solo_threshold(monkeypatch):
    monkeypatch.delenv("PCC_SOLO_CELLS", raising=False)
    assert bench.schedule(300_000) == "solo"
    assert bench.schedule(100_000) == "lockstep"
    monkeypatch.setenv("PCC_SOLO_CELLS", "0")
    assert bench.schedule(300_000) == "lockstep"


def test_roofline_names_the_kernel_with_the_largest_share():
    tim = {"body_apply": {"launches": 100, "total_ms": 6.0, "bytes_per_launch": 3.0e8},
           "constraint_apply": {"launches": 100, "total_ms": 5.0, "bytes_per_launch": 2.7e8}}
    r = bench.roofline(tim)
    assert r["kernel"] == "body_apply"
    assert r["avg_us"] == pytest.approx(60.0)
    assert r["gbs"] == pytest.approx(3.0e8 / 60e-6 / 1e9)
    tim["body_apply"]["launches"] = 0
    assert bench.roofline(tim)["kernel"] == "constraint_apply"


def test_latency_percentiles_in_ms():
    p50, p99 = bench.latency_stats([0.001 * k for k in range(1, 101)])
    assert p50 == pytest.approx(50.5) and p99 == pytest.approx(99.01)
