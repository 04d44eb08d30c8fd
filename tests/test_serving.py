"""Serving helpers on the CPU (no GPU): memory estimate, TI estimate, job stream."""

import numpy as np
import pytest

from timemachine_amd import serving


def test_estimate_system_bytes_grows_and_covers_tiles():
    a, b = serving.estimate_system_bytes(2000), serving.estimate_system_bytes(36000)
    assert 0 < a < b
    nb = -(-36000 // 64)
    assert b >= (nb * (nb + 1) // 2 + nb) * 64 * 4


def test_ti_estimate_trapezoid():
    lam = np.array([0.0, 0.25, 1.0])
    du_dl = np.array([[1.0, 3.0], [4.0, 4.0], [0.0, 2.0]])  # means 2, 4, 1
    assert serving.ti_estimate(lam, du_dl) == pytest.approx(0.25 * 3 + 0.75 * 2.5)


def test_synthetic_stream_shapes():
    jobs = serving.synthetic_stream(3, rate_per_s=2.0, seed=1, atoms=(1800, 2600), windows=(2, 4), steps=(20, 60),
                                    heavy=(8, 10))
    arr = [j.arrival for j in jobs]
    assert arr == sorted(arr) and arr[0] > 0
    for j in jobs:
        assert 1700 <= j.system.N <= 2700
        assert j.system.box[0, 0] >= 2.62
        assert 2 <= j.windows <= 4 and j.n_steps % 10 == 0


def test_latency_stats():
    rs = [serving.JobResult(i, np.zeros((1, 1)), 0.0, 0.0, 0.0, float(i + 1)) for i in range(100)]
    st = serving.latency_stats(rs)
    assert st["p50_s"] == pytest.approx(50.5) and st["p99_s"] == pytest.approx(99.01)
