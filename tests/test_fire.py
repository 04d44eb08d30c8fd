"""FIRE minimisation on the CPU: the oracle (oracle/fire.py) against closed
forms and a quasi-Newton minimum, and the arithmetic header shared with the
HIP kernel (timemachine_amd/csrc/fire.h, host build) against the oracle.
"""

import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest
from scipy.optimize import minimize

from oracle import fire

ROOT = Path(__file__).resolve().parents[1]
P = ctypes.POINTER


@pytest.fixture(scope="module")
def hc(tmp_path_factory):
    out = tmp_path_factory.mktemp("hcf") / "host_check.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC",
                    f"-I{ROOT / 'timemachine_amd' / 'csrc'}", str(ROOT / "tests" / "native" / "host_check.cpp"),
                    "-o", str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    lib.hc_half_round.restype = ctypes.c_float
    lib.hc_half_round.argtypes = [ctypes.c_float]
    lib.hc_fire_harmonic.argtypes = [ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double), ctypes.c_double,
                                     ctypes.c_int, P(ctypes.c_double), P(ctypes.c_float), P(ctypes.c_float),
                                     P(ctypes.c_int)]
    return lib


def test_half_round_matches_numpy_float16(hc):
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.normal(0, 1, 4000), rng.uniform(-7e4, 7e4, 2000), rng.uniform(-1e-4, 1e-4, 3000),
        rng.uniform(-1e-7, 1e-7, 1000),
        # exact ties between neighbouring float16 values, both parities
        (np.arange(1, 2000, dtype=np.float64) + 0.5) * 2.0**-24, (np.arange(1024, 2048) + 0.5) * 2.0**-10,
        [0.0, -0.0, 65504.0, 65519.0, 65520.0, 65536.0, 6.1e-5, 6.103515625e-05, 5.96e-8, 2.98e-8, 1e-30],
    ]).astype(np.float32)
    got = np.array([hc.hc_half_round(float(v)) for v in vals], dtype=np.float32)
    with np.errstate(over="ignore"):
        exp = vals.astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(got, exp)


def test_constants_are_float16(hc):
    from oracle.fire import ALPHA_START, DT_MAX, DT_START, F_ALPHA, F_INC

    for v, ref in ((DT_START, 1e-5), (DT_MAX, 1e-3), (F_INC, 1.1), (ALPHA_START, 0.1), (F_ALPHA, 0.99)):
        assert np.min_scalar_type(ref) == np.float16  # how _vendored/fire.py:103-105 casts them
        assert float(v) == hc.hc_half_round(ref)


def _harmonic(k, c):
    return lambda x, lam: -k * (x - c)


def test_first_two_steps_closed_form():
    """V starts at 0 and float16(dt0^2) = 0: the first step does not move; the
    second moves by dt0 * V1 with V1 = dt0 F0 (parallel to F, so the mixing keeps it)."""
    rng = np.random.default_rng(1)
    x0 = rng.normal(size=(4, 3))
    c = rng.normal(size=(4, 3))
    k = 250.0
    F0 = -k * (x0 - c)
    dt0 = float(np.float16(1e-5))
    x1 = fire.fire_minimize(x0, _harmonic(k, c), np.zeros(2))
    np.testing.assert_array_equal(x1, x0)
    x2 = fire.fire_minimize(x0, _harmonic(k, c), np.zeros(3))
    np.testing.assert_allclose(x2 - x0, dt0 * dt0 * F0, rtol=0, atol=1e-15)  # x0 ~ 1: rounding of x


def test_step_size_schedule_under_steady_descent():
    """Constant force: P > 0 every step, so dt stays at dt0 for n_min steps and then
    grows by f_inc per step (float16 rounding) up to dt_max; alpha shrinks by f_alpha."""
    tr = []
    g = np.array([[3.0, -1.0, 2.0]])
    fire.fire_minimize(np.zeros((1, 3)), lambda x, lam: g, np.zeros(80), trace=tr)
    dt0 = float(np.float16(1e-5))
    for k, t in enumerate(tr, start=1):
        assert t["P"] > 0 and t["n_pos"] == k
        grow = max(0, k - 5)
        ideal = min(dt0 * 1.1**grow, 1e-3)
        assert float(t["dt"]) == pytest.approx(ideal, rel=(grow + 1) * 2.0**-10 + 1e-3 * (ideal == 1e-3))
        assert float(t["alpha"]) == pytest.approx(0.1 * 0.99**grow, rel=(grow + 1) * 2.0**-10)
    assert float(tr[-1]["dt"]) == float(np.float16(1e-3))


def test_uphill_step_resets():
    """Overshooting a stiff minimum makes P < 0: dt halves, alpha resets, n_pos = 0."""
    tr = []
    fire.fire_minimize(np.array([[0.3, 0.0, 0.0]]), _harmonic(4e6, np.zeros((1, 3))), np.zeros(120), trace=tr)
    hits = [i for i, t in enumerate(tr) if t["P"] < 0]
    assert hits
    for i in hits:
        assert tr[i]["n_pos"] == 0
        assert tr[i]["alpha"] == np.float16(0.1)
        assert tr[i]["dt"] == np.float16(tr[i - 1]["dt"] * np.float16(0.5))


def test_converges_to_quasi_newton_minimum():
    """A random spring network: the FIRE minimum energy equals BFGS's from the same start."""
    rng = np.random.default_rng(3)
    n = 8
    x0 = rng.uniform(0, 1.0, size=(n, 3))
    pairs = np.array([(i, j) for i in range(n) for j in range(i + 1, n) if rng.uniform() < 0.6])
    r0 = rng.uniform(0.2, 0.5, len(pairs))
    kb = 2e4

    def energy(x):
        d = x[pairs[:, 0]] - x[pairs[:, 1]]
        r = np.linalg.norm(d, axis=1)
        return 0.5 * kb * np.sum((r - r0) ** 2)

    def force(x, lam=0.0):
        d = x[pairs[:, 0]] - x[pairs[:, 1]]
        r = np.linalg.norm(d, axis=1)
        g = (kb * (r - r0) / r)[:, None] * d
        f = np.zeros_like(x)
        np.add.at(f, pairs[:, 0], -g)
        np.add.at(f, pairs[:, 1], g)
        return f

    xf = fire.fire_minimize(x0, force, np.zeros(4000))
    res = minimize(lambda v: energy(v.reshape(n, 3)), x0.reshape(-1), jac=lambda v: -force(v.reshape(n, 3)).reshape(-1),
                   method="BFGS", options={"gtol": 1e-10})
    assert energy(xf) < 1e-3 * energy(x0)
    assert energy(xf) == pytest.approx(res.fun, abs=1e-6 * energy(x0))
    assert np.abs(force(xf)).max() < 1e-2 * np.abs(force(x0)).max()


def test_header_matches_oracle(hc):
    """fire.h (the kernel's arithmetic, host build, sequential sums) against the
    oracle on the same harmonic problem: identical dt / alpha / n_pos sequences,
    positions to rounding."""
    rng = np.random.default_rng(5)
    n3 = 6
    x0 = rng.normal(size=(2, 3))
    c = rng.normal(size=(2, 3))
    k = 3e5
    n_sched = 300
    tr = []
    xo = fire.fire_minimize(x0, _harmonic(k, c), np.zeros(n_sched), trace=tr)
    xh = np.zeros(n3)
    dt = np.zeros(n_sched - 1, np.float32)
    al = np.zeros(n_sched - 1, np.float32)
    npos = np.zeros(n_sched - 1, np.int32)
    pd = lambda a: a.ctypes.data_as(P(ctypes.c_double))  # noqa: E731
    hc.hc_fire_harmonic(n3, pd(np.ascontiguousarray(x0)), pd(np.ascontiguousarray(c)), k, n_sched, pd(xh),
                        dt.ctypes.data_as(P(ctypes.c_float)), al.ctypes.data_as(P(ctypes.c_float)),
                        npos.ctypes.data_as(P(ctypes.c_int)))
    assert any(t["P"] < 0 for t in tr)  # the run exercises the reset branch
    np.testing.assert_array_equal(dt, np.array([t["dt"] for t in tr], dtype=np.float32))
    np.testing.assert_array_equal(al, np.array([t["alpha"] for t in tr], dtype=np.float32))
    np.testing.assert_array_equal(npos, [t["n_pos"] for t in tr])
    np.testing.assert_allclose(xh.reshape(2, 3), xo, rtol=0, atol=1e-12)
