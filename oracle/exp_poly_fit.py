"""Where the exp polynomial's coefficients come from (orc_expf in
lumen_oracle.c and splat_exp / splat_exp_neg in lumen_amd/csrc/splat_math.hpp).

p(r) = 1 + c1 r + ... + c6 r^6 approximates e^r on |r| <= ln2/2, the range
left by the Cody-Waite reduction.  c1..c6 minimise the maximum relative error
(Lawson's iteratively reweighted least squares on 4000 Chebyshev nodes, which
converges to the minimax fit), then are rounded to float32; the constant term
stays exactly 1 so that exp(0) = 1.  Prints the coefficients as the C sources
spell them and the fit's error before rounding.

    python oracle/exp_poly_fit.py

The float32 evaluation (Horner with fmaf, as both sources run it) is checked
against exp in float64 by tests/test_oracle_kats.py::test_exp_accuracy.
"""
import numpy as np


def fit(deg: int = 6, nodes: int = 4000, iters: int = 200):
    a = np.log(2.0) / 2.0
    r = np.cos(np.pi * (np.arange(nodes) + 0.5) / nodes) * a
    f = np.exp(r)
    A = np.stack([r ** k for k in range(1, deg + 1)], axis=1)
    w = np.ones(nodes)
    for _ in range(iters):
        W = np.sqrt(w) / f
        c, *_ = np.linalg.lstsq(A * W[:, None], (f - 1.0) * W, rcond=None)
        err = (1.0 + A @ c - f) / f
        w = w * np.abs(err)
        w /= w.sum()
    return c, float(np.max(np.abs(err)))


if __name__ == "__main__":
    c, e = fit()
    print("max relative error before rounding: %.3e" % e)
    for k, x in enumerate(c.astype(np.float32), start=1):
        print("c%d = %.9ef" % (k, x))
