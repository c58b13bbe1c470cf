"""Coefficients of the FP32 erfc used by the local energy's real-space Ewald
term (vmc_device.hpp, erfc_fit): erfc(x) = exp(-x^2) P(t), t = (x - 1.5) /
(x + 1.5), for 0 <= x <= 2.5 (alpha r below the cutoff radius L/2 with
alpha = 5/L).  P is the degree-8 least-squares Chebyshev fit of the scaled
complementary error function erfcx(x) = exp(x^2) erfc(x) on 40001 points,
converted to the power basis in t.

    python3 fit_erfc.py    # prints the coefficients and the FP32 error
"""
import numpy as np
from scipy.special import erfcx

K, XMAX, DEG = 1.5, 2.5, 8


def coefficients():
    x = np.linspace(0.0, XMAX, 40001)
    t = (x - K) / (x + K)
    cheb = np.polynomial.chebyshev.Chebyshev.fit(t, erfcx(x), DEG, domain=[t.min(), t.max()])
    poly = cheb.convert(kind=np.polynomial.Polynomial, domain=[t.min(), t.max()],
                        window=[t.min(), t.max()])
    return poly.coef


def fp32_error(coef):
    """Largest relative error of the FP32 Horner evaluation of P(t) against
    erfcx, t formed in FP32 (the exponential factor is not included)."""
    x = np.linspace(0.0, XMAX, 200001).astype(np.float32)
    t = ((x - np.float32(K)) * (np.float32(1) / (x + np.float32(K)))).astype(np.float32)
    c = coef.astype(np.float32)
    acc = np.full_like(x, c[-1])
    for cc in c[-2::-1]:
        acc = (acc * t + cc).astype(np.float32)
    ref = erfcx(x.astype(np.float64))
    return float(np.max(np.abs(acc.astype(np.float64) - ref) / ref))


if __name__ == "__main__":
    c = coefficients()
    for i, v in enumerate(c):
        print(f"c{i} = {np.float32(v)!r:>16}  ({v:.17g})")
    print("max relative error of P in FP32:", fp32_error(c))
