"""Coefficients of the FP32 erfc used by the local energy's real-space Ewald
term (vmc_device.hpp, erfc_fit): erfc(x) = exp(-x^2) P(t), t = (x - 1.5) /
(x + 1.5), for 0 <= x <= 2.5 (alpha r below the cutoff radius L/2 with
alpha = 5/L).  P is the degree-8 least-squares Chebyshev fit of the scaled
complementary error function erfcx(x) = exp(x^2) erfc(x) on 40001 points,
converted to the power basis in t.

The FP64 erfc of the same term (erfc_fit64) is exp(-x^2) Q(s),
s = (x - 3) / (x + 3), with Q the degree-14 polynomial interpolating erfcx at
the Chebyshev nodes of the s interval, solved in 40-digit arithmetic
(mpmath) and rounded to double: Q within 6e-16 relative of erfcx.

    python3 fit_erfc.py    # prints the coefficients and their errors
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


K64, DEG64 = 3, 14


def coefficients64():
    """Monomial coefficients in s of Q, lowest first, as doubles."""
    import mpmath as mp

    with mp.workdps(40):
        s0, s1 = mp.mpf(-1), (mp.mpf(XMAX) - K64) / (mp.mpf(XMAX) + K64)
        n = DEG64 + 1
        nodes = [(s1 - s0) / 2 * mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) + (s1 + s0) / 2 for k in range(n)]
        xs = [K64 * (1 + v) / (1 - v) for v in nodes]
        vals = [mp.erfc(x) * mp.exp(x * x) for x in xs]
        a = mp.matrix([[v ** j for j in range(n)] for v in nodes])
        c = mp.lu_solve(a, mp.matrix(vals))
        return [float(c[j]) for j in range(n)]


def fp64_error(coef, points=2001):
    """Largest relative error against erfc of exp(-x^2) Q(s) in double
    arithmetic (Horner with each step rounded once, as an FMA; x^2 split into
    its rounded value and the rounding), exp taken correctly rounded."""
    import mpmath as mp

    worst = 0.0
    with mp.workdps(40):
        for x in np.linspace(0.0, XMAX, points):
            x = float(x)
            s = (x - K64) / (x + K64)
            q = coef[-1]
            for c in coef[-2::-1]:
                q = float(mp.mpf(q) * s + c)
            h = x * x
            lo = float(mp.mpf(x) * x - h)
            e = float(mp.exp(-h))
            got = float(mp.mpf(e) * (1 - lo)) * q
            ref = mp.erfc(mp.mpf(x))
            worst = max(worst, float(abs((got - ref) / ref)))
    return worst


if __name__ == "__main__":
    c = coefficients()
    for i, v in enumerate(c):
        print(f"c{i} = {np.float32(v)!r:>16}  ({v:.17g})")
    print("max relative error of P in FP32:", fp32_error(c))
    c64 = coefficients64()
    for i, v in enumerate(c64):
        print(f"d{i} = {v!r}")
    print("max relative error of the FP64 erfc:", fp64_error(c64))
