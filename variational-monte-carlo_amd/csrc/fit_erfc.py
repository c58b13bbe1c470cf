"""Coefficients of the FP32 erfc used by the local energy's real-space Ewald
term (vmc_device.hpp, erfc_fit): erfc(x) = exp(-x^2) P(t), t = (x - 1.5) /
(x + 1.5), for 0 <= x <= 2.5 (alpha r below the cutoff radius L/2 with
alpha = 5/L).  P is the degree-8 least-squares Chebyshev fit of the scaled
complementary error function erfcx(x) = exp(x^2) erfc(x) on 40001 points,
converted to the power basis in t.

The FP64 erfc of the same term (erfc_fit64) is a polynomial Q(t) of degree
24 in t = 0.8 x - 1 (the interval mapped onto [-1, 1]): the Chebyshev
interpolant of erfc at 25 nodes, computed in 50-digit arithmetic (mpmath),
converted to the power basis in t and rounded to double.  Its truncation
error is below 2e-16 and its FP64 evaluation, Q(t) = E(t^2) + t O(t^2) with
the even and odd coefficients in two Horner chains, stays within 4e-16 of
erfc in absolute terms (the Ewald sum needs absolute accuracy: each pair adds
erfc(alpha r) / r), with no exponential and no division.

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


DEG64 = 24


def coefficients64():
    """Power-basis coefficients in t = 0.8 x - 1 of the degree-24 Chebyshev
    interpolant of erfc on [0, 2.5], lowest first, as doubles."""
    import mpmath as mp

    n = DEG64 + 1
    with mp.workdps(50):
        nodes = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n)]
        vals = [mp.erfc(mp.mpf(XMAX) / 2 * (1 + t)) for t in nodes]
        cheb = [(2 if j else 1) * mp.fsum(vals[k] * mp.cos(mp.pi * j * (k + mp.mpf(1) / 2) / n)
                                          for k in range(n)) / n for j in range(n)]
        # T_j in the power basis by T_{j+1} = 2 t T_j - T_{j-1}
        T = [[mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]]
        for j in range(2, n):
            a = [mp.mpf(0)] + [2 * v for v in T[j - 1]]
            T.append([x - (T[j - 2][k] if k < len(T[j - 2]) else 0) for k, x in enumerate(a)])
        out = [mp.mpf(0)] * n
        for j, c in enumerate(cheb):
            for k, v in enumerate(T[j]):
                out[k] += c * v
        return [float(v) for v in out]


def fp64_error(coef, points=2001):
    """Largest absolute error against erfc of the FP64 evaluation as the
    kernel does it: t = 0.8 x - 1 and u = t^2 each rounded once, the even
    chain E(u) from c24 and the odd chain O(u) from c23 with every Horner step
    rounded once (as FMAs), and Q = O t + E rounded once."""
    import mpmath as mp

    fma = lambda a, b, c: float(mp.mpf(a) * b + c)  # noqa: E731
    worst = 0.0
    with mp.workdps(40):
        for x in np.linspace(0.0, XMAX, points):
            x = float(x)
            t = fma(x, 0.8, -1.0)
            u = float(mp.mpf(t) * t)
            e, o = coef[24], coef[23]
            for k in range(22, -1, -2):
                e = fma(e, u, coef[k])
            for k in range(21, 0, -2):
                o = fma(o, u, coef[k])
            q = fma(o, t, e)
            worst = max(worst, float(abs(q - mp.erfc(mp.mpf(x)))))
    return worst


if __name__ == "__main__":
    c = coefficients()
    for i, v in enumerate(c):
        print(f"c{i} = {np.float32(v)!r:>16}  ({v:.17g})")
    print("max relative error of P in FP32:", fp32_error(c))
    c64 = coefficients64()
    for i, v in enumerate(c64):
        print(f"d{i} = {v!r}")
    print("max absolute error of the FP64 erfc:", fp64_error(c64))
