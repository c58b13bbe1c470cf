"""The FP32 and FP64 erfc of the local energy's real-space Ewald term
(erfc_fit, erfc_fit64 in vmc_device.hpp) against erfc: the polynomial
coefficients in the source are the fits that fit_erfc.py makes; the FP32
evaluation of exp(-x^2) P(t) stays within 3e-7 relative of erfc on
0 <= x <= 2.5 (the exponential is taken in float64 here; the device's
v_exp_f32 adds about one rounding), and the FP64 evaluation of the degree-24
polynomial Q(0.8 x - 1) within 4e-16 of erfc in absolute terms (40-digit
erfc as the reference)."""
import importlib.util
import pathlib
import re

import numpy as np
from scipy.special import erfc

CSRC = pathlib.Path(__file__).resolve().parents[1] / "variational-monte-carlo_amd" / "csrc"


def source_coefficients():
    text = (CSRC / "vmc_device.hpp").read_text()
    body = text[text.index("__device__ inline float erfc_fit(float x)"):]
    body = body[: body.index("constexpr float kLog2e")]
    vals = [float(v) for v in re.findall(r"([-+]?\d\.\d+e[-+]\d+)f", body)]
    return vals[::-1]  # c0 .. c8 (Horner starts from the highest)


def test_coefficients_are_the_fit():
    spec = importlib.util.spec_from_file_location("fit_erfc", CSRC / "fit_erfc.py")
    fit = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fit)
    want = np.asarray(fit.coefficients(), dtype=np.float32)
    got = np.asarray(source_coefficients(), dtype=np.float32)
    assert got.shape == (9,)
    np.testing.assert_array_equal(got, want)


def test_fp32_evaluation_tracks_erfc():
    c = np.asarray(source_coefficients(), dtype=np.float32)
    x = np.linspace(0.0, 2.5, 100001).astype(np.float32)
    t = ((x - np.float32(1.5)) * (np.float32(1) / (x + np.float32(1.5)))).astype(np.float32)
    p = np.full_like(x, c[-1])
    for cc in c[-2::-1]:
        p = (p * t + cc).astype(np.float32)
    x64 = x.astype(np.float64)
    got = np.exp(-x64 * x64) * p.astype(np.float64)
    rel = np.abs(got - erfc(x64)) / erfc(x64)
    assert rel.max() < 3e-7
    # A dropped or mistyped coefficient fails by orders of magnitude.
    bad = c.copy()
    bad[3] *= 1.01
    q = np.full_like(x, bad[-1])
    for cc in bad[-2::-1]:
        q = (q * t + cc).astype(np.float32)
    assert (np.abs(np.exp(-x64 * x64) * q - erfc(x64)) / erfc(x64)).max() > 1e-5


def fit_module():
    spec = importlib.util.spec_from_file_location("fit_erfc", CSRC / "fit_erfc.py")
    fit = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fit)
    return fit


def source_coefficients64():
    text = (CSRC / "vmc_device.hpp").read_text()
    body = text[text.index("__device__ inline double erfc_fit64(double x)"):]
    body = body[body.index("constexpr double c[25] = {"): body.index("};")]
    body = body[body.index("{") + 1:]
    return [float(v) for v in body.replace("\n", " ").split(",")]


def test_fp64_coefficients_are_the_fit():
    fit = fit_module()
    want = fit.coefficients64()
    got = source_coefficients64()
    assert len(got) == 25
    assert got == want


def test_fp64_evaluation_tracks_erfc():
    fit = fit_module()
    c = source_coefficients64()
    assert fit.fp64_error(c, points=401) < 4e-16
    # A coefficient off in its eighth digit already fails.
    bad = list(c)
    bad[9] *= 1 + 1e-7
    assert fit.fp64_error(bad, points=401) > 1e-13
