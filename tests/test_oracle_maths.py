"""cumtrapz and CSpline of the oracle against the reference's own tests and exact cases.

test/test_maths.jl:77-97 (cumtrapz cos→sin, rtol 1e-6), :167-207 (CSpline).
"""
import math

import numpy as np


def test_cumtrapz_cos_to_sin(orc):
    x = np.linspace(0, 8 * np.pi, 2 ** 14) * 1e-15
    y = np.cos(x * 1e15)
    yi = np.sin(x * 1e15) / 1e15
    got = orc.cumtrapz(y, x[1] - x[0])
    np.testing.assert_allclose(got, yi, rtol=1e-6, atol=1e-6 * np.max(np.abs(yi)))


def test_cumtrapz_exact_for_linear(orc):
    dx = 0.25
    y = 3.0 + 2.0 * np.arange(1000) * dx  # trapezoid rule is exact for linear data
    x = np.arange(1000) * dx
    exact = 3.0 * x + x ** 2
    np.testing.assert_allclose(orc.cumtrapz(y, dx), exact, rtol=1e-13, atol=1e-12)


def test_cumtrapz_matches_long_double(orc):
    rng = np.random.default_rng(3)
    y = rng.standard_normal(16384)
    dx = 0.37
    ref = np.zeros(len(y), dtype=np.longdouble)
    acc = np.longdouble(0)
    for i in range(1, len(y)):
        acc += (np.longdouble(0.5) * (np.longdouble(y[i - 1]) + np.longdouble(y[i]))) * np.longdouble(dx)
        ref[i] = acc
    got = orc.cumtrapz(y, dx)
    assert np.max(np.abs(got - ref.astype(float))) < 1e-11


def test_cspline_sin(orc, pkg):
    from luna_jl_amd.maths import cspline_derivatives, cspline_index
    x = np.linspace(0.0, 2 * np.pi, 100)
    y = np.sin(x)
    D = cspline_derivatives(y)
    # at the nodes (src test: < 5e-18)
    assert np.max(np.abs(orc.cspline(x, y, D, x) - y)) < 5e-18
    x2 = np.linspace(0.0, 2 * np.pi, 300)
    assert np.max(orc.cspline(x, y, D, x2) - np.sin(x2)) < 5e-8
    # uniform index rule == brute force "first node greater than x0"
    for x0 in x2:
        slow = 2 if x0 <= x[0] else (len(x) if x0 >= x[-1] else int(np.argmax(x > x0)) + 1)
        assert cspline_index(x0, x[0], x[-1], len(x)) == slow


def test_cspline_cubic_exact(orc, pkg):
    # a Hermite cubic with exact derivatives reproduces a cubic exactly
    x = np.linspace(-1, 2, 31)
    f = lambda u: 0.5 * u ** 3 - u ** 2 + 2 * u - 1
    df = lambda u: 1.5 * u ** 2 - 2 * u + 2
    h = x[1] - x[0]
    x0 = np.linspace(-0.99, 1.99, 97)
    got = orc.cspline(x, f(x), df(x) * h, x0)
    np.testing.assert_allclose(got, f(x0), rtol=1e-12, atol=1e-12)
