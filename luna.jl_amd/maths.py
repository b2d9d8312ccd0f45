"""Host-side numerical helpers used at set-up time.

Restates ``planck_taper`` (``Maths.jl:404-452``), the ``CSpline`` construction
(``Maths.jl:812-833``) and its uniform index rule (``Maths.jl:764-789``), and
``converge_series`` (``Maths.jl:721-738``).  Spline *evaluation* is on the hot
path and lives in the HIP kernels; the host only builds the node/derivative
tables they read.
"""
import math
import numpy as np
from scipy.linalg import solve_banded


def _taper(xc, x1, x2, x3, x4):
    out = np.zeros_like(xc, dtype=float)
    m = (x1 < xc) & (xc < x2)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        z12 = (x2 - x1) / (xc[m] - x1) + (x2 - x1) / (xc[m] - x2)
        out[m] = 1.0 / (1.0 + np.exp(z12))
        out[(x2 <= xc) & (xc <= x3)] = 1.0
        m = (x3 < xc) & (xc < x4)
        z34 = (x3 - x4) / (xc[m] - x3) + (x3 - x4) / (xc[m] - x4)
        out[m] = 1.0 / (1.0 + np.exp(z34))
    return out


def planck_taper(x, left0, left1, right1, right0):
    """4-point Planck taper window, ``Maths.jl:429-452``."""
    x = np.asarray(x, dtype=float)
    x0 = (right0 + left0) / 2
    xc = x - x0
    X = right0 - left0
    eleft = abs(left1 - left0) / X
    eright = abs(right0 - right1) / X
    x1 = -X / 2
    x2 = -X / 2 * (1 - 2 * eleft)
    x3 = X / 2 * (1 - 2 * eright)
    x4 = X / 2
    return _taper(xc, x1, x2, x3, x4)


def cspline_derivatives(y):
    """Solve the CSpline tridiagonal system (``Maths.jl:814-832``).

    R = 3·(central differences, one-sided at the ends); M = tridiag(1, 4, 1) with
    2 at both ends of the diagonal; D = M \\ R.
    """
    y = np.asarray(y, dtype=float)
    n = len(y)
    R = np.empty(n)
    R[0] = y[1] - y[0]
    R[1:-1] = y[2:] - y[:-2]
    R[-1] = y[-1] - y[-2]
    R *= 3
    ab = np.zeros((3, n))
    ab[0, 1:] = 1.0
    ab[1, :] = 4.0
    ab[1, 0] = 2.0
    ab[1, -1] = 2.0
    ab[2, :-1] = 1.0
    return solve_banded((1, 1), ab, R)


def cspline_index(x0, xmin, xmax, n):
    """Uniform-grid 1-based index rule of ``make_spline_ifun`` (``Maths.jl:771-775``)."""
    if x0 <= xmin:
        return 2
    if x0 >= xmax:
        return n
    return int(math.ceil((x0 - xmin) / (xmax - xmin) * (n - 1))) + 1


def cspline_eval(x, y, D, x0):
    """Scalar CSpline evaluation, ``Maths.jl:840-848`` (host reference for tables)."""
    n = len(x)
    i = cspline_index(x0, x[0], x[-1], n) - 1  # 0-based index of the upper node
    if x0 == x[i]:
        return y[i]
    if x0 == x[i - 1]:
        return y[i - 1]
    t = (x0 - x[i - 1]) / (x[i] - x[i - 1])
    return (y[i - 1] + D[i - 1] * t + (3 * (y[i] - y[i - 1]) - 2 * D[i - 1] - D[i]) * (t * t)
            + (2 * (y[i - 1] - y[i]) + D[i - 1] + D[i]) * (t * t * t))


def cspline_eval_vec(x, y, D, x0):
    """``cspline_eval`` at an array of points (the same index rule and cubic)."""
    x = np.asarray(x, dtype=float)
    x0 = np.asarray(x0, dtype=float)
    n = len(x)
    xmin, xmax = x[0], x[-1]
    with np.errstate(invalid="ignore"):
        r = np.ceil((x0 - xmin) / (xmax - xmin) * (n - 1))
    r = np.where(np.isfinite(r), r, 0.0)
    i1 = np.where(x0 <= xmin, 2, np.where(x0 >= xmax, n, r.astype(np.int64) + 1))
    i = np.clip(i1, 2, n) - 1  # 0-based index of the upper node
    xl, xu, yl, yu, Dl, Du = x[i - 1], x[i], y[i - 1], y[i], D[i - 1], D[i]
    t = (x0 - xl) / (xu - xl)
    v = yl + Dl * t + (3 * (yu - yl) - 2 * Dl - Du) * (t * t) + (2 * (yl - yu) + Dl + Du) * (t * t * t)
    return np.where(x0 == xu, yu, np.where(x0 == xl, yl, v))


def converge_series(f, x0, n0=0, rtol=1e-6, maxiter=10000):
    """Brute-force series limit, ``Maths.jl:721-738``."""
    n = n0
    x1 = x0
    success = False
    while (not success) and n < maxiter:
        x1 = f(x0, n)
        if 2 * abs(x1 - x0) / abs(x1 + x0) < rtol:
            success = True
        n += 1
        x0 = x1
    return x1, success, n


def julia_range(start, stop, length):
    """Julia ``range(start, stop, length)`` values (TwicePrecision lerp, correctly rounded).

    Evaluated in extended precision as start + i·(stop-start)/(length-1).
    """
    i = np.arange(length, dtype=np.longdouble)
    s = np.longdouble(start)
    e = np.longdouble(stop)
    vals = (s * (length - 1 - i) + e * i) / np.longdouble(length - 1)
    return vals.astype(np.float64)


def simpson_even(dx, y):
    """NumericalIntegration.jl ``SimpsonEven`` on a uniform grid (extended Simpson rule),
    along the first axis."""
    y = np.asarray(y)
    ret = (17 * y[0] + 59 * y[1] + 43 * y[2] + 49 * y[3]
           + 49 * y[-4] + 43 * y[-3] + 59 * y[-2] + 17 * y[-1]) / 48
    ret = ret + np.sum(y[4:-4], axis=0)
    return dx * ret


def hilbert(x, axis=0):
    """Analytic signal via FFT along ``axis``, ``Maths.jl:560-569``."""
    x = np.moveaxis(np.asarray(x, dtype=float), axis, 0)
    xf = np.fft.fft(x, axis=0)
    n = xf.shape[0]
    n1 = n // 2
    xf[1:n1] *= 2
    xf[n1:] = 0
    return np.moveaxis(np.fft.ifft(xf, axis=0), 0, axis)


def fwhm(x, y, minmax="min"):
    """``Maths.fwhm(x, y; method=:linear, minmax)`` (``src/Maths.jl:128-214``)
    at level ½·max(y): the crossings nearest the (first) maximum for "min",
    the outermost for "max", each interpolated linearly between the samples
    either side; NaN where a crossing or its neighbour is missing (the
    reference's try/catch)."""
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    if minmax not in ("min", "max"):
        raise ValueError("minmax has to be min or max")
    i0 = int(np.argmax(y))
    val = 0.5 * y[i0]
    xm = x[i0]
    n = len(x)

    def xing(i1, i2):  # linterpx: x where the line through (x1, y1), (x2, y2) reaches val
        slope = (y[i2] - y[i1]) / (x[i2] - x[i1])
        return (val - (y[i1] - slope * x[i1])) / slope

    if minmax == "min":
        lo = np.flatnonzero((x < xm) & (y < val))
        hi = np.flatnonzero((x > xm) & (y < val))
        if len(lo) == 0 or len(hi) == 0:
            return math.nan
        li, ri = lo[-1], hi[0]
        return abs(xing(ri - 1, ri) - xing(li, li + 1))
    lo = np.flatnonzero((x < xm) & (y > val))
    hi = np.flatnonzero((x > xm) & (y > val))
    if len(lo) == 0 or len(hi) == 0 or lo[0] < 1 or hi[-1] + 1 >= n:
        return math.nan
    li, ri = lo[0], hi[-1]
    return abs(xing(ri, ri + 1) - xing(li - 1, li))
