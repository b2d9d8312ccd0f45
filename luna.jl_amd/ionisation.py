"""Ionisation rates and the PPT lookup table (host-side set-up).

Restates ``src/Ionisation.jl``: the ADK rate (``:22-63``), the PPT rate
(``ionrate_fun_PPT`` ``:317-390`` with ``φ`` ``:403-429``), the table builder
``makePPTcache`` (``:185-200``) with ``barrier_suppression`` (``:208-212``), and
``makePPTaccel`` (``:247-271``), which keeps the rate>0 points and builds a
cubic spline of log(rate).  The spline is evaluated on the GPU; this module
returns the node/value/derivative arrays the kernels read.

Special functions come from scipy (``gamma``, ``dawsn``, ``hyp1f1``) and mpmath
for the large-argument branch of φ (the reference uses BigFloat quadrature
there).  Table values are pinned to the reference's golden values
(``test/test_ionisation.jl:6-16``) in the tests.
"""
import math
import numpy as np
from scipy import special

from . import physdata as pd
from .maths import converge_series, cspline_derivatives, julia_range


def ionrate_fun_adk(ionpot, threshold=True, cycle_average=False):
    """Vectorised ADK rate, ``Ionisation.jl:22-63``."""
    nstar = math.sqrt(0.5 / (ionpot / pd.au_energy))
    cn_sq = 2 ** (2 * nstar) / (nstar * math.gamma(nstar + 1) * math.gamma(nstar))
    w_p = ionpot / pd.hbar
    w_t_prefac = pd.electron / math.sqrt(2 * pd.m_e * ionpot)
    Ip_au = ionpot / pd.au_energy
    F0 = (2 * Ip_au) ** 1.5 * pd.au_Efield
    avfac = math.sqrt(3 / (math.pi * F0))

    def raw(E):
        aE = np.abs(np.asarray(E, dtype=float))
        with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
            r = (w_p * cn_sq * (4 * w_p / (w_t_prefac * aE)) ** (2 * nstar - 1)
                 * np.exp(-4 / 3 * w_p / (w_t_prefac * aE)))
        if cycle_average:
            r = r * avfac * np.sqrt(aE)
        return np.where(np.isfinite(r), r, 0.0)

    thr = 0.0
    if threshold:
        E = 1e3
        while raw(np.array([E]))[0] == 0:
            E *= 1.01
        thr = E

    def rate(E):
        E = np.asarray(E, dtype=float)
        return np.where(np.abs(E) >= thr, raw(E), 0.0)

    return rate


def ionrate_adk(gas, E, **kw):
    return ionrate_fun_adk(pd.ionisation_potential(gas), **kw)(E)


def _phi(m, x):
    """PPT φ_m(x), ``Ionisation.jl:403-429``."""
    if m == 0:
        return special.dawsn(x)
    mabs = abs(m)
    if x <= 26:
        return (math.exp(-x * x) * math.sqrt(math.pi) * x ** (2 * mabs + 1) * math.gamma(mabs + 1)
                * special.hyp1f1(0.5, 1.5 + mabs, x * x) / (2 * math.gamma(1.5 + mabs)))
    import mpmath
    mpmath.mp.dps = 40
    xm = mpmath.mpf(x)
    val = mpmath.quad(lambda y: (xm ** 2 - y ** 2) ** mabs * mpmath.exp(y ** 2), [0, xm])
    return float(mpmath.exp(-xm ** 2) * val)


def ionrate_fun_ppt(ionpot, lam0, Z, l, sum_tol=1e-6, cycle_average=False,
                    sum_integral=False, dalpha=0.0, alpha_ion=0.0, msum=True, Cnl=None,
                    occupancy=2):
    """Scalar PPT rate closure, ``Ionisation.jl:317-390``."""
    dalpha = dalpha or 0.0
    alpha_ion_au = (alpha_ion or 0.0) / pd.au_polarisability

    def rate(E):
        Ip_au = (ionpot + dalpha / 2 * E * E) / pd.au_energy
        ns = Z / math.sqrt(2 * Ip_au)
        ls = ns - 1
        Cnl2 = (2 ** (2 * ns) / (ns * math.gamma(ns + ls + 1) * math.gamma(ns - ls))
                if Cnl is None else Cnl ** 2)
        w0 = 2 * math.pi * pd.c / lam0
        w0_au = pd.au_time * w0
        E0_au = (2 * Ip_au) ** 1.5
        E_au = abs(E) / pd.au_Efield
        g = w0_au * math.sqrt(2 * Ip_au) / E_au
        g2 = g * g
        beta = 2 * g / math.sqrt(1 + g2)
        alpha = 2 * (math.asinh(g) - g / math.sqrt(1 + g2))
        Up_au = E_au ** 2 / (4 * w0_au ** 2)
        Uit_au = Ip_au + Up_au
        v = Uit_au / w0_au
        ret = 0.0
        mrange = range(-l, l + 1) if msum else (0,)
        for m in mrange:
            mabs = abs(m)
            flm = ((2 * l + 1) * math.factorial(l + mabs)
                   / (2 ** mabs * math.factorial(mabs) * math.factorial(l - mabs)))
            lret = 4 * math.sqrt(2) / math.pi * Cnl2
            lret *= (2 * E0_au / (E_au * math.sqrt(1 + g2))) ** (2 * ns - mabs - 1.5)
            lret *= flm / math.factorial(mabs)
            lret *= math.exp(-2 * v * (math.asinh(g) - g * math.sqrt(1 + g2) / (1 + 2 * g2)))
            lret *= Ip_au * g2 / (1 + g2)
            if not cycle_average:
                lret *= math.sqrt(math.pi * E0_au / (3 * E_au))
            n0 = math.ceil(v)
            if sum_integral:
                s = (math.sqrt(math.pi) * math.factorial(mabs) * beta ** mabs
                     / (2 * (alpha + beta) ** (mabs + 1)) * math.sqrt(beta / alpha))
            else:
                def term(x, n):
                    diff = n - v
                    return x + math.exp(-alpha * diff) * _phi(m, math.sqrt(beta * diff))
                s, _, _ = converge_series(term, 0.0, n0=n0, rtol=sum_tol, maxiter=math.inf)
            lret *= s
            ret += occupancy * lret
        if alpha_ion_au != 0:
            ret *= math.exp(-2 * alpha_ion_au * E_au)
        return ret / pd.au_time

    return rate


def ppt_options(gas, stark_shift=True, dipole_corr=True):
    _, l, Z = pd.quantum_numbers(gas)
    da = pd.polarisability_difference(gas) if stark_shift else 0.0
    ai = pd.polarisability(gas, True) if dipole_corr else 0.0
    return dict(ionpot=pd.ionisation_potential(gas), Z=Z, l=l, dalpha=da or 0.0, alpha_ion=ai or 0.0)


def ionrate_ppt(gas, lam0, E, stark_shift=True, dipole_corr=True, **kw):
    o = ppt_options(gas, stark_shift, dipole_corr)
    f = ionrate_fun_ppt(o["ionpot"], lam0, o["Z"], o["l"], dalpha=o["dalpha"],
                        alpha_ion=o["alpha_ion"], **kw)
    E = np.atleast_1d(np.asarray(E, dtype=float))
    return np.array([f(e) for e in E])


def barrier_suppression(ionpot, Z):
    """Barrier-suppression field, ``Ionisation.jl:208-212``."""
    Ip_au = ionpot / pd.au_energy
    ns = Z / math.sqrt(2 * Ip_au)
    return Z ** 3 / (16 * ns ** 4) * pd.au_Efield


def _phi_vec(m, x):
    if m == 0:
        return special.dawsn(x)
    mabs = abs(m)
    out = np.empty_like(x)
    small = x <= 26
    xs = x[small]
    out[small] = (np.exp(-xs * xs) * math.sqrt(math.pi) * xs ** (2 * mabs + 1) * math.gamma(mabs + 1)
                  * special.hyp1f1(0.5, 1.5 + mabs, xs * xs) / (2 * math.gamma(1.5 + mabs)))
    for i in np.nonzero(~small)[0]:
        out[i] = _phi(m, float(x[i]))
    return out


def ionrate_ppt_vec(ionpot, lam0, Z, l, E, sum_tol=1e-6, dalpha=0.0, alpha_ion=0.0,
                    occupancy=2, Cnl=None):
    """Vectorised ``ionrate_fun_PPT`` (same arithmetic per element, ``Ionisation.jl:317-390``).

    The ``converge_series`` loop runs in lock-step over all field values, each
    element stopping by the reference's own relative-change criterion.
    """
    E = np.asarray(E, dtype=float)
    alpha_ion_au = (alpha_ion or 0.0) / pd.au_polarisability
    Ip_au = (ionpot + (dalpha or 0.0) / 2 * E * E) / pd.au_energy
    ns = Z / np.sqrt(2 * Ip_au)
    ls = ns - 1
    if Cnl is None:
        Cnl2 = 2 ** (2 * ns) / (ns * special.gamma(ns + ls + 1) * special.gamma(ns - ls))
    else:
        Cnl2 = np.full_like(E, Cnl ** 2)
    w0 = 2 * math.pi * pd.c / lam0
    w0_au = pd.au_time * w0
    E0_au = (2 * Ip_au) ** 1.5
    E_au = np.abs(E) / pd.au_Efield
    g = w0_au * np.sqrt(2 * Ip_au) / E_au
    g2 = g * g
    beta = 2 * g / np.sqrt(1 + g2)
    alpha = 2 * (np.arcsinh(g) - g / np.sqrt(1 + g2))
    Up_au = E_au ** 2 / (4 * w0_au ** 2)
    v = (Ip_au + Up_au) / w0_au
    ret = np.zeros_like(E)
    for m in range(-l, l + 1):
        mabs = abs(m)
        flm = ((2 * l + 1) * math.factorial(l + mabs)
               / (2 ** mabs * math.factorial(mabs) * math.factorial(l - mabs)))
        lret = 4 * math.sqrt(2) / math.pi * Cnl2
        lret = lret * (2 * E0_au / (E_au * np.sqrt(1 + g2))) ** (2 * ns - mabs - 1.5)
        lret = lret * (flm / math.factorial(mabs))
        lret = lret * np.exp(-2 * v * (np.arcsinh(g) - g * np.sqrt(1 + g2) / (1 + 2 * g2)))
        lret = lret * (Ip_au * g2 / (1 + g2))
        lret = lret * np.sqrt(math.pi * E0_au / (3 * E_au))
        n = np.ceil(v)
        x0 = np.zeros_like(E)
        s = np.zeros_like(E)
        active = np.arange(len(E))
        while active.size:
            diff = n[active] - v[active]
            x1 = x0[active] + np.exp(-alpha[active] * diff) * _phi_vec(m, np.sqrt(beta[active] * diff))
            with np.errstate(divide="ignore", invalid="ignore"):
                done = 2 * np.abs(x1 - x0[active]) / np.abs(x1 + x0[active]) < sum_tol
            s[active] = x1
            x0[active] = x1
            n[active] += 1
            active = active[~done]
        ret = ret + occupancy * (lret * s)
    if alpha_ion_au != 0:
        ret = ret * np.exp(-2 * alpha_ion_au * E_au)
    return ret / pd.au_time


def ppt_device_table(ionpot, lam0, Z, l, E, sum_tol=1e-6, dalpha=0.0, alpha_ion=0.0, occupancy=2,
                     Cnl=None, timing=None):
    """The same rates computed on the GPU (``lh_ppt_table``: one thread per
    field value, each running the converge_series loop itself)."""
    import ctypes as C
    from . import _capi
    E = np.ascontiguousarray(E, dtype=float)
    d = _capi.LhPptDesc(ionpot=float(ionpot), lambda0=float(lam0), Z=float(Z), l=int(l), msum=1,
                        dalpha=float(dalpha or 0.0), alpha_ion_au=float(alpha_ion or 0.0) / pd.au_polarisability,
                        sum_tol=float(sum_tol), occupancy=float(occupancy), cycle_average=0, reserved=0,
                        cnl=float(Cnl) if Cnl is not None else 0.0, c=pd.c, au_energy=pd.au_energy,
                        au_time=pd.au_time, au_efield=pd.au_Efield)
    rate = np.empty_like(E)
    ms = C.c_double(0.0)
    _capi.check(_capi.lib().lh_ppt_table(C.byref(d), len(E), _capi._ptr(E), _capi._ptr(rate), C.byref(ms)),
                "lh_ppt_table")
    if timing is not None:
        timing["kernel_ms"] = ms.value
    return rate


def make_ppt_cache(gas, lam0, N=2 ** 16, Emax=None, stark_shift=True, dipole_corr=True, device=False, **kw):
    """(E, rate) table, ``makePPTcache`` ``Ionisation.jl:185-200``; ``device``:
    computed by the GPU kernel instead of the host loop."""
    o = ppt_options(gas, stark_shift, dipole_corr)
    if Emax is None:
        Emax = 2 * barrier_suppression(o["ionpot"], o["Z"])
    Emin = Emax / 5000
    E = julia_range(Emin, Emax, N)
    f = ppt_device_table if device else ionrate_ppt_vec
    rate = f(o["ionpot"], lam0, o["Z"], o["l"], E, dalpha=o["dalpha"], alpha_ion=o["alpha_ion"], **kw)
    return E, rate


class PPTSpline:
    """Log-rate cubic spline of ``makePPTaccel`` (``Ionisation.jl:247-271``).

    Holds the node positions ``x`` (E values with rate > 0), ``y = log(rate)`` and
    the CSpline derivatives ``d``.  Evaluation: |E| < x[0] → 0; |E| > x[-1] →
    error; otherwise exp(spline(|E|)).
    """

    def __init__(self, E, rate):
        keep = rate > 0
        self.x = np.ascontiguousarray(E[keep], dtype=float)
        self.y = np.ascontiguousarray(np.log(rate[keep]), dtype=float)
        self.d = np.ascontiguousarray(cspline_derivatives(self.y), dtype=float)
        dx = self.x[1] - self.x[0]
        assert np.allclose(np.diff(self.x), dx, rtol=1e-6, atol=0)

    @property
    def emin(self):
        return float(self.x[0])

    @property
    def emax(self):
        return float(self.x[-1])


_PPT_CACHE = {}


def ppt_spline_cached(gas, lam0, N=2 ** 16, cachedir=None, device=False, **kw):
    """``ionrate_fun!_PPTcached`` (``Ionisation.jl:142-175``): table cached on disk (.npz)."""
    import hashlib
    import os
    key = (gas, float(lam0), int(N), tuple(sorted(kw.items())))
    if key in _PPT_CACHE:
        return _PPT_CACHE[key]
    if cachedir is None:
        cachedir = os.environ.get("LUNA_HIP_CACHE", os.path.join(os.path.expanduser("~"), ".luna_hip", "pptcache"))
    h = hashlib.sha1(repr(key).encode()).hexdigest()[:16]
    path = os.path.join(cachedir, f"ppt_{h}.npz")
    if os.path.isfile(path):
        dat = np.load(path)
        E, rate = dat["E"], dat["rate"]
    else:
        E, rate = make_ppt_cache(gas, lam0, N=N, device=device, **kw)
        try:
            os.makedirs(cachedir, exist_ok=True)
            tmp = path + f".{os.getpid()}.tmp.npz"
            np.savez(tmp, E=E, rate=rate)
            os.replace(tmp, path)
        except OSError:
            pass
    spl = PPTSpline(E, rate)
    _PPT_CACHE[key] = spl
    return spl
