"""Waveguide modes for mode-averaged propagation (host-side set-up).

``MarcatiliMode`` restates ``src/Capillary.jl:81-296`` (HE/TE/TM hollow-capillary
modes, ``:full``/``:reduced`` models, loss switch, ``Aeff_Jintg``);
``SimpleMode`` restates ``src/SimpleFibre.jl:9-53`` (Taylor-series β, constant
loss and Aeff).  ``beta``/``alpha`` follow ``src/Modes.jl:136-148``;
``dispersion`` replaces the FiniteDifferences derivative of ``Maths.derivative``
(``src/Maths.jl:15-31``) with an 8th-order central stencil (β1 only fixes the
frame velocity of the time window).
"""
import math

import numpy as np
from scipy import integrate, special

from . import physdata as pd

_FD8 = np.array([1 / 280, -4 / 105, 1 / 5, -4 / 5, 0.0, 4 / 5, -1 / 5, 4 / 105, -1 / 280])


# 8th-order central second-derivative weights at offsets -4..4 (times 1/h²)
_FD8_2 = np.array([-1 / 560, 8 / 315, -1 / 5, 8 / 5, -205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560])


def derivative1(f, x, rel_h=1e-3):
    h = abs(x) * rel_h if x != 0 else rel_h
    pts = x + h * np.arange(-4, 5)
    vals = np.array([f(p) for p in pts], dtype=float)
    return float(np.dot(_FD8, vals) / h)


def beta2(mode, omega, rel_h=1e-3):
    """GVD β2 = d²β/dω² at ω (``Modes.dispersion(mode, 2, ω)``, ``src/Modes.jl:188-200``),
    by the 8th-order central stencil."""
    h = abs(omega) * rel_h
    pts = omega + h * np.arange(-4, 5)
    return float(np.dot(_FD8_2, mode.beta(pts)) / (h * h))


def zdw(mode, lam_min=100e-9, lam_max=3000e-9):
    """Zero-dispersion wavelength, ``Modes.zdw(m; λmin, λmax)`` (``src/Modes.jl:207-216``):
    the root of β2(ω) in (2πc/λmax, 2πc/λmin), NaN when β2 does not change sign
    there (the reference returns missing, which the statistics turn into NaN)."""
    from scipy.optimize import brentq
    lo, hi = 2 * math.pi * pd.c / lam_max, 2 * math.pi * pd.c / lam_min
    flo, fhi = beta2(mode, lo), beta2(mode, hi)
    if not (np.isfinite(flo) and np.isfinite(fhi)) or flo * fhi > 0:
        return float("nan")
    w0 = brentq(lambda w: beta2(mode, w), lo, hi, xtol=1e-12 * hi, rtol=1e-15, maxiter=200)
    return 2 * math.pi * pd.c / w0


class MarcatiliMode:
    """``Capillary.MarcatiliMode(a, gas, P; n=1, m=1, kind=:HE, T, model, clad, loss)``."""

    def __init__(self, a, gas, P, n=1, m=1, kind="HE", T=pd.roomtemp, model="full",
                 clad="SiO2", loss=True):
        self.a = float(a)
        self.n, self.m, self.kind = n, m, kind
        self.model, self.loss = model, loss
        self.gas, self.P, self.T = gas, P, T
        self._coren = pd.ref_index_fun(gas, P, T)
        self._cladn = pd.ref_index_fun(clad)
        if kind == "HE":
            self.unm = float(special.jn_zeros(n - 1, m)[-1])
        else:
            self.unm = float(special.jn_zeros(1, m)[-1])
        self.aeff_intg = self._aeff_jintg()

    def _aeff_jintg(self):
        """``Aeff_Jintg``, ``Capillary.jl:289-296``."""
        u, n = self.unm, self.n
        den, _ = integrate.quad(lambda r: r * special.jv(n - 1, u * r) ** 4, 0, 1,
                                epsabs=0, epsrel=1e-13, limit=200)
        np1 = n if self.kind == "HE" else 2
        num = 0.25 * special.jv(np1, u) ** 4
        return 2 * math.pi * num / den

    def Aeff(self, z=0.0):
        return self.a ** 2 * self.aeff_intg

    def aeff_onaxis(self):
        """P/I(r = 0) of the mode: the area that turns power into the on-axis
        intensity of Modes.to_space at (0, 0) with the y component
        (src/Modes.jl:458-493, Exy = field/√N, src/Modes.jl:73): for HE1m,
        E_y(0) = J0(0)/√N with N = π/2 a² J1(u)² √(ε0/μ0) (src/Capillary.jl:
        270-287), so I(0) = cε0/2·P/N and the area is 2N/(cε0) = π a² J1(u)².
        Other modes vanish on the axis."""
        if self.kind != "HE" or self.n != 1:
            raise NotImplementedError("on-axis statistics need an HE1m mode (the others vanish on the axis)")
        N = math.pi / 2 * self.a ** 2 * special.jv(1, self.unm) ** 2 * math.sqrt(pd.eps_0 / pd.mu_0)
        return 2 * N / (pd.c * pd.eps_0)

    def neff(self, omega):
        """``neff(m::MarcatiliMode, ω)``, ``Capillary.jl:136-186``."""
        omega = np.asarray(omega, dtype=float)
        lam = pd.wlfreq(omega)
        eps_cl = self._cladn(lam) ** 2
        eps_co = self._coren(lam) ** 2
        if self.kind == "HE":
            vn = (eps_cl + 1) / (2 * np.sqrt(eps_cl - 1 + 0j))
        elif self.kind == "TE":
            vn = 1 / np.sqrt(eps_cl - 1 + 0j)
        else:
            vn = eps_cl / np.sqrt(eps_cl - 1 + 0j)
        a = self.a
        if self.model == "full":
            k = omega / pd.c
            n = np.sqrt(eps_co - (self.unm / (k * a)) ** 2 * (1 - 1j * vn / (k * a)) ** 2 + 0j)
            if self.loss:
                n = np.where(n.real < 1e-3, 1e-3 + 1j * np.clip(n.imag, 0, np.inf), n)
                return n
            nr = n.real
            return np.where(nr < 1e-3, 1e-3, nr)
        # reduced model
        val = (1 + (eps_co - 1) / 2 - pd.c ** 2 * self.unm ** 2 / (2 * omega ** 2 * a ** 2))
        if self.loss:
            return val + 1j * (pd.c ** 3 * self.unm ** 2) / (a ** 3 * omega ** 3) * vn
        return np.real(val)

    def vn(self, omega):
        lam = pd.wlfreq(np.asarray(omega, dtype=float))
        eps_cl = self._cladn(lam) ** 2
        if self.kind == "HE":
            return (eps_cl + 1) / (2 * np.sqrt(eps_cl - 1 + 0j))
        if self.kind == "TE":
            return 1 / np.sqrt(eps_cl - 1 + 0j)
        return eps_cl / np.sqrt(eps_cl - 1 + 0j)

    def neff_wg(self, omega):
        """Waveguide part of the effective index, ``neff_wg`` (``Capillary.jl:188-214``):
        neff = sqrt(εco - nwg) (:full) or 1 + (εco - 1)/2 - nwg (:reduced)."""
        omega = np.asarray(omega, dtype=float)
        vn, a = self.vn(omega), self.a
        if self.model == "full":
            k = omega / pd.c
            return (self.unm / (k * a)) ** 2 * (1 - 1j * vn / (k * a)) ** 2
        val = pd.c ** 2 * self.unm ** 2 / (2 * omega ** 2 * a ** 2) + 0j
        if self.loss:
            val = val + 1j * (pd.c ** 3 * self.unm ** 2) / (a ** 3 * omega ** 3) * vn
        return val

    def beta(self, omega):
        return np.asarray(omega) / pd.c * np.real(self.neff(omega))

    def alpha(self, omega):
        return 2 * np.asarray(omega) / pd.c * np.imag(self.neff(omega))

    def dispersion1(self, omega0):
        return derivative1(lambda w: float(self.beta(np.array([w]))[0]), omega0)

    def losslength(self, omega):
        return 1 / self.alpha(omega)

    def dB_per_m(self, omega):
        return 10 / math.log(10) * self.alpha(omega)


class SimpleMode:
    """``SimpleFibre.SimpleMode(ωref, βs; Aeff=1.0, loss=0.0)``, ``SimpleFibre.jl:23-53``."""

    def __init__(self, omega_ref, betas, Aeff=1.0, loss=0.0):
        self.omega_ref = float(omega_ref)
        self.coeffs = np.array([b / math.factorial(i) for i, b in enumerate(betas)], dtype=float)
        self._aeff = Aeff
        self._alpha = math.log(10) / 10 * loss

    def _poly(self, coeffs, x):
        # Horner evaluation (Base.evalpoly)
        x = np.asarray(x, dtype=float)
        ex = np.full_like(x, coeffs[-1])
        for c in coeffs[-2::-1]:
            ex = x * ex + c
        return ex

    def beta(self, omega):
        return self._poly(self.coeffs, np.asarray(omega) - self.omega_ref)

    def alpha(self, omega):
        return np.full_like(np.asarray(omega, dtype=float), self._alpha)

    def dispersion1(self, omega0):
        d = np.array([self.coeffs[i] * i for i in range(1, len(self.coeffs))])
        if len(d) == 0:
            return 0.0
        return float(self._poly(d, np.array([omega0 - self.omega_ref]))[0])

    def Aeff(self, z=0.0):
        return self._aeff
