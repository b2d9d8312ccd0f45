"""Nonlinear response descriptors, mirroring ``src/Nonlinear.jl`` and ``src/Raman.jl``.

These objects carry the parameters each reference response closure captures;
the arithmetic runs in the HIP kernels:
  - ``Kerr_field(γ3)``            Nonlinear.jl:24-34  (P += ρ ε0 γ3 E³)
  - ``Kerr_env(γ3)``              Nonlinear.jl:63-73  (P += ¾ ρ ε0 γ3 |E|² E)
  - ``PlasmaCumtrapz(t, E, rate, ionpot; preionfrac)``  Nonlinear.jl:88-177
  - ``RamanPolarEnv(t, response)``                       Nonlinear.jl:262-333
  - ``RamanPolarField(t, response)``                     Nonlinear.jl:183-247, 284-333
Raman response functions: ``RamanRespNormedSingleDampedOscillator``
(``Raman.jl:53-65``) combined with the Planck window of
``CombinedRamanResponse`` (``Raman.jl:242-262``).
"""
import numpy as np

from .maths import planck_taper


class Kerr_field:
    def __init__(self, gamma3):
        self.gamma3 = float(gamma3)


class Kerr_env:
    def __init__(self, gamma3):
        self.gamma3 = float(gamma3)


class PlasmaCumtrapz:
    def __init__(self, t, ratefunc, ionpot, preionfrac=0.0):
        if not (0.0 <= preionfrac <= 1.0):
            raise ValueError("preionfrac must be between 0 and 1")  # DomainError, Nonlinear.jl:113
        self.dt = float(t[1] - t[0])
        self.ratefunc = ratefunc  # ionisation.PPTSpline
        self.ionpot = float(ionpot)
        self.preionfrac = float(preionfrac)


class RamanRespNormedSingleDampedOscillator:
    """``Raman.jl:53-65``: K ← K(Ω²+1/τ2²)/Ω; h(t) = K sin(Ωt) e^{-t/τ2} for t>0."""

    def __init__(self, K, Omega, tau2):
        self.K = K * (Omega ** 2 + 1 / tau2 ** 2) / Omega
        self.Omega = Omega
        self.tau2 = tau2

    def hrpre(self, t):
        return np.where(t > 0.0, self.K * np.sin(self.Omega * t), 0.0)

    def hrdamp(self, rho):
        return self.tau2


class CombinedRamanResponse:
    """``Raman.jl:242-262``."""

    def __init__(self, t, Rs):
        N = len(t)
        self.tt = np.arange(N, dtype=float) * (t[1] - t[0])
        self.Rs = list(Rs)
        self.hpres = [R.hrpre(self.tt) for R in self.Rs]
        te = self.tt[-1]
        self.w = planck_taper(self.tt, -te, -te * 0.7, te * 0.7, te)

    def __call__(self, rho):
        ht = np.zeros_like(self.tt)
        for hp, R in zip(self.hpres, self.Rs):
            ht = ht + hp * np.exp(-self.tt / R.hrdamp(rho))
        return ht * self.w


class RamanPolarEnv:
    def __init__(self, t, response):
        self.dt = float(t[1] - t[0])
        self.response = response

    def h(self, rho):
        return self.response(rho)


class RamanPolarField(RamanPolarEnv):
    """``RamanPolarField(t, r; thg=true)`` (``Nonlinear.jl:183-247``): the
    field-resolved Raman polarisation ρ·E·(h ⊛ E²), the convolution over the
    zero-padded 2·Nt grid.  Only the form with third-harmonic terms (E², the
    default) is provided."""

    def __init__(self, t, response, thg=True):
        if not thg:
            raise NotImplementedError("RamanPolarField(thg=false) (the Hilbert-transform form) is not provided")
        super().__init__(t, response)
