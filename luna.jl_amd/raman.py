"""Raman response functions of molecular gases, restating ``src/Raman.jl``.

``RamanRespSingleDampedOscillator`` (``Raman.jl:39-70``), the vibrational
model ``RamanRespVibrational`` (``:124-158``), the non-rigid rotor
``RamanRespRotationalNonRigid`` (``:162-240``), ``CombinedRamanResponse``
with its Planck window (``:242-262``), ``molecular_raman_response`` and
``raman_response`` (``:274-326``) over the gas table ``raman_parameters``
(``PhysData.jl:968-1098``).  The response h(t) is evaluated here on the host
(once per density); the convolution with E² runs in the HIP kernels.
"""
import math

import numpy as np

from . import physdata as pd
from .maths import planck_taper
from .nonlinear import CombinedRamanResponse as _Combined
from .nonlinear import RamanRespNormedSingleDampedOscillator

EPS0 = 8.854187817620389e-12


class RamanRespSingleDampedOscillator:
    """h(t) = K sin(Ωt) e^{-t/τ2(ρ)} for t > 0 (``Raman.jl:39-70``)."""

    def __init__(self, K, Omega, tau2rho):
        self.K, self.Omega, self.tau2rho = float(K), float(Omega), tau2rho

    def hrpre(self, t):
        t = np.asarray(t, dtype=float)
        return np.where(t > 0.0, self.K * np.sin(self.Omega * t), 0.0)

    def hrdamp(self, rho):
        return self.tau2rho(rho)

    def __call__(self, t, rho):
        return self.hrpre(t) * np.exp(-np.asarray(t) / self.hrdamp(rho))


def _tau2rho(tau2, Brho, Arho, C=None, rotational=False):
    if Brho is None:
        if tau2 is None:
            raise ValueError("one of `τ2` or `Bρ` must be specified")
        return lambda rho: tau2
    if tau2 is not None:
        raise ValueError("only one of `τ2` or `Bρ` must be specified")
    if Arho is None:
        raise ValueError("if `Bρ` is specified you must also specify `Aρ`")
    amg = pd.amg
    if rotational:
        return lambda rho: 1 / (math.pi * (Arho / (rho / amg) + Brho * rho / amg)) if rho > 0 else math.inf
    C = 0.0 if C is None else C
    return lambda rho: 1 / (math.pi * (C + Arho / (rho / amg) + Brho * rho / amg))


def RamanRespVibrational(Omega_v, dadQ, mu, tau2=None, Brho=None, Arho=None, C=0.0):
    """``Raman.jl:142-158``: K = (4πε0)² (dα/dQ)² / (4 μ Ωv)."""
    K = (4 * math.pi * EPS0) ** 2 * dadQ ** 2 / (4 * mu * Omega_v)
    return RamanRespSingleDampedOscillator(K, Omega_v, _tau2rho(tau2, Brho, Arho, C))


class RamanRespRotationalNonRigid:
    """``Raman.jl:195-240``: a sum of damped oscillators over the J → J+2
    transitions of a non-rigid rotor, weighted by the thermal populations."""

    def __init__(self, B, dalpha, qJodd, qJeven, D=0.0, minJ=0, maxJ=50, temp=pd.roomtemp, tau2=None,
                 Brho=None, Arho=None):
        hbar, c, kB = pd.hbar, pd.c, pd.k_B
        J = np.arange(minJ, maxJ + 1, dtype=float)
        EJ = 2 * math.pi * hbar * c * (B * J * (J + 1) - D * (J * (J + 1)) ** 2)
        neg = np.flatnonzero(np.diff(EJ) < 0.0)
        if len(neg):
            mJ = neg[0] + 1  # Julia's findfirst, 1-based, of diff(EJ) < 0
            if mJ - minJ + 1 <= 2:
                raise ValueError("Raman rigid rotation model cannot sum over levels")
            J = J[:mJ]
            EJ = EJ[:mJ]
        DJ = np.where(J % 2 == 1, qJodd, qJeven)
        rho_J = DJ * (2 * J + 1) * np.exp(-EJ / (kB * temp))
        rho_J = rho_J / np.sum(rho_J)
        Js = J[:-2]
        Om = (EJ[2:] - EJ[:-2]) / hbar
        K = -(4 * math.pi * EPS0) ** 2 * 2 / 15 * dalpha ** 2 / hbar
        self.tau2rho = _tau2rho(tau2, Brho, Arho, rotational=True)
        self.Rs = [RamanRespSingleDampedOscillator(
            K * (Js[i] + 1) * (Js[i] + 2) / (2 * Js[i] + 3) * (rho_J[i + 2] / (2 * Js[i] + 5) - rho_J[i] / (2 * Js[i] + 1)),
            Om[i], self.tau2rho) for i in range(len(Js))]

    def hrpre(self, t):
        out = np.zeros_like(np.asarray(t, dtype=float))
        for R in self.Rs:
            out = out + R.hrpre(t)
        return out

    def hrdamp(self, rho):
        return self.tau2rho(rho)


class CombinedRamanResponse(_Combined):
    """``Raman.jl:242-262`` (the windowed sum over responses)."""


def raman_parameters(material):
    """``PhysData.raman_parameters`` (``PhysData.jl:968-1098``) for the
    molecular gases."""
    c, m_u = pd.c, pd.m_u
    if material == "N2":
        return dict(kind="molecular", rotation="nonrigid", vibration="sdo", B=199.0, D=5.74e-4, qJodd=1,
                    qJeven=2, dalpha=6.7e-31, Brho_r=3.3e9, Arho_r=0.0, dadQ=1.75e-20,
                    Omega_v=2 * math.pi * 2330.0 * 100.0 * c, mu=(m_u * 14.0067) ** 2 / (2 * m_u * 14.0067),
                    tau2_v=6e-12)
    if material == "H2":
        return dict(kind="molecular", rotation="nonrigid", vibration="sdo", B=5890.0, D=5.0, qJodd=3,
                    qJeven=1, dalpha=3e-31, Brho_r=114e6, Arho_r=6.15e6, dadQ=1.3e-20,
                    Omega_v=2 * math.pi * 124.5669e12, mu=(m_u * 1.00784) ** 2 / (2 * m_u * 1.00784),
                    Brho_v=52.2e6, Arho_v=309e6)
    if material == "D2":
        return dict(kind="molecular", rotation="nonrigid", vibration="sdo", B=2930.0, D=2.1, qJodd=1,
                    qJeven=2, dalpha=3e-31, Brho_r=4e-3 * 100.0 * c, Arho_r=0.0, dadQ=1.4e-20,
                    Omega_v=2 * math.pi * 2987 * 100.0 * c, mu=(m_u * 2.014) ** 2 / (2 * m_u * 2.014),
                    Brho_v=120e6, Arho_v=101e6)
    if material == "N2O":
        return dict(kind="molecular", rotation="nonrigid", vibration="none", B=41.0, D=0.0, qJodd=1, qJeven=1,
                    dalpha=28.1e-31, tau2_r=23.8e-12, Omega_v=2 * math.pi * 1285 * 100.0 * c)
    if material == "CH4":
        return dict(kind="molecular", rotation="none", vibration="sdo", dadQ=1.04e-20,
                    Omega_v=2 * math.pi * 2914 * 100.0 * c, mu=(1.00784 * m_u) / 4, Brho_v=384e6, Arho_v=0.0,
                    Cv=8220e6)
    if material == "SF6":
        return dict(kind="molecular", rotation="none", vibration="sdo", dadQ=1.23e-20,
                    Omega_v=2 * math.pi * 775 * 100.0 * c, mu=(18.998403 * m_u) / 6, tau2_v=6.6e-12)
    raise ValueError(f"Unknown material {material}")


def molecular_raman_response(t, rp, rotation=True, vibration=True, minJ=0, maxJ=50, temp=pd.roomtemp):
    """``Raman.jl:274-306``."""
    Rs = []
    if rotation and rp["rotation"] != "none":
        if rp["rotation"] != "nonrigid":
            raise ValueError(f"Unknown Rotational Raman model {rp['rotation']}")
        if "Brho_r" in rp:
            hr = RamanRespRotationalNonRigid(rp["B"], rp["dalpha"], rp["qJodd"], rp["qJeven"], Brho=rp["Brho_r"],
                                             Arho=rp["Arho_r"], D=rp["D"], minJ=minJ, maxJ=maxJ, temp=temp)
        else:
            hr = RamanRespRotationalNonRigid(rp["B"], rp["dalpha"], rp["qJodd"], rp["qJeven"], tau2=rp["tau2_r"],
                                             D=rp["D"], minJ=minJ, maxJ=maxJ, temp=temp)
        Rs.append(hr)
    if vibration and rp["vibration"] != "none":
        if rp["vibration"] != "sdo":
            raise ValueError(f"Unknown Vibrational Raman model {rp['vibration']}")
        if "Brho_v" in rp:
            hv = RamanRespVibrational(rp["Omega_v"], rp["dadQ"], rp["mu"], Brho=rp["Brho_v"], Arho=rp["Arho_v"],
                                      C=rp.get("Cv", 0.0))
        else:
            hv = RamanRespVibrational(rp["Omega_v"], rp["dadQ"], rp["mu"], tau2=rp["tau2_v"])
        Rs.append(hv)
    return CombinedRamanResponse(t, Rs)


def raman_response(t, material, rotation=True, vibration=True, minJ=0, maxJ=50, temp=pd.roomtemp):
    """``Raman.raman_response(t, material; kwargs...)`` for the molecular gases."""
    return molecular_raman_response(t, raman_parameters(material), rotation=rotation, vibration=vibration,
                                    minJ=minJ, maxJ=maxJ, temp=temp)


__all__ = ["RamanRespSingleDampedOscillator", "RamanRespVibrational", "RamanRespRotationalNonRigid",
           "CombinedRamanResponse", "RamanRespNormedSingleDampedOscillator", "raman_parameters",
           "molecular_raman_response", "raman_response", "planck_taper"]
