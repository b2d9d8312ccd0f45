"""Physical constants and gas/glass material data (host-side setup only).

Restates the parts of the reference's ``src/PhysData.jl`` that the propagation
hot path consumes at set-up time: CODATA-2014 constants (``PhysData.jl:14-48``),
gas Sellmeier expansions (``PhysData.jl:95-245``), the SiO2 Sellmeier formula
(``PhysData.jl:253-260``), refractive-index closures (``PhysData.jl:448-518``),
the Lehmeier γ3 model (``PhysData.jl:588-657``), ionisation potentials, quantum
numbers and polarisabilities (``PhysData.jl:731-859``).

Number density: the reference calls CoolProp (``PhysData.jl:695-699``), a C++
library that is not available here.  ``density`` below is an ideal gas with a
second/third virial correction calibrated to the REFPROP values held in the
reference's ``test/test_physdata.jl:62-75``.  The propagation kernels take the
density as an explicit number, so parity between the GPU path and the oracle
does not depend on this model; its absolute value is pinned only to ~1e-4.
"""
import math
import numpy as np

# CODATA 2014 (PhysData.jl:14-48; PhysicalConstants.CODATA2014)
c = 299792458.0
atm = 101325.0
bar = 100000.0
k_B = 1.38064852e-23
mu_0 = 4e-7 * math.pi
eps_0 = 1.0 / (mu_0 * c * c)
electron = 1.6021766208e-19
m_e = 9.10938356e-31
e_ratio = electron ** 2 / m_e
hbar = 1.054571800e-34
alpha_fs = 7.2973525664e-3
a_0 = 0.52917721067e-10
au_energy = hbar * c * alpha_fs / a_0
au_time = hbar / au_energy
au_Efield = au_energy / (electron * a_0)
roomtemp = 293.15
N_A = 6.022140857e23
amg = atm / (k_B * 273.15)
m_u = 1.660539040e-27
au_polarisability = electron ** 2 * a_0 ** 2 / au_energy

GASES = ("He", "HeJ", "HeB", "Ne", "Ar", "ArB", "Kr", "Xe", "N2")


def wlfreq(x):
    """ω ↔ λ conversion, ``PhysData.jl:78``."""
    return 2 * math.pi * c / x


# --- density -----------------------------------------------------------------
# Second/third virial coefficients (m^3/mol, m^6/mol^2) at 294 K, fitted to the
# REFPROP values in test_physdata.jl:62-75, with a linear temperature slope for B
# taken from the usual literature tables (He: B nearly flat, Ar: dB/dT > 0).
_VIRIAL = {
    "He": (1.1848656766756638e-05, -0.0025e-6, 1.181192181582509e-10, 294.0),
    "Ar": (-1.6337103599085788e-05, 0.2476e-6, 1.0424498572262705e-09, 294.0),
    "Ne": (11.3e-6, 0.03e-6, 2.2e-10, 294.0),
    "Kr": (-52.0e-6, 0.55e-6, 2.7e-9, 294.0),
    "Xe": (-133.0e-6, 1.3e-6, 7.0e-9, 294.0),
    "N2": (-5.0e-6, 0.2e-6, 1.4e-9, 294.0),
}


def _base(gas):
    if gas in ("He", "HeJ", "HeB"):
        return "He"
    if gas in ("Ar", "ArB"):
        return "Ar"
    return gas


def density(gas, P=1.0, T=roomtemp):
    """Number density [m^-3] at pressure P [bar] and temperature T [K].

    Replaces ``PhysData.density`` (``PhysData.jl:695-699``, CoolProp).  Solves
    P = ρm R T (1 + B ρm + C ρm²) for the molar density ρm by Newton iteration.
    """
    if P == 0:
        return 0.0
    B0, dBdT, C, T0 = _VIRIAL[_base(gas)]
    B = B0 + dBdT * (T - T0)
    R = k_B * N_A
    p = bar * P
    rho = p / (R * T)
    for _ in range(50):
        f = rho * R * T * (1 + B * rho + C * rho * rho) - p
        df = R * T * (1 + 2 * B * rho + 3 * C * rho * rho)
        step = f / df
        rho -= step
        if abs(step) < 1e-16 * rho:
            break
    return rho * N_A


def pressure(gas, density, T=roomtemp):
    """Pressure [bar] of a number density [m^-3]: the inverse of ``density``
    (replaces ``PhysData.pressure``, ``PhysData.jl:708-711``, CoolProp), i.e.
    P = ρm R T (1 + B ρm + C ρm²) evaluated directly."""
    if density == 0:
        return 0.0
    B0, dBdT, C, T0 = _VIRIAL[_base(gas)]
    B = B0 + dBdT * (T - T0)
    R = k_B * N_A
    rho = density / N_A
    return rho * R * T * (1 + B * rho + C * rho * rho) / bar


def density_spline(gas, Pmax, Pmin=0.0, N=2 ** 10, T=roomtemp):
    """Nodes (P, ρ, D) of ``densityspline`` (``PhysData.jl:719-723``): a CSpline
    of the density over N uniform pressures range(Pmin, Pmax, N)."""
    from .maths import cspline_derivatives, julia_range
    P = julia_range(Pmin, Pmax, N)
    rho = np.array([density(gas, float(p), T) for p in P])
    return P, rho, cspline_derivatives(rho)


def dens_1bar_0degC(gas):
    return density(gas, 1.0, 273.15)


def dens_1atm_0degC(gas):
    return density(gas, atm / bar, 273.15)


# --- Sellmeier ---------------------------------------------------------------
def sellmeier_gas(gas):
    """Linear polarisability γ(λ[μm]) per particle, ``PhysData.jl:146-220``."""
    dens = dens_1bar_0degC(gas)
    if gas == "HeB":
        B1, C1, B2, C2 = 4977.77e-8, 28.54e-6, 1856.94e-8, 7.76e-3
        return _borzsonyi(B1 / dens, C1, B2 / dens, C2)
    if gas in ("He", "HeJ"):
        B1, C1 = 2.16463842e-05, -6.80769781e-04
        B2, C2 = 2.10561127e-07, 5.13251289e-03
        B3, C3 = 4.75092720e-05, 3.18621354e-03
        return _jct(B1 / dens, C1, B2 / dens, C2, B3 / dens, C3)
    if gas == "Ne":
        return _borzsonyi(9154.48e-8 / dens, 656.97e-6, 4018.63e-8 / dens, 5.728e-3)
    if gas == "Ar":
        B1, C1 = 0.00032323117217767093, 0.0045416501944977915
        B2, C2 = 0.00011557814904827939, 0.011120847461156543
        B3, C3 = 0.00010909808164540697, 0.0006827046691889898
        return _jct(B1 / dens, C1, B2 / dens, C2, B3 / dens, C3)
    if gas == "ArB":
        return _borzsonyi(20332.29e-8 / dens, 206.12e-6, 34458.31e-8 / dens, 8.066e-3)
    if gas == "Kr":
        return _borzsonyi(26102.88e-8 / dens, 2.01e-6, 56946.82e-8 / dens, 10.043e-3)
    if gas == "Xe":
        return _borzsonyi(103701.61e-8 / dens, 12750e-6, 31228.61e-8 / dens, 0.561e-3)
    if gas == "N2":
        return _borzsonyi(39209.95e-8 / dens, 1146.24e-6, 18806.48e-8 / dens, 13.476e-3)
    raise ValueError(f"Unknown gas {gas}")


def _borzsonyi(B1, C1, B2, C2):
    return lambda um: B1 * um ** 2 / (um ** 2 - C1) + B2 * um ** 2 / (um ** 2 - C2)


def _jct(B1, C1, B2, C2, B3, C3):
    return lambda um: (B1 * um ** 2 / (um ** 2 - C1) + B2 * um ** 2 / (um ** 2 - C2)
                       + B3 * um ** 2 / (um ** 2 - C3))


def sellmeier_sio2(um):
    """Fused silica (Malitson 1965), ``PhysData.jl:253-260`` (complex sqrt)."""
    um = np.asarray(um, dtype=float)
    eps = (1 + 0.6961663 / (1 - (0.0684043 / um) ** 2)
           + 0.4079426 / (1 - (0.1162414 / um) ** 2)
           + 0.8974794 / (1 - (9.896161 / um) ** 2))
    return np.sqrt(eps.astype(complex))


def ref_index_fun(material, P=1.0, T=roomtemp):
    """Refractive index n(λ[m]), ``PhysData.jl:448-470``.

    Gases: sqrt(1 + χ1) with χ1 = γ(λ)·ρ(P, T).  SiO2: Sellmeier (the reference
    defaults to a tabulated B-spline for SiO2; the cladding index enters only the
    Marcatili loss term and a ~1e-9 correction to β).
    """
    if material == "SiO2":
        return lambda lam: sellmeier_sio2(np.asarray(lam) * 1e6)
    gamma = sellmeier_gas(material)
    dens = density(material, P, T)
    return lambda lam: np.sqrt(1 + (gamma(np.asarray(lam, dtype=float) * 1e6) * dens).astype(complex))


def gamma3_gas(gas):
    """Third-order hyperpolarisability, Lehmeier model ``PhysData.jl:588-616``."""
    dens = dens_1atm_0degC(gas)
    fac = {"He": 1, "HeB": 1, "HeJ": 1, "Ne": 1.8, "Ar": 23.5, "ArB": 23.5,
           "Kr": 64.0, "Xe": 188.2, "N2": 21.1}[gas]
    return 4 * fac * 3.43e-28 / dens


def ionisation_potential(gas, unit="SI"):
    """First ionisation potential, ``PhysData.jl:731-771``."""
    Ip = {"He": 0.9036, "HeB": 0.9036, "HeJ": 0.9036, "Ne": 0.7925, "Ar": 0.5792,
          "ArB": 0.5792, "Kr": 0.5142, "Xe": 0.4458, "H": 0.5, "N2": 0.5726}[gas]
    if unit == "atomic":
        return Ip
    if unit == "eV":
        return 27.21138602 * Ip
    return 27.21138602 * electron * Ip


def quantum_numbers(gas):
    """(n, l, Z) for the PPT rate, ``PhysData.jl:773-795``."""
    if gas in ("Ar", "ArB"):
        return 3, 1, 1
    if gas == "Ne":
        return 2, 1, 1
    if gas == "Kr":
        return 4, 1, 1
    if gas == "Xe":
        return 5, 1, 1
    if gas in ("He", "HeB", "HeJ"):
        return 1, 0, 1
    if gas == "N2":
        return 2, 0, 0.9
    raise ValueError(gas)


def polarisability(gas, ion=False):
    """Static dipole polarisability (SI), ``PhysData.jl:797-830``; None if unknown."""
    tab = {"He": (1.3207, 0.2811), "Ne": (2.376, 1.2417), "Ar": (10.762, 6.807)}
    g = _base(gas)
    if g not in tab:
        return None
    return (tab[g][1] if ion else tab[g][0]) * au_polarisability


def polarisability_difference(gas):
    a, b = polarisability(gas, False), polarisability(gas, True)
    if a is None:
        return None
    return a - b


def chi1_fun(gas, P=1.0, T=roomtemp):
    gamma = sellmeier_gas(gas)
    dens = density(gas, P, T)
    return lambda lam: gamma(np.asarray(lam, dtype=float) * 1e6) * dens
