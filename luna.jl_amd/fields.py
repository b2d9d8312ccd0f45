"""Input fields, restating ``src/Fields.jl``.

``GaussField``/``SechField`` (``Fields.jl:66-108``) evaluated by the
``PulseField`` call (``:129-144``) with energy normalisation via
``energyfuncs`` (``:510-520``) or peak-power normalisation; ``ShotNoise``
(``:290-320``) with the random phases passed in explicitly (the reference draws
them from Julia's global RNG, which cannot be reproduced here).
"""
import math

import numpy as np

from . import physdata as pd
from .maths import hilbert, simpson_even


def gauss(x, fwhm, power=2):
    sigma = fwhm / (2 * (2 * math.log(2)) ** (1 / power))
    return np.exp(-0.5 * (x / sigma) ** power)


class PulseField:
    def __init__(self, lam0, itshape, energy=None, power=None):
        if (energy is None) == (power is None):
            raise ValueError("exactly one of energy or power must be given")
        self.lam0, self.itshape, self.energy, self.power = lam0, itshape, energy, power

    def make_Et(self, grid):
        t = grid.t
        w0 = pd.wlfreq(self.lam0)
        if grid.kind == "real":
            return np.sqrt(self.itshape(t)) * np.cos(w0 * t)
        dw = pd.wlfreq(self.lam0) - grid.omega0
        return np.sqrt(self.itshape(t)) * np.exp(1j * (dw * t))

    def __call__(self, grid):
        Et = self.make_Et(grid)
        dt = grid.t[1] - grid.t[0]
        if self.energy is not None:
            It = np.abs(hilbert(Et)) ** 2 if grid.kind == "real" else np.abs(Et) ** 2
            Et = Et * (math.sqrt(self.energy) / math.sqrt(simpson_even(dt, It)))
        else:
            It = np.abs(hilbert(Et)) ** 2 if grid.kind == "real" else np.abs(Et) ** 2
            Et = Et * (math.sqrt(self.power) / math.sqrt(It.max()))
        if grid.kind == "real":
            return np.fft.rfft(Et)
        return np.fft.fft(Et)


def GaussField(lam0, tau_fwhm, energy=None, power=None, m=1):
    return PulseField(lam0, lambda t: gauss(t, tau_fwhm, power=2 * m), energy=energy, power=power)


def SechField(lam0, tau_fwhm=None, tau_w=None, energy=None, power=None):
    if tau_w is None:
        tau_w = tau_fwhm / (2 * math.log(1 + math.sqrt(2)))
    def sech2(t):
        e = np.exp(-2 * np.abs(t / tau_w))
        return 4 * e / (1 + e) ** 2  # sech² without overflow in cosh

    return PulseField(lam0, sech2, energy=energy, power=power)


class ShotNoise:
    """One-photon-per-mode noise with explicit uniform phases φ∈[0,1) (×2π)."""

    def __init__(self, seed=0, phases=None):
        self.seed, self.phases = seed, phases

    def __call__(self, grid):
        dw = grid.omega[1] - grid.omega[0]
        dt = grid.t[1] - grid.t[0]
        u = self.phases if self.phases is not None else np.random.default_rng(self.seed).random(len(grid.omega))
        phi = 2 * math.pi * u
        if grid.kind == "real":
            amp = np.sqrt(pd.hbar * grid.omega / dw)
            return math.sqrt(2 * math.pi) / (2 * dt) * amp * np.exp(1j * phi)
        amp = np.zeros_like(grid.omega)
        amp[grid.sidx] = np.sqrt(pd.hbar * grid.omega[grid.sidx] / dw)
        return math.sqrt(2 * math.pi) / dt * amp * np.exp(1j * phi)


def energy_t(grid, Et):
    dt = grid.t[1] - grid.t[0]
    It = np.abs(hilbert(Et)) ** 2 if grid.kind == "real" else np.abs(Et) ** 2
    return simpson_even(dt, It)


def energyfuncs(grid, q=None):
    """``Fields.energyfuncs(grid[, q])`` (``src/Fields.jl:510-565``): the pair
    (energy_t(Et), energy_ω(Eω)); with a QDHT q the radial forms, Et and Eω of
    shape (N, Nr) in real and reciprocal space."""
    dw = grid.omega[1] - grid.omega[0]
    if q is None:
        if grid.kind == "real":
            pre = 2 * math.pi / grid.omega[-1] ** 2
            return (lambda Et: energy_t(grid, Et)), (lambda Ew: pre * simpson_even(dw, np.abs(Ew) ** 2))
        pre = 2 * math.pi * dw / (len(grid.omega) * dw) ** 2
        return (lambda Et: energy_t(grid, Et)), (lambda Ew: pre * np.sum(np.abs(Ew) ** 2))
    if grid.kind != "real":
        # Fields.jl:556-570: |Et|² of the envelope, Σ|Eω|² over ω
        dw = grid.omega[1] - grid.omega[0]
        pre = 2 * math.pi * pd.c * pd.eps_0 / 2 * 2 * math.pi * dw / (len(grid.omega) * dw) ** 2
        return (lambda Et: energy_t_radial(grid, q, Et)), \
            (lambda Ew: pre * q.integrateK(np.sum(np.abs(Ew) ** 2, axis=0)))
    return (lambda Et: energy_t_radial(grid, q, Et)), (lambda Ew: energy_omega_radial(grid, q, Ew))


# --------------------------------------------------------------------------
# Radially symmetric inputs on a QDHT grid (field grid only)
# --------------------------------------------------------------------------
def gauss_sigma(x, sigma):
    """``Maths.gauss(x, σ)``."""
    return np.exp(-0.5 * (x / sigma) ** 2)


def energy_t_radial(grid, q, Et):
    """``energyfuncs(grid, q::QDHT)[1]``, ``Fields.jl:536-541`` (field grid: of
    the analytic signal) and ``:556-559`` (envelope grid); Et is (Nt, Nr)."""
    dt = grid.t[1] - grid.t[0]
    Eta = hilbert(Et, axis=0) if grid.kind == "real" else Et
    tintg = simpson_even(dt, np.abs(Eta) ** 2)
    return 2 * math.pi * pd.c * pd.eps_0 / 2 * q.integrateR(tintg)


def energy_omega_radial(grid, q, Ew):
    """``energyfuncs(grid::RealGrid, q::QDHT)[2]``, ``Fields.jl:543-548``; Ew is (Nω, Nr)."""
    dw = grid.omega[1] - grid.omega[0]
    prefac = 2 * math.pi * pd.c * pd.eps_0 / 2 * 2 * math.pi / (grid.omega[-1] ** 2)
    wintg = simpson_even(dw, np.abs(Ew) ** 2)
    return prefac * q.integrateK(wintg)


def prop_radial(Ewk, z, grid, q):
    """Free-space propagation by z in (ω, k) space, ``Fields.jl:401-406`` (in place)."""
    k0 = grid.omega / pd.c
    kzsq = k0[:, None] ** 2 - (q.k ** 2)[None, :]
    kzsq[kzsq < 0] = 0
    kz = np.sqrt(kzsq)
    Ewk *= np.exp(-1j * z * (kz - k0[:, None]))
    return Ewk


class GaussGaussField:
    """``Fields.GaussGaussField`` (``Fields.jl:343-364``) evaluated as a
    ``SpatioTemporalField`` (``Fields.jl:366-398``) on a field grid and QDHT:
    Gaussian in time (intensity FWHM τfwhm, super-Gaussian order m) and in r
    (waist w0), scaled to ``energy``, transformed to (ω, k) and propagated by
    ``propz`` from the waist."""

    def __init__(self, lam0, tau_fwhm, energy, w0, phi=0.0, tau0=0.0, m=1, propz=0.0):
        self.lam0, self.tau_fwhm, self.energy, self.w0 = lam0, tau_fwhm, energy, w0
        self.phi, self.tau0, self.m, self.propz = phi, tau0, m, propz

    def make_Etr(self, grid, q):
        """``make_Etr``, ``Fields.jl:366-376``: cos(ω0 t + ϕ) on the field grid,
        exp(i(ϕ + Δω t)) with Δω = ω(λ0) − ω0 on the envelope grid."""
        t = grid.t - self.tau0
        I = gauss(t, self.tau_fwhm, power=2 * self.m)[:, None] * gauss_sigma(q.r, self.w0 / 2)[None, :]
        if grid.kind != "real":
            dw = pd.wlfreq(self.lam0) - grid.omega0
            return np.sqrt(I) * np.exp(1j * (self.phi + dw * t))[:, None]
        w0 = pd.wlfreq(self.lam0)
        return np.sqrt(I) * np.cos(w0 * t + self.phi)[:, None]

    def __call__(self, grid, q):
        Etr = self.make_Etr(grid, q)
        Etr = Etr * (math.sqrt(self.energy) / math.sqrt(energy_t_radial(grid, q, Etr)))
        ft = np.fft.rfft if grid.kind == "real" else np.fft.fft  # FFTW's forward plans (unnormalised)
        Ewk = q.forward(ft(Etr, axis=0))
        if self.propz != 0.0:
            prop_radial(Ewk, self.propz, grid, q)
        return Ewk
