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
