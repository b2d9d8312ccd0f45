"""Named workloads: the capillary (UPPE, field grid), GNLSE and scan set-ups.

``capillary_instance`` is ``prop_capillary(125e-6, flength, :He, P; λ0=800e-9,
τfwhm=10e-15, energy, λlims=(150e-9, 4e-6), trange, shotnoise=false)`` with
Kerr + PPT plasma.  ``scan_instances`` draws the (energy, pressure) sweep of
the batched scan: energy U[50, 200] μJ, pressure U[0.5, 3] bar, half of the
points on a 0.5 ps window (Nto = 8192) and half on 1 ps (Nto = 16384).
"""
from dataclasses import dataclass, field

import numpy as np

from . import physdata as pd
from .fields import GaussGaussField
from .grid import EnvGrid, RealGrid
from .hankel import QDHT
from .interface import prop_capillary_args, prop_gnlse_args, setup_radial
from .ionisation import ppt_spline_cached
from .linops import make_const_linop_radial
from .nonlinear import Kerr_env, Kerr_field, PlasmaCumtrapz
from .nonlinearrhs import const_norm_radial


@dataclass
class Instance:
    holder: object
    Ew0: np.ndarray
    zmax: float
    grid: object
    linop: np.ndarray
    transform: object
    sidx: np.ndarray
    params: dict = field(default_factory=dict)


def capillary_instance(energy=120e-6, pressure=1.0, trange=1e-12, flength=1.0, radius=125e-6,
                       gas="He", lambda0=800e-9, tau_fwhm=10e-15, lambda_lims=(150e-9, 4e-6),
                       ppt_n=2 ** 16, plasma=True, kerr=True):
    Ew, grid, linop, transform, _, _ = prop_capillary_args(
        radius, flength, gas, pressure, lambda0=lambda0, lambda_lims=lambda_lims, trange=trange,
        tau_fwhm=tau_fwhm, energy=energy, shotnoise=False, plasma=plasma, kerr=kerr, ppt_N=ppt_n)
    holder = transform.descriptor(linop)
    return Instance(holder=holder, Ew0=np.ascontiguousarray(Ew), zmax=flength, grid=grid, linop=linop,
                    transform=transform, sidx=grid.sidx.copy(),
                    params=dict(energy=energy, pressure=pressure, trange=trange, flength=flength,
                                n_to=grid.n_to, n_omega=grid.n_omega))


def scan_points(n=256, seed=1):
    """(energy, pressure, trange) for the batched sweep, seeded."""
    rng = np.random.default_rng(seed)
    energy = rng.uniform(50e-6, 200e-6, n)
    pressure = rng.uniform(0.5, 3.0, n)
    trange = np.where(np.arange(n) < n // 2, 0.5e-12, 1e-12)
    return energy, pressure, trange


def scan_instances(n=256, seed=1, flength=1.0, ppt_n=2 ** 16):
    E, P, T = scan_points(n, seed)
    return [capillary_instance(energy=float(e), pressure=float(p), trange=float(t), flength=flength,
                               ppt_n=ppt_n) for e, p, t in zip(E, P, T)]


def gnlse_instance(trange=6e-12, flength=0.15, noise_seed=0, shotnoise=True):
    """Supercontinuum in solid-core fibre (Dudley et al. RMP 78, 1135 (2006), Fig. 3)."""
    betas = [0.0, 0.0, -11.830e-27, 8.1038e-41, -9.5205e-56, 2.0737e-70, -5.3943e-85,
             1.3486e-99, -2.5495e-114, 3.0524e-129, -1.7140e-144]
    Ew, grid, linop, transform, _, _ = prop_gnlse_args(
        0.11, flength, betas, lambda0=835e-9, lambda_lims=(400e-9, 2400e-9), trange=trange,
        tau_fwhm=50e-15, power=1e4, pulseshape="sech", shotnoise=shotnoise, noise_seed=noise_seed,
        raman=True, ramanmodel="sdo", tau1=12.2e-15, tau2=32e-15, fr=0.18, shock=True, loss=0.0)
    holder = transform.descriptor(linop)
    return Instance(holder=holder, Ew0=np.ascontiguousarray(Ew), zmax=flength, grid=grid, linop=linop,
                    transform=transform, sidx=grid.sidx.copy(), params=dict(n_to=grid.n_to))


def radial_instance(gas="Ar", pressure=0.5, flength=0.5, lambda0=800e-9, tau_fwhm=30e-15, energy=200e-6,
                    w0=40e-6, propz=-0.157, R=6e-3, N=256, trange=1e-12,
                    lambda_lims=(400e-9, 2000e-9), kerr=True, plasma=True, ppt_n=2 ** 16):
    """Radially symmetric free-space propagation (TransRadial), as in the
    reference's radial example: a Gaussian beam whose 1 mm waist at z = 0
    (``propz`` before a 40 μm focus) is focused into Ar; Kerr + PPT plasma.
    With the defaults Nt = 4096, Nto = 8192, Nω = 2049 and Nr = 256."""
    grid = RealGrid(flength, lambda0, lambda_lims, trange)
    q = QDHT(R, N, dim=2)
    rho = pd.density(gas, pressure)
    resp = []
    if kerr:
        resp.append(Kerr_field(pd.gamma3_gas(gas)))
    if plasma:
        spl = ppt_spline_cached(gas, grid.ref_lambda, N=ppt_n)
        resp.append(PlasmaCumtrapz(grid.to, spl, pd.ionisation_potential(gas)))
    nfun = pd.ref_index_fun(gas, pressure)
    linop, beta1 = make_const_linop_radial(grid, q, nfun)
    normfun = const_norm_radial(grid, q, nfun)
    inputs = GaussGaussField(lambda0, tau_fwhm, energy, w0, propz=propz)
    Ewk, transform = setup_radial(grid, q, rho, normfun, resp, inputs)
    holder = transform.descriptor(linop)
    return Instance(holder=holder, Ew0=np.asfortranarray(Ewk), zmax=flength, grid=grid, linop=linop,
                    transform=transform, sidx=grid.sidx.copy(),
                    params=dict(energy=energy, pressure=pressure, trange=trange, flength=flength, R=R, N=N,
                                n_to=grid.n_to, n_omega=grid.n_omega, n_r=N, q=q))


def radial_env_instance(gas="Ar", pressure=1.2, flength=0.6, lambda0=800e-9, tau_fwhm=20e-15, energy=2e-9,
                        w0=40e-6, propz=-0.3, R=4e-3, N=128, trange=1e-12, lambda_lims=(400e-9, 2000e-9)):
    """Radially symmetric propagation on the envelope grid (TransRadial on an
    EnvGrid, Kerr_env only), the reference's free-space envelope example
    (examples/low_level_interface/freespace/radial_env.jl: Ar at 1.2 bar, 20 fs
    at 800 nm, 2 nJ, 40 μm waist 0.3 m after the input, R = 4 mm) with a 1 ps
    window (Nt = Nto = 1024) and N = 128 radial points by default."""
    grid = EnvGrid(flength, lambda0, lambda_lims, trange)
    q = QDHT(R, N, dim=2)
    rho = pd.density(gas, pressure)
    nfun = pd.ref_index_fun(gas, pressure)
    linop, beta1 = make_const_linop_radial(grid, q, nfun)
    normfun = const_norm_radial(grid, q, nfun)
    inputs = GaussGaussField(lambda0, tau_fwhm, energy, w0, propz=propz)
    Ewk, transform = setup_radial(grid, q, rho, normfun, [Kerr_env(pd.gamma3_gas(gas))], inputs)
    holder = transform.descriptor(linop)
    return Instance(holder=holder, Ew0=np.asfortranarray(Ewk), zmax=flength, grid=grid, linop=linop,
                    transform=transform, sidx=grid.sidx.copy(),
                    params=dict(energy=energy, pressure=pressure, trange=trange, flength=flength, R=R, N=N,
                                n_to=grid.n_to, n_omega=grid.n_omega, n_r=N, q=q))
