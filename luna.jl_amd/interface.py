"""Simple interface and driver: ``prop_capillary``/``prop_gnlse`` and ``run``.

Mirrors ``src/Interface.jl`` (``prop_capillary_args`` ``:360-401`` with its
builders ``:453-616, 773-835``; ``prop_gnlse_args`` ``:921-979``) and the driver
``Luna.run`` (``src/Luna.jl:358-391``).  The propagation itself (RK45 loop,
window filter, saves) runs on the GPU via ``lh_solve``.
"""
import math
import warnings

import numpy as np

from . import _capi
from . import stats as Stats
from . import physdata as pd
from .capillary import Gradient, gradient_spec
from .capillary import taper_radius as capillary_taper_radius
from .fields import GaussField, SechField, ShotNoise
from .grid import EnvGrid, RealGrid
from .ionisation import ppt_spline_cached
from .linops import make_const_linop
from .modes import MarcatiliMode, SimpleMode
from .nonlinear import (CombinedRamanResponse, Kerr_env, Kerr_field, PlasmaCumtrapz, RamanPolarEnv,
                        RamanRespNormedSingleDampedOscillator)
from .nonlinearrhs import TransModeAvg, TransRadial, norm_mode_average, norm_mode_average_gnlse


class MemoryOutput:
    """``Output.MemoryOutput(0, zmax, saveN, statsfun)`` with ``GridCondition``
    saves (``Output.jl:15-141``); ``statsfun`` from ``stats.default`` is
    recorded on the device at every accepted step."""

    def __init__(self, tmin, tmax, saveN, statsfun=None):
        self.tmin, self.tmax, self.saveN = tmin, tmax, saveN
        self.statsfun = statsfun
        self.data = {"stats": {}, "meta": {}}

    def __getitem__(self, k):
        return self.data[k]

    def __contains__(self, k):
        return k in self.data


def run(Ew, grid, linop, transform, FT=None, output=None, min_dz=0.0, max_dz=None, init_dz=1e-4,
        z0=0.0, rtol=1e-6, atol=1e-10, safety=0.9, status_period=1):
    """``Luna.run``: propagate to ``grid.zmax`` on the GPU; fills ``output``."""
    saveN = output.saveN if output is not None else 201
    opts = _capi.rk_opts(grid.zmax, z0=z0, rtol=rtol, atol=atol, safety=safety,
                         max_dz=grid.zmax / 2 if max_dz is None else max_dz, min_dz=min_dz,
                         init_dz=init_dz, save_n=saveN)
    ctx = transform.context(linop)
    statsfun = getattr(output, "statsfun", None)
    ctx.set_step_stats(statsfun)
    Ew_out, saves, zs, stats = ctx.solve(Ew, opts)
    if statsfun is not None:
        rec = ctx.step_stats()
        if stats.accepted > len(rec):
            warnings.warn(f"statistics kept for the first {len(rec)} of {stats.accepted} steps "
                          "(raise the capacity of stats.default)", RuntimeWarning)
        output.data["stats"] = statsfun.to_dict(rec)
    if output is not None:
        n = grid.n_omega
        # saves[s] holds the state column-major (ω fastest): → (Nω[, Nr], saveN)
        output.data["Eω"] = saves.reshape(saveN, -1, n).transpose(2, 1, 0).squeeze(1)
        output.data["z"] = zs[:saveN]
        output.data["grid"] = {k: getattr(grid, k) for k in ("t", "omega", "to", "omega_o", "sidx",
                                                              "omega_win", "t_win", "to_win")}
        output.data["meta"]["solver"] = stats.as_dict()
    return Ew_out, stats


def setup_radial(grid, q, densityfun, normfun, responses, inputs):
    """``Luna.setup(grid::RealGrid, q::QDHT, densityfun, normfun, responses, inputs)``,
    ``Luna.jl:239-258``: returns ``(Eωk, transform)``; inputs are summed in (ω, k)."""
    inputs = inputs if isinstance(inputs, (list, tuple)) else [inputs]
    Ewk = np.zeros((grid.n_omega, q.N), dtype=np.complex128)
    for f in inputs:
        Ewk = Ewk + f(grid, q)
    transform = TransRadial(grid, q, responses, densityfun, normfun)
    return Ewk, transform


def prop_capillary_args(radius, flength, gas, pressure, *, lambda0, lambda_lims, trange,
                        tau_fwhm=None, energy=None, power=None, pulseshape="gauss",
                        shotnoise=True, noise_seed=0, kerr=True, plasma=None, loss=True,
                        model="full", temperature=pd.roomtemp, preionfrac=0.0, saveN=201,
                        ppt_N=2 ** 16, density=None):
    """Returns ``(Eω, grid, linop, transform, FT, output)`` like ``Interface.jl:360-401``.

    ``pressure``: a number (constant fill), ``(p0, p1)`` (a gradient from z = 0
    to ``flength``, ``Capillary.gradient(gas, L, p0, p1)``) or ``(Z, P)`` (a
    multi-point gradient), as ``makemode_s``/``makedensity`` accept
    (``Interface.jl:491-532``).  ``density`` overrides the number density of a
    constant fill (the reference obtains it from CoolProp, which is not
    available here)."""
    plasma = True if plasma is None else plasma
    grid = RealGrid(flength, lambda0, lambda_lims, trange)
    grad = None
    taper = None
    if callable(radius):
        # a(z) as the reference takes it (src/Capillary.jl:60-81): sampled on
        # 1025 nodes over the fibre, piecewise linear between them
        Zt = np.linspace(0.0, flength, 1025)
        radius = (Zt, np.array([float(radius(z)) for z in Zt]))
    if isinstance(radius, (tuple, list)):
        # (a0, a1): linear from a0 at z = 0 to a1 at the fibre's end; (Z, A):
        # piecewise linear through the nodes.  Either runs through the
        # z-dependent fill, so a constant pressure becomes a flat one.
        if np.ndim(radius[0]) == 0:
            taper = (float(radius[0]), float(radius[1]), float(flength))
            radius = taper[0]
        else:
            taper = (np.asarray(radius[0], dtype=float), np.asarray(radius[1], dtype=float))
            radius = float(capillary_taper_radius(taper, 0.0))
        if not isinstance(pressure, (tuple, list)):
            pressure = (float(pressure), float(pressure))
    if isinstance(pressure, (tuple, list)):
        if np.ndim(pressure[0]) == 0:
            grad = Gradient.linear(gas, flength, float(pressure[0]), float(pressure[1]), T=temperature)
        else:
            grad = Gradient(gas, pressure[0], pressure[1], T=temperature)
        pressure = float(grad.pressure(0.0))
    mode = MarcatiliMode(radius, gas, pressure, T=temperature, model=model, loss=loss)
    if grad is not None:
        rho = grad.density(0.0)
    else:
        rho = pd.density(gas, pressure, temperature) if density is None else density
    resp = []
    if kerr:
        resp.append(Kerr_field(pd.gamma3_gas(gas)))
    if plasma:
        spl = ppt_spline_cached(gas, grid.ref_lambda, N=ppt_N)
        resp.append(PlasmaCumtrapz(grid.to, spl, pd.ionisation_potential(gas), preionfrac=preionfrac))
    if pulseshape == "gauss":
        inputs = [GaussField(lambda0, tau_fwhm, energy=energy, power=power)]
    else:
        inputs = [SechField(lambda0, tau_fwhm=tau_fwhm, energy=energy, power=power)]
    if shotnoise:
        inputs.append(ShotNoise(seed=noise_seed))
    linop, beta, beta1 = make_const_linop(grid, mode, grid.ref_lambda)
    aeff = mode.Aeff()
    norm = norm_mode_average(grid, beta, aeff)
    spec = gradient_spec(grid, mode, grad, grid.ref_lambda, taper=taper) if grad is not None else None
    transform = TransModeAvg(grid, resp, rho, norm, aeff, gradient=spec)
    Ew = sum(f(grid) for f in inputs)
    stats = Stats.default(grid, Ew, mode, linop, transform, gas=gas)
    output = MemoryOutput(0, grid.zmax, saveN, stats)
    return Ew, grid, linop, transform, None, output


def prop_capillary(*args, **kw):
    Ew, grid, linop, transform, FT, output = prop_capillary_args(*args, **kw)
    run(Ew, grid, linop, transform, FT, output)
    return output


def prop_gnlse_args(gamma, flength, betas, *, lambda0, lambda_lims, trange, tau_fwhm=None,
                    energy=None, power=None, pulseshape="gauss", shotnoise=True, noise_phases=None,
                    noise_seed=0, shock=True, loss=0.0, raman=True, fr=0.18, ramanmodel="sdo",
                    tau1=12.2e-15, tau2=32e-15, saveN=201):
    """``prop_gnlse_args``, ``Interface.jl:921-979`` (envelope grid, Kerr + SDO Raman)."""
    grid = EnvGrid(flength, lambda0, lambda_lims, trange, thg=False)
    mode = SimpleMode(pd.wlfreq(lambda0), betas, loss=loss)
    aeff = 1.0
    linop, beta, beta1 = make_const_linop(grid, mode, lambda0)
    k0 = 2 * math.pi / lambda0
    n2 = gamma / k0 * aeff
    chi3 = 4 / 3 * (1 - fr) * n2 * (pd.eps_0 * pd.c)
    resp = [Kerr_env(chi3)]
    if raman:
        if ramanmodel != "sdo":
            raise ValueError("only the :sdo Raman model is provided")
        chi3R = 2 * fr * n2 * (pd.eps_0 * pd.c)
        R = CombinedRamanResponse(grid.to, [RamanRespNormedSingleDampedOscillator(chi3R * pd.eps_0, 1 / tau1, tau2)])
        resp.append(RamanPolarEnv(grid.to, R))
    if pulseshape == "gauss":
        inputs = [GaussField(lambda0, tau_fwhm, energy=energy, power=power)]
    else:
        inputs = [SechField(lambda0, tau_fwhm=tau_fwhm, energy=energy, power=power)]
    if shotnoise:
        inputs.append(ShotNoise(seed=noise_seed, phases=noise_phases))
    norm = norm_mode_average_gnlse(grid, aeff, shock=shock)
    transform = TransModeAvg(grid, resp, 1.0, norm, aeff)
    Ew = sum(f(grid) for f in inputs)
    stats = Stats.default(grid, Ew, mode, linop, transform)
    output = MemoryOutput(0, grid.zmax, saveN, stats)
    return Ew, grid, linop, transform, None, output


def prop_gnlse(*args, **kw):
    Ew, grid, linop, transform, FT, output = prop_gnlse_args(*args, **kw)
    run(Ew, grid, linop, transform, FT, output)
    return output
