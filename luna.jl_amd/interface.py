"""Simple interface and driver: ``prop_capillary``/``prop_gnlse`` and ``run``.

Mirrors ``src/Interface.jl`` (``prop_capillary_args`` ``:360-401`` with its
builders ``:453-616, 773-835``; ``prop_gnlse_args`` ``:921-979``) and the driver
``Luna.run`` (``src/Luna.jl:358-391``).  The propagation itself (RK45 loop,
window filter, saves) runs on the GPU via ``lh_solve``.
"""
import logging
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
                        RamanPolarField, RamanRespNormedSingleDampedOscillator)
from .nonlinearrhs import TransModeAvg, TransRadial, norm_mode_average, norm_mode_average_gnlse
from .output import HDF5Output, MemoryOutput  # noqa: F401 (re-exported: Output.MemoryOutput)
from .stats import StepRecords

log = logging.getLogger("luna_jl_amd")


def simtype(grid, transform, linop):
    """``simtype`` (``src/Luna.jl:327-336``): the ``simulation_type`` group."""
    variable = callable(linop) or getattr(transform, "gradient", None) is not None
    return {"field": "field-resolved" if grid.kind == "real" else "envelope",
            "transform": str(transform), "linop": "variable" if variable else "constant"}


def run(Ew, grid, linop, transform, FT=None, output=None, min_dz=0.0, max_dz=None, init_dz=1e-4,
        z0=0.0, rtol=1e-6, atol=1e-10, safety=0.9, status_period=1):
    """``Luna.run`` (``src/Luna.jl:358-391``): propagate to ``grid.zmax`` on the
    GPU.  A cached propagation in the output (``HDF5Output``) is resumed from
    its cache; the grid and simulation type are written to the output; the
    device solve fills a ``MemoryOutput`` from its save buffer and streams
    every saving step to any other output (``lh_solve_stream``).  Statistics
    evaluated on the host (``collect_stats``, ``Stats.default``'s userfuns)
    stream every accepted step and are evaluated on its filtered state.
    Device statistics that outgrow ``capacity`` in a ``MemoryOutput`` run are
    recomputed by a second solve with room for every accepted step; other
    outputs keep the first ``capacity`` records and warn."""
    if output is None:
        output = MemoryOutput(0, grid.zmax, 201)
    save_first = 0
    Ewc, zc, dzc = output.check_cache(Ew, z0, init_dz)
    if zc > z0:
        log.info("Found cached propagation. Resuming...")
        Ew, z0, init_dz = Ewc, zc, dzc
        save_first = output.saved
    output(grid.to_dict(), group="grid")
    output(simtype(grid, transform, linop), group="simulation_type")
    statsfun = output.statsfun
    host = statsfun is not None and getattr(statsfun, "has_host", False)
    zdep = getattr(transform, "gradient", None) is not None
    ctx = transform.context(linop)
    host_records = False  # the device records replaced by the same records computed on the host
    if hasattr(statsfun, "desc"):
        try:
            ctx.set_step_stats(statsfun)
        except _capi.LunaHipError as e:
            # the grids staged through HBM record no statistics on the device
            if e.status != _capi.LH_E_ARG or zdep:
                raise
            ctx.set_step_stats(None)
            host = host_records = True
    else:
        ctx.set_step_stats(None)
    if host and zdep:
        raise NotImplementedError("statistics evaluated on the host are not provided with a z-dependent fill")
    opts = _capi.rk_opts(grid.zmax, z0=z0, rtol=rtol, atol=atol, safety=safety,
                         max_dz=grid.zmax / 2 if max_dz is None else max_dz, min_dz=min_dz,
                         init_dz=init_dz, save_n=output.saveN, save_z0=output.tmin, save_first=save_first,
                         every_step=host)
    shape = (grid.n_omega,) + ((transform.n_r,) if transform.n_r > 1 else ())
    if not host and (isinstance(output, MemoryOutput) or zdep):
        # the whole run in the device save buffer, handed over at the end
        # (z-dependent fills are not streamed)
        Ew_out, saves, zs, stats = ctx.solve(Ew, opts)
        if hasattr(statsfun, "desc") and stats.status == _capi.LH_OK and stats.accepted > statsfun.capacity:
            # the reference's statistics arrays grow with the run: solve again
            # (the same trajectory) with a record for every accepted step
            statsfun.capacity = int(stats.accepted)
            ctx.set_step_stats(statsfun)
            Ew_out, saves, zs, stats = ctx.solve(Ew, opts)
        _capi.check(stats.status, "Luna.run")
        rec = ctx.step_stats() if statsfun is not None else np.zeros((0, _capi.LH_NSTAT))
        n = output.saveN - save_first
        # saves[s] holds the state column-major (ω fastest): → (Nω[, Nr], saveN)
        sv = saves[save_first:].reshape(n, -1, grid.n_omega).transpose(2, 1, 0).reshape(shape + (n,), order="F")
        output.on_step(save_first, sv, zs[save_first:output.saveN], stats.z, stats.dz,
                       Ew_out.reshape(shape, order="F"), rec)
    else:
        def on_step(first, saves, zsv, z, dz, y, records):
            saves = saves.reshape(shape + (saves.shape[-1],), order="F")
            y = y.reshape(shape, order="F")
            if host:  # one accepted step per call
                if host_records:
                    records = statsfun.host_record(y, z, dz)[None, :]
                records = StepRecords(records, [statsfun.host_step(y, z, dz)])
            return output.on_step(first, saves, zsv, z, dz, y, records)

        Ew_out, stats = ctx.solve_stream(Ew, opts, on_step)
        if stats.status != _capi.LH_E_STOPPED:
            _capi.check(stats.status, "Luna.run")
    if hasattr(statsfun, "desc") and not host_records and stats.accepted > statsfun.capacity:
        warnings.warn(f"statistics kept for the first {statsfun.capacity} of {stats.accepted} steps "
                      "(raise the capacity of stats.default)", RuntimeWarning)
    output("solver", stats.as_dict(), meta=True, force=True)
    return Ew_out, stats


def setup_radial(grid, q, densityfun, normfun, responses, inputs):
    """``Luna.setup(grid, q::QDHT, densityfun, normfun, responses, inputs)``,
    ``Luna.jl:239-279`` (field and envelope grids): returns ``(Eωk,
    transform)``; inputs are summed in (ω, k)."""
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
                        ppt_N=2 ** 16, density=None, raman=None, rotation=True, vibration=True):
    """Returns ``(Eω, grid, linop, transform, FT, output)`` like ``Interface.jl:360-401``.

    ``pressure``: a number (constant fill), ``(p0, p1)`` (a gradient from z = 0
    to ``flength``, ``Capillary.gradient(gas, L, p0, p1)``) or ``(Z, P)`` (a
    multi-point gradient), as ``makemode_s``/``makedensity`` accept
    (``Interface.jl:491-532``).  ``density`` overrides the number density of a
    constant fill (the reference obtains it from CoolProp, which is not
    available here).  ``raman`` (default: on for the molecular gases, as
    ``makeresponse``, ``Interface.jl:534-558``) adds ``RamanPolarField`` with
    the gas's rotational and vibrational response."""
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
    if raman is None:
        raman = gas in ("N2", "H2", "D2", "N2O", "CH4", "SF6")
    if raman:
        from .raman import raman_response
        resp.append(RamanPolarField(grid.to, raman_response(grid.to, gas, rotation=rotation, vibration=vibration,
                                                            temp=temperature)))
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
