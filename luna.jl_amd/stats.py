"""Per-step statistics of a mode-averaged propagation: the drop-in for
``Stats.default`` (``src/Stats.jl:476-532``) and the ``collect_stats`` closure
it returns (``src/Stats.jl:454-468``).

The reference evaluates these functions on the host after every accepted step
(``MemoryOutput``, ``src/Output.jl:65-67``), including an inverse FFT for the
analytic field and a PPT-rate pass.  Here they run on the GPU inside the
propagation kernel, right after the step's window filter, and write one record
of ``_capi.LH_NSTAT`` values per accepted step:

  ω0 (``:18-25``), energy (``:32-43`` with ``Fields.energyfuncs``),
  peakpower (``:85-99``), peakintensity (``:150-154``), fwhm_t_min/max
  (``:182-199``), electrondensity and peak_ionisation_rate (``:219-241``, field
  grid with ``PlasmaCumtrapz`` only), density (``:325-329``), z and dz
  (``zdz!``, ``:409-412``).

``windows`` (``energy_λ``/``energy_window``, ``:52-79``): up to four
wavelength windows, each a Planck-taper window over ω whose energy
energyfun_ω(Eω .* window) is recorded on the device as Σ_k energy_w[k]·win_k²·
|Eω_k|², keyed ``energy_<λmin>nm_<λmax>nm`` as the reference labels them.

``onaxis`` (``peakintensity(grid, (mode,))`` and ``electrondensity(..., (mode,))``,
``:159-174, 243-262``): the field at r = 0 of a single HE1m mode is the
mode-averaged field scaled by a constant, so both statistics are the
mode-averaged ones with Aeff replaced by the mode's P/I(0) (``aeff_onaxis``).

``pressure`` (``:335-339``, PhysData.pressure of the density) is filled on the
host from each record's density with the gas model, and ``zdw`` (``zdw_linop``,
``:571-582``: ``Modes.zdw``, a root search on the mode's dispersion) from the
mode: once for a constant fill, at each record's z for a gradient fill.

``userfuns`` (functions f(d, Eω, Et, z, dz) that add entries to d, ``:523-529``)
and ``collect_stats`` (``:476-493``) with any functions — the statistics of a
radial propagation, which ``Stats.default`` does not cover — run on the host:
the solve then hands every accepted step's filtered state to the host
(``lh_solve_stream`` with ``every_step``), which evaluates them on it and on its
analytic field (``plan_analytic``, ``:423-470``).  Host restatements of the
reference's statistics functions for that use: ``omega0``, ``energy``,
``energy_window``, ``energy_lambda``, ``peakpower``, ``fwhm_t``, ``density``,
``pressure``, ``zdz``.
"""
import math

import numpy as np

from . import _capi
from . import physdata as pd
from .maths import cspline_eval_vec
from .maths import fwhm as _fwhm
from .nonlinear import PlasmaCumtrapz


def planck_taper(x, left0, left1, right1, right0):
    """``Maths.planck_taper(x, left0, left1, right1, right0)``
    (``src/Maths.jl:422-449``): 0 below left0, rising to 1 between left0 and
    left1, 1 up to right1 (both ends included), falling to 0 at right0."""
    xc = np.asarray(x, dtype=float) - (right0 + left0) / 2
    X = right0 - left0
    el = abs(left1 - left0) / X
    er = abs(right0 - right1) / X
    x1, x2, x3, x4 = -X / 2, -X / 2 * (1 - 2 * el), X / 2 * (1 - 2 * er), X / 2
    out = np.zeros_like(xc)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        m = (x1 < xc) & (xc < x2)
        out[m] = 1 / (1 + np.exp((x2 - x1) / (xc[m] - x1) + (x2 - x1) / (xc[m] - x2)))
        out[(x2 <= xc) & (xc <= x3)] = 1.0
        m = (x3 < xc) & (xc < x4)
        out[m] = 1 / (1 + np.exp((x3 - x4) / (xc[m] - x3) + (x3 - x4) / (xc[m] - x4)))
    return out


def energy_window_spec(grid, lambda_lims, winwidth=0.0):
    """(label, window over grid.omega) of ``energy_λ(grid, energyfun_ω, λlims;
    winwidth)``, ``src/Stats.jl:52-61``."""
    lam = np.asarray(lambda_lims, dtype=float)
    wmin, wmax = sorted(pd.wlfreq(lam))
    win = planck_taper(grid.omega, wmin - winwidth, wmin, wmax, wmax + winwidth)
    label = "%.2fnm_%.2fnm" % (1e9 * lam.min(), 1e9 * lam.max())
    return label, win


def energy_weights(grid):
    """Weights w with energy_ω(Eω) = Σ w |Eω|², ``Fields.energyfuncs``
    (``src/Fields.jl:510-534``).

    Field grid: 2π/ω[end]² · ∫|Eω|² dω with NumericalIntegration's SimpsonEven
    (extended Simpson rule, end weights 17, 59, 43, 49 over 48).  Envelope grid:
    2π δω/Δω² · Σ|Eω|² with Δω = Nω δω."""
    w = np.asarray(grid.omega, dtype=float)
    n = len(w)
    dw = w[1] - w[0]
    if grid.kind == "real":
        if n < 8:
            raise ValueError("SimpsonEven needs at least 8 points here")
        prefac = 2 * math.pi / (w[-1] ** 2)
        q = np.ones(n)
        q[:4] = np.array([17, 59, 43, 49]) / 48
        q[-4:] = np.array([49, 43, 59, 17]) / 48
        return prefac * dw * q
    Dw = n * dw
    prefac = 2 * math.pi * dw / (Dw ** 2)
    return np.full(n, prefac)


class StatsFun:
    """What ``Stats.default`` returns: callable as ``stats(Eω, z, dz)`` → dict,
    and the description the propagation kernel uses to record the same values
    at every accepted step (``desc()``)."""

    def __init__(self, grid, transform, capacity=16384, gas=None, zdw=None, windows=None, aeff=None,
                 userfuns=()):
        if getattr(transform, "kind", None) != _capi.LH_TRANS_MODEAVG:
            raise NotImplementedError("Stats.default is provided for mode-averaged transforms "
                                      "(radial propagations: collect_stats)")
        self.grid = grid
        self.userfuns = list(userfuns)
        self.has_host = bool(self.userfuns)  # evaluated on the host at every accepted step
        self._ukeys = []
        self.transform = transform
        self.capacity = int(capacity)
        self.energy_w = np.ascontiguousarray(energy_weights(grid))
        self.t = np.ascontiguousarray(grid.t, dtype=float)
        self.aeff = float(transform.aeff if aeff is None else aeff)  # on-axis statistics: P/I(0)
        self.plasma = grid.kind == "real" and any(isinstance(r, PlasmaCumtrapz) for r in transform.responses)
        self.gas = gas
        self.zdw = zdw  # None, a number, or a function of z
        # energy windows: λlims pairs, or (λlims, winwidth)
        self.windows = []
        for w in (windows or ()):
            lims, ww = (w[0], w[1]) if np.ndim(w[0]) == 1 else (w, 0.0)
            self.windows.append(energy_window_spec(grid, lims, ww))
        if len(self.windows) > _capi.LH_STAT_MAX_WINDOWS:
            raise ValueError(f"at most {_capi.LH_STAT_MAX_WINDOWS} energy windows")
        self.window_w = (np.ascontiguousarray([self.energy_w * win ** 2 for _, win in self.windows])
                         if self.windows else None)

    def desc(self):
        d = _capi.LhStatsDesc()
        d.capacity = self.capacity
        d.energy_w = _capi._ptr(self.energy_w)
        d.t = _capi._ptr(self.t)
        d.aeff = self.aeff
        d.n_windows = len(self.windows)
        d.window_w = _capi._ptr(self.window_w) if self.windows else None
        self._keep = d  # the arrays stay owned by self
        return d

    def keys(self):
        keys = [k for k in _capi.STAT_KEYS if not k.startswith("window")]
        if not self.plasma:
            keys = [k for k in keys if k not in ("electrondensity", "peak_ionisation_rate")]
        return keys

    def host_record(self, Ew, z, dz):
        """The record the propagation kernel writes (the same values and
        formulas), evaluated on the host: for the grids whose transforms are
        staged through HBM, where the kernel records no statistics."""
        K = _capi.STAT_KEYS.index
        g = self.grid
        Ew = np.asarray(Ew).reshape(-1)
        y = np.abs(Ew) ** 2
        rec = np.full(_capi.LH_NSTAT, np.nan)
        rec[K("z")], rec[K("dz")] = z, dz
        rec[K("ω0")] = np.sum(np.asarray(g.omega, dtype=float) * y) / np.sum(y)
        rec[K("energy")] = np.sum(self.energy_w * y)
        for q in range(len(self.windows)):
            rec[K(f"window{q}")] = np.sum(self.window_w[q] * y)
        Et = analytic(g, Ew)
        P = np.abs(Et) ** 2
        pmax = float(np.max(P))
        rec[K("peakpower")] = pmax
        rec[K("peakintensity")] = pmax / self.aeff
        rec[K("fwhm_t_min")] = _fwhm(self.t, P)
        rec[K("fwhm_t_max")] = _fwhm(self.t, P, "max")
        rho = self.transform.density
        rec[K("density")] = rho
        if self.plasma:
            # electrondensity (src/Stats.jl:219-241): the PPT rate of the real
            # field in V/m, its cumulative integral, 1 − exp(−∫rate) at the end
            E = np.abs(np.real(Et) / math.sqrt(pd.eps_0 * pd.c * self.aeff / 2))
            spl = next(r.ratefunc for r in self.transform.responses if isinstance(r, PlasmaCumtrapz))
            if not np.all(E <= spl.emax):
                raise RuntimeError("field strength exceeds maximum for PPT ionisation rate")
            rate = np.where(E < spl.emin, 0.0, np.exp(cspline_eval_vec(spl.x, spl.y, spl.d, E)))
            dt = float(self.t[1] - self.t[0])
            F = np.sum(0.5 * (rate[1:] + rate[:-1])) * dt
            rec[K("electrondensity")] = (1 - math.exp(-F)) * rho
            rec[K("peak_ionisation_rate")] = float(np.max(rate))
        return rec

    def names(self):
        keys = list(self.to_dict(np.zeros((0, _capi.LH_NSTAT))).keys())
        return keys + [k for k in self._ukeys if k not in keys]

    def host_step(self, Ew, z, dz):
        """The userfuns' values at one accepted step."""
        d = {}
        if self.userfuns:
            Et = analytic(self.grid, Ew)
            for uf in self.userfuns:
                uf(d, Ew, Et, z, dz)
            if not self._ukeys:
                self._ukeys = list(d)
        return d

    def to_dict(self, records):
        """Records [n][LH_NSTAT] (or a ``StepRecords``) → {name: array over
        steps} (plus ``pressure`` from the recorded density when a gas is
        given, and the userfuns' values, which replace a device value of the
        same name as a later function does in the reference)."""
        host = ()
        if isinstance(records, StepRecords):
            host, records = records.host, records.device
        records = np.asarray(records, dtype=float).reshape(-1, _capi.LH_NSTAT)
        out = {k: records[:, _capi.STAT_KEYS.index(k)].copy() for k in self.keys()}
        for q, (label, _) in enumerate(self.windows):
            out["energy_" + label] = records[:, _capi.STAT_KEYS.index(f"window{q}")].copy()
        if self.gas is not None:
            out["pressure"] = np.array([pd.pressure(self.gas, r) for r in out["density"]])
        if callable(self.zdw):
            out["zdw"] = np.array([self.zdw(z) for z in out["z"]])
        elif self.zdw is not None:
            out["zdw"] = np.full(len(records), float(self.zdw))
        out.update(_stack(host))
        return out

    def __call__(self, Ew, z, dz):
        ctx = self.transform.context()
        ctx.set_step_stats(self)
        rec = ctx.stats_eval(Ew, z, dz)
        d = {k: v[0] for k, v in self.to_dict(rec[None, :]).items()}
        d.update(self.host_step(np.asarray(Ew), z, dz))
        return d


def default(grid, Ew, mode, linop, transform, *, windows=None, gas=None, onaxis=False, userfuns=(),
            capacity=16384):
    """``Stats.default(grid, Eω, mode, linop, transform; gas, windows, onaxis,
    userfuns)``, ``src/Stats.jl:492-532``, for a mode-averaged transform."""
    aeff = None
    if onaxis:
        if mode is None or not hasattr(mode, "aeff_onaxis"):
            raise NotImplementedError("on-axis statistics need a capillary (Marcatili) mode")
        if getattr(transform, "gradient", None) is not None:
            raise NotImplementedError("on-axis statistics with a gradient fill are not provided")
        aeff = mode.aeff_onaxis()
    zdw = None
    if mode is not None and hasattr(mode, "beta"):
        from .modes import MarcatiliMode, zdw as mode_zdw
        spec = getattr(transform, "gradient", None)
        if spec is not None and isinstance(mode, MarcatiliMode):
            g = spec.gradient  # the mode at the fill pressure (and taper radius) of each z
            zdw = lambda z: mode_zdw(MarcatiliMode(spec.radius(z) or mode.a, g.gas, g.pressure(z), n=mode.n,
                                                   m=mode.m, kind=mode.kind, T=g.T, model=mode.model,
                                                   loss=mode.loss))
        else:
            zdw = mode_zdw(mode)
    return StatsFun(grid, transform, capacity=capacity, gas=gas, zdw=zdw, windows=windows, aeff=aeff,
                    userfuns=userfuns)


# --------------------------------------------------------------------------
# Statistics evaluated on the host at every accepted step
# --------------------------------------------------------------------------
class StepRecords:
    """The statistics one output call receives: the device records
    [n][LH_NSTAT] of the steps since the previous call and one dict of host
    values per step (every-step streaming: one step per call)."""

    def __init__(self, device=None, host=()):
        self.device = (np.zeros((0, _capi.LH_NSTAT)) if device is None
                       else np.asarray(device, dtype=float).reshape(-1, _capi.LH_NSTAT))
        self.host = list(host)

    def __len__(self):
        return max(len(self.device), len(self.host))


def _stack(host):
    """Per-step dicts → {name: values over steps (first axis)}."""
    if not host:
        return {}
    return {k: np.array([h[k] for h in host]) for k in host[0]}


def analytic(grid, Ew):
    """``plan_analytic(grid, Eω)``'s transform (``src/Stats.jl:423-470``): the
    analytic time-domain field.  Envelope grid: the inverse FFT of Eω along its
    first axis.  Field grid: the Nω rFFT samples doubled into the first Nω - 1
    bins of a 2(Nω - 1)-point buffer (the +fs/2 sample dropped,
    ``copyto_fft!``), then the inverse FFT."""
    Ew = np.asarray(Ew)
    if grid.kind != "real":
        return np.fft.ifft(Ew, axis=0)
    n = Ew.shape[0] - 1
    buf = np.zeros((2 * n,) + Ew.shape[1:], dtype=np.complex128)
    buf[:n] = 2 * Ew[:n]
    return np.fft.ifft(buf, axis=0)


def zdz(d, Ew, Et, z, dz):
    """``zdz!`` (``src/Stats.jl:411-414``)."""
    d["z"] = z
    d["dz"] = dz


def omega0(grid):
    """``ω0(grid)`` (``src/Stats.jl:18-29``): the first moment of |Eω|² over ω,
    one value per column of a 2-D field."""
    w = np.asarray(grid.omega, dtype=float)

    def addstat(d, Ew, Et, z, dz):
        P = np.abs(Ew) ** 2
        m = np.tensordot(w, P, axes=(0, 0)) / np.sum(P, axis=0)
        d["ω0"] = float(m) if np.ndim(m) == 0 else m
    return addstat


def energy(grid, energyfun_w):
    """``energy(grid, energyfun_ω)`` (``src/Stats.jl:35-44``)."""
    def addstat(d, Ew, Et, z, dz):
        if np.ndim(Ew) > 1:
            d["energy"] = np.array([energyfun_w(Ew[:, i]) for i in range(Ew.shape[1])])
        else:
            d["energy"] = energyfun_w(Ew)
    return addstat


def energy_window(grid, energyfun_w, window, label):
    """``energy_window(grid, energyfun_ω, window; label)`` (``src/Stats.jl:69-79``)."""
    key = "energy_" + label
    window = np.asarray(window)

    def addstat(d, Ew, Et, z, dz):
        if np.ndim(Ew) > 1:
            d[key] = np.array([energyfun_w(Ew[:, i] * window) for i in range(Ew.shape[1])])
        else:
            d[key] = energyfun_w(Ew * window)
    return addstat


def energy_lambda(grid, energyfun_w, lambda_lims, label=None, winwidth=0.0):
    """``energy_λ(grid, energyfun_ω, λlims; label, winwidth)`` (``src/Stats.jl:52-61``)."""
    lab, win = energy_window_spec(grid, lambda_lims, winwidth)
    return energy_window(grid, energyfun_w, win, lab if label is None else label)


def peakpower(grid):
    """``peakpower(grid)`` (``src/Stats.jl:86-99``): max |Et|² (per column of a
    2-D field, and ``peakpower_allmodes``, the maximum over t of the sum over
    columns)."""
    def addstat(d, Ew, Et, z, dz):
        P = np.abs(Et) ** 2
        if np.ndim(Et) > 1:
            d["peakpower"] = np.max(P, axis=0)
            d["peakpower_allmodes"] = float(np.max(np.sum(P, axis=1)))
        else:
            d["peakpower"] = float(np.max(P))
    return addstat


def fwhm_t(grid):
    """``fwhm_t(grid)`` (``src/Stats.jl:182-199``): FWHM of |Et|² at its
    narrowest (``min``) and widest (``max``) crossings, per column of a 2-D
    field (and of the column sum, ``_allmodes``)."""
    t = np.asarray(grid.t, dtype=float)

    def addstat(d, Ew, Et, z, dz):
        P = np.abs(Et) ** 2
        if np.ndim(Et) > 1:
            Ps = np.sum(P, axis=1)
            d["fwhm_t_min"] = np.array([_fwhm(t, P[:, i]) for i in range(P.shape[1])])
            d["fwhm_t_max"] = np.array([_fwhm(t, P[:, i], "max") for i in range(P.shape[1])])
            d["fwhm_t_min_allmodes"] = _fwhm(t, Ps)
            d["fwhm_t_max_allmodes"] = _fwhm(t, Ps, "max")
        else:
            d["fwhm_t_min"] = _fwhm(t, P)
            d["fwhm_t_max"] = _fwhm(t, P, "max")
    return addstat


def density(dfun):
    """``density(dfun)`` (``src/Stats.jl:323-327``)."""
    def addstat(d, Ew, Et, z, dz):
        d["density"] = dfun(z)
    return addstat


def pressure(dfun, gas):
    """``pressure(dfun, gas)`` (``src/Stats.jl:335-339``)."""
    def addstat(d, Ew, Et, z, dz):
        d["pressure"] = pd.pressure(gas, dfun(z))
    return addstat


class HostStats:
    """``collect_stats(grid, Eω, funcs...)`` (``src/Stats.jl:476-493``): every
    function f(d, Eω, Et, z, dz) at every accepted step, on the host, with
    ``zdz`` added when missing."""

    has_host = True

    def __init__(self, grid, Ew, funcs):
        self.grid = grid
        self.funcs = list(funcs)
        if zdz not in self.funcs:
            self.funcs.append(zdz)
        self._keys = []

    def host_step(self, Ew, z, dz):
        d = {}
        Et = analytic(self.grid, Ew)
        for f in self.funcs:
            f(d, Ew, Et, z, dz)
        if not self._keys:
            self._keys = list(d)
        return d

    def __call__(self, Ew, z, dz):
        return self.host_step(np.asarray(Ew), z, dz)

    def names(self):
        return list(self._keys)

    def to_dict(self, records):
        return _stack(records.host if isinstance(records, StepRecords) else ())


def collect_stats(grid, Ew, *funcs):
    """``Stats.collect_stats(grid, Eω, funcs...)``."""
    return HostStats(grid, Ew, funcs)
