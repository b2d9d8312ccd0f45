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
mode: once for a constant fill, at each record's z for a gradient fill.  The
``userfuns`` option is not provided.
"""
import math

import numpy as np

from . import _capi
from . import physdata as pd
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


def energy_window(grid, lambda_lims, winwidth=0.0):
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

    def __init__(self, grid, transform, capacity=16384, gas=None, zdw=None, windows=None, aeff=None):
        if getattr(transform, "kind", None) != _capi.LH_TRANS_MODEAVG:
            raise NotImplementedError("per-step statistics are provided for mode-averaged transforms")
        self.grid = grid
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
            self.windows.append(energy_window(grid, lims, ww))
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

    def to_dict(self, records):
        """Records [n][LH_NSTAT] → {name: array over steps} (plus ``pressure``
        from the recorded density when a gas is given)."""
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
        return out

    def __call__(self, Ew, z, dz):
        ctx = self.transform.context()
        ctx.set_step_stats(self)
        rec = ctx.stats_eval(Ew, z, dz)
        return {k: v[0] for k, v in self.to_dict(rec[None, :]).items()}


def default(grid, Ew, mode, linop, transform, *, windows=None, gas=None, onaxis=False, userfuns=(),
            capacity=16384):
    """``Stats.default(grid, Eω, mode, linop, transform; gas, windows, onaxis,
    userfuns)``, ``src/Stats.jl:492-532``, for a mode-averaged transform."""
    if userfuns:
        raise NotImplementedError("userfuns statistics are not provided")
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
    return StatsFun(grid, transform, capacity=capacity, gas=gas, zdw=zdw, windows=windows, aeff=aeff)
