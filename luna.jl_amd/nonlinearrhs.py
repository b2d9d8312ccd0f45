"""Transforms E(ω) → P_nl(ω): the drop-in for ``src/NonlinearRHS.jl``.

``TransModeAvg`` (``NonlinearRHS.jl:323-373``) and ``TransRadial``
(``:406-474``) keep the reference's callable contract ``t(nl, Eω, z)`` (writes
``nl`` in place) but evaluate it in ``libluna_hip.so`` (``lh_rhs``).  The
normalisation builders ``norm_mode_average`` (``:375-387``),
``norm_mode_average_gnlse`` (``:389-399``) and ``const_norm_radial``
(``:482-524``) return the per-bin factor arrays the kernels apply.
"""
import math

import numpy as np

from . import _capi
from . import physdata as pd
from .nonlinear import Kerr_env, Kerr_field, PlasmaCumtrapz, RamanPolarEnv, RamanPolarField

NLSCALE = math.sqrt(pd.eps_0 * pd.c / 2)  # NonlinearRHS.jl:360


def _aeff_value(aeff):
    return float(aeff(0.0)) if callable(aeff) else float(aeff)


def _density_value(densityfun):
    return float(densityfun(0.0)) if callable(densityfun) else float(densityfun)


def norm_mode_average(grid, beta, aeff, shock=True):
    """Complex factor (pre/β)·√Aeff of ``norm_mode_average``, ``NonlinearRHS.jl:375-387``."""
    w = grid.omega
    shockterm = w * w if shock else w * pd.wlfreq(grid.ref_lambda)
    pre_im = -(((shockterm) / 4) / NLSCALE) / pd.c
    s = math.sqrt(_aeff_value(aeff))
    f = np.zeros(len(w), dtype=np.complex128)
    sid = grid.sidx
    f[sid] = (0.0 + 1j * (pre_im[sid] / beta[sid])) * s
    return f


def norm_mode_average_gnlse(grid, aeff, shock=True):
    """Complex factor pre·√Aeff of ``norm_mode_average_gnlse``, ``NonlinearRHS.jl:389-399``."""
    w = grid.omega
    shockterm = w * w if shock else w * pd.wlfreq(grid.ref_lambda)
    with np.errstate(divide="ignore", invalid="ignore"):
        pre_im = (-shockterm / (2 * pd.c ** 1.5 * math.sqrt(2 * pd.eps_0))) / (w / pd.c)
    s = math.sqrt(_aeff_value(aeff))
    f = np.zeros(len(w), dtype=np.complex128)
    sid = grid.sidx
    f[sid] = (1j * pre_im[sid]) * s
    return f


def norm_radial(grid, q, nfun_omega, z=0.0):
    """Real (Nω, Nr) factor of ``norm_radial``, ``NonlinearRHS.jl:497-524``:
    kz/(μ0 ω) where kz² = (n ω/c)² - k_r² > 0, else 1 (also at ω = 0)."""
    w = grid.omega
    k2 = np.zeros(len(w))
    s = grid.sidx
    k2[s] = (np.real(nfun_omega(w[s])) * w[s] / pd.c) ** 2
    out = np.ones((len(w), q.N))
    kr2 = q.k ** 2
    for ir in range(q.N):
        bsq = k2 - kr2[ir]
        m = (w != 0) & (bsq > 0)
        out[m, ir] = np.sqrt(bsq[m]) / (pd.mu_0 * w[m])
    return out


def const_norm_radial(grid, q, nfun_lambda):
    """``const_norm_radial(grid, q, nfun)``, ``NonlinearRHS.jl:482-493``: nfun of
    wavelength, factor evaluated once and returned by a closure of z."""
    out = norm_radial(grid, q, lambda w: nfun_lambda(pd.wlfreq(w)))
    return lambda z: out


class _TransformBase:
    kind = None

    def __init__(self, grid, responses, densityfun):
        self.grid = grid
        self.responses = tuple(responses)
        self.density = _density_value(densityfun)
        self._ctx = None
        self._ctx_linop_id = None

    # -- descriptor ---------------------------------------------------------
    def _response_fields(self):
        f = dict(kerr=_capi.LH_KERR_NONE, plasma=0, gamma3=0.0, ppt_n=0, raman=0,
                 ionpot=0.0, preionfrac=0.0)
        for r in self.responses:
            if isinstance(r, Kerr_field):
                f["kerr"], f["gamma3"] = _capi.LH_KERR_FIELD, r.gamma3
            elif isinstance(r, Kerr_env):
                f["kerr"], f["gamma3"] = _capi.LH_KERR_ENV, r.gamma3
            elif isinstance(r, PlasmaCumtrapz):
                spl = r.ratefunc
                f.update(plasma=1, ppt_n=len(spl.x), ppt_x=spl.x, ppt_y=spl.y, ppt_d=spl.d,
                         ionpot=r.ionpot, preionfrac=r.preionfrac)
            elif isinstance(r, RamanPolarEnv):  # and RamanPolarField, its field-grid form
                if (self.grid.kind == "real") != isinstance(r, RamanPolarField):
                    raise TypeError("RamanPolarField belongs to a field grid, RamanPolarEnv to an envelope grid")
                f.update(raman=1, raman_h=np.ascontiguousarray(r.h(self.density), dtype=float))
            else:
                raise TypeError(f"unsupported response {type(r).__name__}")
        order = [type(r) for r in self.responses]
        # Et_to_Pt! accumulates in tuple order; the kernels apply Kerr first.
        if Kerr_field in order and order.index(Kerr_field) != 0:
            raise ValueError("Kerr response must come first in the response tuple")
        return f

    def descriptor(self, linop=None):
        g = self.grid
        n_r = self.n_r
        if linop is None:
            linop = np.zeros(g.n_omega * n_r, dtype=np.complex128)
        linop = np.asarray(linop, dtype=np.complex128).reshape(-1, order="F")
        fields = dict(
            grid_kind=_capi.LH_GRID_REAL if g.kind == "real" else _capi.LH_GRID_ENV,
            transform=self.kind, n_omega=g.n_omega, n_t=g.n_t, n_to=g.n_to, n_r=n_r,
            dt_o=g.dt_o, omega=g.omega, sidx=g.sidx.astype(np.uint8), omega_win=g.omega_win,
            t_win=g.t_win, to_win=g.to_win, linop=linop, norm=self.norm_array(),
            field_div=self.field_div(), density=self.density)
        fields.update(self._response_fields())
        fields.update(self._extra_fields())
        return _capi.DescHolder(**fields)

    def _extra_fields(self):
        return {}

    # -- device context -------------------------------------------------------
    def context(self, linop=None):
        from .runtime import Context
        key = None if linop is None else id(linop)
        if self._ctx is None or key != self._ctx_linop_id:
            self._ctx = Context(self.descriptor(linop))
            self._ctx_linop_id = key
        return self._ctx

    def __call__(self, nl, Ew, z):
        out = self.context().rhs(Ew, z)
        nl[...] = out.reshape(nl.shape, order="F")

    def __str__(self):
        """``show(io, t::TransModeAvg)`` / ``TransRadial`` (``NonlinearRHS.jl:336-342,
        420-428``), the ``transform`` entry of the output's ``simulation_type``."""
        g = self.grid
        parts = [type(self).__name__, f"grid type: {type(g).__name__}", f"time grid size: {g.n_t} / {g.n_to}"]
        if self.kind == _capi.LH_TRANS_RADIAL:
            parts += [f"radial points: {self.q.N}", f"aperture: {self.q.R}"]
        parts.append("responses: " + "\n    ".join(type(r).__name__ for r in self.responses))
        return "\n  ".join(parts)


class TransModeAvg(_TransformBase):
    """``TransModeAvg(grid, FT, resp, densityfun, norm!, aeff)`` (FT implied)."""

    kind = _capi.LH_TRANS_MODEAVG
    n_r = 1

    def __init__(self, grid, responses, densityfun, norm, aeff, gradient=None):
        """``gradient``: a ``capillary.GradientSpec`` for a z-dependent fill, whose
        density, operator and normalisation the kernels re-evaluate at every z
        (``densityfun`` and ``norm`` then give the z = 0 values)."""
        super().__init__(grid, responses, densityfun)
        self.norm = np.asarray(norm, dtype=np.complex128)
        self.aeff = _aeff_value(aeff)
        self.gradient = gradient

    def descriptor(self, linop=None):
        holder = super().descriptor(linop)
        holder.gradient = self.gradient
        return holder

    def norm_array(self):
        return self.norm

    def field_div(self):
        return NLSCALE * math.sqrt(self.aeff)


class TransRadial(_TransformBase):
    """``TransRadial(grid, HT, FT, responses, densityfun, normfun)``, ``NonlinearRHS.jl:428-456``."""

    kind = _capi.LH_TRANS_RADIAL

    def __init__(self, grid, q, responses, densityfun, normfun):
        super().__init__(grid, responses, densityfun)
        self.q = q
        self.n_r = q.N
        nf = normfun(0.0) if callable(normfun) else normfun  # (Nω, Nr) real
        w = grid.omega[:, None]
        fac = (grid.omega_win[:, None] * (-1j * w)) / (2 * nf)
        self.norm = np.asarray(fac, dtype=np.complex128).reshape(-1, order="F")

    def norm_array(self):
        return self.norm

    def field_div(self):
        return 1.0

    def _extra_fields(self):
        return dict(qdht=np.ascontiguousarray(self.q.T, dtype=float),
                    qdht_fwd=self.q.R / self.q.K, qdht_inv=self.q.K / self.q.R)
