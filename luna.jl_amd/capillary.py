"""Gas-filled capillary with a pressure gradient (host-side set-up).

Restates ``Capillary.gradient`` (``src/Capillary.jl:306-340``): the pressure
profile p(z) of a two-point fill (``gradient(gas, L, p0, p1)``) or of a
multi-point fill (``gradient(gas, Z, P)``), the density ρ(z) =
``densityspline``(p(z)) (``src/PhysData.jl:719-723``) and the core index
√(1 + γ(λ)ρ(z)).  ``gradient_spec`` packs what the propagation kernels need to
re-evaluate the linear operator of ``LinearOps.make_linop``
(``src/LinearOps.jl:337-358``) and the normalisation of ``norm_mode_average``
(``src/NonlinearRHS.jl:375-387``) at every new z: γ(λ) and the waveguide
index ``neff_wg`` per frequency bin (neither depends on z for a constant core
radius, as in ``neff_β_grid``, ``src/Capillary.jl:345-360``), the density
spline, and the β1 stencil around ω0.
"""
import math

import numpy as np

from . import _capi
from . import physdata as pd
from .maths import cspline_eval
from .modes import _FD8


class Gradient:
    """``Capillary.gradient(gas, L, p0, p1)`` (``two_point``) or ``gradient(gas, Z, P)``."""

    def __init__(self, gas, Z, P, T=pd.roomtemp, two_point=False):
        self.gas, self.T = gas, T
        self.Z = np.ascontiguousarray(Z, dtype=float)
        self.P = np.ascontiguousarray(P, dtype=float)
        self.two_point = bool(two_point)
        if len(self.Z) != len(self.P) or len(self.Z) < 2 or np.any(np.diff(self.Z) <= 0):
            raise ValueError("Z must increase and match P")
        if self.two_point and (len(self.Z) != 2 or self.Z[0] != 0.0):
            raise ValueError("a two-point gradient runs from z = 0 to L")
        lo, hi = float(self.P.min()), float(self.P.max())
        self.dens_p, self.dens_rho, self.dens_d = pd.density_spline(gas, hi, 0.0 if lo == hi else lo, T=T)

    @classmethod
    def linear(cls, gas, L, p0, p1, T=pd.roomtemp):
        return cls(gas, [0.0, L], [p0, p1], T=T, two_point=True)

    def pressure(self, z):
        Z, P = self.Z, self.P
        if self.two_point:
            L, p0, p1 = Z[1], P[0], P[1]
            return p1 if z > L else p0 if z <= 0 else math.sqrt(p0 * p0 + z / L * (p1 * p1 - p0 * p0))
        if z <= Z[0]:
            return float(P[0])
        if z >= Z[-1]:
            return float(P[-1])
        i = int(np.nonzero(Z < z)[0][-1])
        return math.sqrt(P[i] * P[i] + (z - Z[i]) / (Z[i + 1] - Z[i]) * (P[i + 1] * P[i + 1] - P[i] * P[i]))

    def density(self, z):
        return float(cspline_eval(self.dens_p, self.dens_rho, self.dens_d, self.pressure(z)))


def gradient_spec(grid, mode, gradient, ref_lambda, shock=True, taper=None):
    """The arrays of ``lh_gradient_desc`` for a Marcatili mode (``mode`` supplies
    the radius, cladding, model and loss switch; its own fill pressure is not
    used).  ``taper = (a0, a1, L)``: the radius runs linearly from a0 at z = 0
    to a1 at z = L (constant beyond); ``taper = (Z, A)`` (two sequences): the
    piecewise-linear radius through those nodes, constant outside.  With a
    taper the mode's radius is ignored."""
    if grid.kind != "real":
        raise NotImplementedError("gradient fills are provided on the field grid")
    w = grid.omega
    s = grid.sidx
    gamma_fun = pd.sellmeier_gas(gradient.gas)
    gamma = np.zeros(len(w))
    gamma[s] = gamma_fun(pd.wlfreq(w[s]) * 1e6)
    nwg = np.zeros(len(w), dtype=np.complex128)
    nwg[s] = mode.neff_wg(w[s])
    shockterm = w * w if shock else w * pd.wlfreq(grid.ref_lambda)
    from .nonlinearrhs import NLSCALE
    pre = np.zeros(len(w), dtype=np.complex128)
    pre[s] = 1j * (-(((shockterm[s]) / 4) / NLSCALE) / pd.c)
    w0 = pd.wlfreq(ref_lambda)
    h = abs(w0) * 1e-3
    wj = w0 + h * np.arange(-4, 5)
    fd = np.zeros((9, 4))
    fd[:, 0] = wj
    fd[:, 1] = gamma_fun(pd.wlfreq(wj) * 1e6)
    nw = mode.neff_wg(wj)
    fd[:, 2], fd[:, 3] = nw.real, nw.imag
    tp = None
    if taper is not None:
        # neff_wg(ω, a) from U = u/k and V = v_n/k (src/Capillary.jl:188-214)
        U = np.zeros(len(w))
        V = np.zeros(len(w), dtype=np.complex128)
        k = w[s] / pd.c
        U[s] = mode.unm / k
        V[s] = mode.vn(w[s]) / k
        kj = wj / pd.c
        fdt = np.stack([mode.unm / kj, (mode.vn(wj) / kj).real, (mode.vn(wj) / kj).imag], axis=1)
        tp = dict(sqrt_intg=math.sqrt(mode.aeff_intg), U=U, V=V, fdt=fdt)
        if len(taper) == 2:
            tp.update(kind=2, Z=np.ascontiguousarray(taper[0], dtype=float),
                      A=np.ascontiguousarray(taper[1], dtype=float))
        else:
            tp.update(kind=1, a0=float(taper[0]), a1=float(taper[1]), L=float(taper[2]))
    return GradientSpec(gradient, mode, gamma, nwg, pre, math.sqrt(mode.Aeff()), h, fd, taper=tp)


def taper_radius(taper, z):
    """a(z) of a linear taper (a0, a1, L): a0 + (a1 - a0)·clamp(z/L, 0, 1), or
    of a piecewise-linear one (Z, A), constant outside the nodes."""
    if len(taper) == 2:
        Z, A = np.asarray(taper[0], dtype=float), np.asarray(taper[1], dtype=float)
        if z <= Z[0]:
            return float(A[0])
        if z >= Z[-1]:
            return float(A[-1])
        i = int(np.searchsorted(Z, z, side="left")) - 1  # Z[i] < z <= Z[i+1]
        return float(A[i] + (A[i + 1] - A[i]) * ((z - Z[i]) / (Z[i + 1] - Z[i])))
    a0, a1, L = taper
    return a0 + (a1 - a0) * min(max(z / L, 0.0), 1.0)


class GradientSpec:
    """Owns the arrays an ``LhGradientDesc`` points into."""

    def __init__(self, gradient, mode, gamma, nwg, pre, sqrt_aeff, fd_h, fd, taper=None):
        self.gradient = gradient
        self.model = 0 if mode.model == "full" else 1
        self.loss = 1 if mode.loss else 0
        self.gamma = np.ascontiguousarray(gamma, dtype=float)
        self.nwg = np.ascontiguousarray(nwg, dtype=np.complex128)
        self.pre = np.ascontiguousarray(pre, dtype=np.complex128)
        self.sqrt_aeff = float(sqrt_aeff)
        self.fd_h = float(fd_h)
        self.fd = np.ascontiguousarray(fd, dtype=float)
        self.taper = taper
        if taper is not None:
            self.wgU = np.ascontiguousarray(taper["U"], dtype=float)
            self.wgV = np.ascontiguousarray(taper["V"], dtype=np.complex128)
            self.fdt = np.ascontiguousarray(taper["fdt"], dtype=float)
        g = gradient
        self.desc = _capi.LhGradientDesc(
            n_points=len(g.Z), two_point=int(g.two_point), Z=_capi._ptr(g.Z), P=_capi._ptr(g.P),
            dens_n=len(g.dens_p), model=self.model, dens_p=_capi._ptr(g.dens_p),
            dens_rho=_capi._ptr(g.dens_rho), dens_d=_capi._ptr(g.dens_d), loss=self.loss,
            taper=taper["kind"] if taper is not None else 0,
            gamma=_capi._ptr(self.gamma), nwg=_capi._ptr(self.nwg.view(np.float64)),
            pre=_capi._ptr(self.pre.view(np.float64)), sqrt_aeff=self.sqrt_aeff, fd_h=self.fd_h,
            fd=_capi._ptr(self.fd))
        if taper is not None:
            d = self.desc
            if taper["kind"] == 1:
                d.taper_a0, d.taper_a1, d.taper_L = taper["a0"], taper["a1"], taper["L"]
            else:
                self.taper_Z, self.taper_A = taper["Z"], taper["A"]
                d.taper_n = len(self.taper_Z)
                d.taper_Z, d.taper_A = _capi._ptr(self.taper_Z), _capi._ptr(self.taper_A)
            d.sqrt_aeff_intg = taper["sqrt_intg"]
            d.wgU = _capi._ptr(self.wgU)
            d.wgV = _capi._ptr(self.wgV.view(np.float64))
            d.fdt = _capi._ptr(self.fdt)

    def radius(self, z):
        """a(z): the taper's radius, or None for a constant radius."""
        if self.taper is None:
            return None
        if self.taper["kind"] == 2:
            return taper_radius((self.taper["Z"], self.taper["A"]), z)
        return taper_radius((self.taper["a0"], self.taper["a1"], self.taper["L"]), z)

    @property
    def ref(self):
        return _capi.C.byref(self.desc)
