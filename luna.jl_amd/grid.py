"""Time/frequency grids, restating ``src/Grid.jl``.

``RealGrid`` (field-resolved, ``Grid.jl:38-81``) and ``EnvGrid`` (envelope,
``Grid.jl:115-190``).  Arrays use the reference's 0-based equivalents of its
1-based layouts: ``omega`` is the coarse state grid (rfft order for RealGrid, FFT
order for EnvGrid), ``to``/``omega_o`` the oversampled grid the RHS runs on.
"""
import math
import numpy as np

from . import physdata as pd
from .maths import planck_taper


def _next_pow2_exp(x):
    return int(math.ceil(math.log2(x)))


class RealGrid:
    """``Grid.RealGrid(zmax, referenceλ, λ_lims, trange, δt=1)``, ``Grid.jl:38-81``."""

    kind = "real"

    def __init__(self, zmax, ref_lambda, lambda_lims, trange, dt=1.0):
        f_lims = (pd.c / lambda_lims[0], pd.c / lambda_lims[1])
        dto = min(1 / (6 * max(f_lims)), dt)
        samples = 2 ** _next_pow2_exp(trange / dto)
        trange_even = dto * samples
        dwo = 2 * math.pi / trange_even
        n_to = np.arange(samples, dtype=float)
        self.to = (n_to - samples / 2) * dto
        self.omega_o = np.arange(samples // 2 + 1, dtype=float) * dwo
        wmin = 2 * math.pi * min(f_lims)
        wmax = 2 * math.pi * max(f_lims)
        wmax_win = 1.1 * wmax
        cropidx = int(np.argmax(self.omega_o > wmax_win)) + 1  # 1-based findfirst
        cropidx = 2 ** _next_pow2_exp(cropidx) + 1
        self.omega = self.omega_o[:cropidx].copy()
        dt_coarse = math.pi / self.omega.max()
        tsamples = (cropidx - 1) * 2
        self.t = (np.arange(tsamples, dtype=float) - tsamples / 2) * dt_coarse
        self.omega_win = planck_taper(self.omega, wmin / 2, wmin, wmax, wmax_win)
        self.t_win = planck_taper(self.t, self.t.min(), -trange / 2, trange / 2, self.t.max())
        self.to_win = planck_taper(self.to, self.to.min(), -trange / 2, trange / 2, self.to.max())
        self.sidx = (self.omega > wmin / 2) & (self.omega < wmax_win)
        self.zmax = float(zmax)
        self.ref_lambda = ref_lambda
        self.omega0 = pd.wlfreq(ref_lambda)
        self.trange = trange
        dtt = self.t[1] - self.t[0]
        dtto = self.to[1] - self.to[0]
        assert math.isclose(dtt / dtto, len(self.to) / len(self.t), rel_tol=1e-8)

    @property
    def n_omega(self):
        return len(self.omega)

    @property
    def n_t(self):
        return len(self.t)

    @property
    def n_to(self):
        return len(self.to)

    @property
    def dt_o(self):
        return self.to[1] - self.to[0]

    def to_dict(self):
        """``Grid.to_dict`` (``Grid.jl:238-244``): the struct's fields by name."""
        return {"zmax": self.zmax, "referenceλ": self.ref_lambda, "t": self.t, "ω": self.omega,
                "to": self.to, "ωo": self.omega_o, "sidx": self.sidx, "ωwin": self.omega_win,
                "twin": self.t_win, "towin": self.to_win}


class EnvGrid:
    """``Grid.EnvGrid(zmax, referenceλ, λ_lims, trange; δt=1, thg=false)``, ``Grid.jl:115-190``."""

    kind = "env"

    def __init__(self, zmax, ref_lambda, lambda_lims, trange, dt=1.0, thg=False):
        fmin = pd.c / max(lambda_lims)
        fmax = pd.c / min(lambda_lims)
        fmax_win = 1.1 * fmax
        omega0 = 2 * math.pi * pd.c / ref_lambda
        if thg:
            ffac = 6
            f_lims = (fmin, fmax)
        else:
            ffac = 2
            f_lims = (fmin, fmax_win)
        rf = [f - pd.c / ref_lambda for f in f_lims]
        dt_f = 1 / (ffac * max(rf))
        if dt_f <= dt:
            dto = dt_f
            oversampling = thg
        else:
            dto = dt
            oversampling = True
        samples = 2 ** _next_pow2_exp(trange / dto)
        trange_even = dto * samples
        dwo = 2 * math.pi / trange_even
        no = np.arange(samples, dtype=float)
        self.to = (no - samples / 2) * dto
        vo = (no - samples / 2) * dwo
        vo = np.fft.fftshift(vo)
        self.omega_o = vo + omega0
        wmin = 2 * math.pi * fmin
        wmax = 2 * math.pi * fmax
        wmax_win = 2 * math.pi * fmax_win
        if oversampling:
            cropidx = int(np.argmax(self.omega_o >= (wmax_win - dwo))) + 1
            cropidx = 2 ** _next_pow2_exp(cropidx)
            v = np.concatenate([vo[:cropidx], vo[-cropidx:]])
        else:
            v = vo
        dt_coarse = -math.pi / v.min()
        tsamples = len(v)
        self.t = (np.arange(tsamples, dtype=float) - tsamples / 2) * dt_coarse
        self.omega = v + omega0
        self.sidx = (self.omega > wmin / 2) & (self.omega < wmax_win)
        self.omega_win = planck_taper(self.omega, wmin / 2, wmin, wmax, wmax_win)
        self.t_win = planck_taper(self.t, self.t.min(), -trange / 2, trange / 2, self.t.max())
        self.to_win = planck_taper(self.to, self.to.min(), -trange / 2, trange / 2, self.to.max())
        self.zmax = float(zmax)
        self.ref_lambda = ref_lambda
        self.omega0 = omega0
        self.trange = trange
        self.oversampling = oversampling

    n_omega = RealGrid.n_omega
    n_t = RealGrid.n_t
    n_to = RealGrid.n_to
    dt_o = RealGrid.dt_o

    def to_dict(self):
        """``Grid.to_dict`` (``Grid.jl:238-244``): the struct's fields by name."""
        return {"zmax": self.zmax, "referenceλ": self.ref_lambda, "ω0": self.omega0, "t": self.t,
                "ω": self.omega, "to": self.to, "ωo": self.omega_o, "sidx": self.sidx,
                "ωwin": self.omega_win, "twin": self.t_win, "towin": self.to_win}
