"""Closed-form checks of the oracle's TransModeAvg chain.

These pin the scaling chain of src/NonlinearRHS.jl:16-53, 362-387 (to_time!
scale (No-1)/(N-1), FFTW 1/No normalisation, field divisor, to_freq! crop
scale) and the PlasmaCumtrapz recursion of src/Nonlinear.jl:121-133 against
results fixed by the mathematics, not by re-typing the oracle.
"""
import math

import numpy as np
import pytest

EPS0 = 8.854187817620389e-12
E_RATIO = 1.6021766208e-19 ** 2 / 9.10938356e-31


def _plain_desc(capi, n_to=4096, kerr=True, plasma=False, ppt=None, gamma3=1e-40, rho=1e25,
                field_div=1.0, ionpot=0.0):
    n_t = n_to // 2
    N = n_t // 2 + 1
    f = dict(grid_kind=capi.LH_GRID_REAL, transform=capi.LH_TRANS_MODEAVG, n_omega=N, n_t=n_t, n_to=n_to,
             n_r=1, dt_o=1e-16, omega=np.arange(N, dtype=float), sidx=np.ones(N, np.uint8),
             omega_win=np.ones(N), t_win=np.ones(n_t), to_win=np.ones(n_to),
             linop=np.zeros(N, complex), norm=np.ones(N, complex), field_div=field_div, density=rho,
             kerr=capi.LH_KERR_FIELD if kerr else 0, gamma3=gamma3, plasma=1 if plasma else 0,
             ionpot=ionpot, preionfrac=0.0)
    if ppt is not None:
        f.update(ppt_n=len(ppt[0]), ppt_x=ppt[0], ppt_y=ppt[1], ppt_d=ppt[2])
    return capi.DescHolder(**f)


def test_kerr_single_line(orc, capi):
    # E(t) = B cos θ  ⇒  E³ = B³(¾ cos θ + ¼ cos 3θ): two spectral lines.
    n_to, k0, a, fd = 4096, 37, 3.0e3, 2.5
    h = _plain_desc(capi, n_to=n_to, field_div=fd)
    N = h.desc.n_omega
    Ew = np.zeros(N, complex)
    Ew[k0] = a
    st, nl = orc.rhs(h, Ew)
    assert st == 0
    No = n_to // 2 + 1
    scale_t = (No - 1) / (N - 1)
    B = 2 * scale_t * a / n_to / fd
    fac = 1e25 * EPS0 * 1e-40
    line1 = (N - 1) / (No - 1) * fac * B ** 3 * 0.75 * n_to / 2
    line3 = (N - 1) / (No - 1) * fac * B ** 3 * 0.25 * n_to / 2
    assert nl[k0] == pytest.approx(line1, rel=1e-12)
    assert nl[3 * k0] == pytest.approx(line3, rel=1e-12)
    other = np.delete(nl, [k0, 3 * k0])
    assert np.max(np.abs(other)) < 1e-12 * abs(line1)


def _const_rate_table(r0, emin=1.0, emax=1e12, n=64):
    x = np.linspace(emin, emax, n)
    return x, np.full(n, math.log(r0)), np.zeros(n)


def test_plasma_constant_rate_closed_form(orc, capi):
    # Constant field E0 and constant rate r0: F = r0 t exactly (trapezoid),
    # f = 1 - e^{-r0 t}, J = ∫ f e²/m E0 + Ip r0 (1-f)/E0, P = ∫ J.
    r0, E0, ip, dt = 2e13, 5e9, 3e-18, 1e-16
    n = 4096
    ppt = _const_rate_table(r0)
    h = _plain_desc(capi, n_to=n, kerr=False, plasma=True, ppt=ppt, rho=1.0, ionpot=ip)
    h.desc.dt_o = dt
    E = np.full(n, E0)
    st, P = orc.responses_real(h, E)
    assert st == 0
    t = np.arange(n) * dt
    f = 1 - np.exp(-r0 * t)
    # exact integrals of the continuous expressions; trapezoid error is O(dt² r0²)
    J = E_RATIO * E0 * (t - (1 - np.exp(-r0 * t)) / r0) + ip * r0 * np.exp(-r0 * t) / E0
    Pex = E_RATIO * E0 * (t ** 2 / 2 - t / r0 + (1 - np.exp(-r0 * t)) / r0 ** 2) \
        + ip / E0 * (1 - np.exp(-r0 * t)) - ip * r0 / E0 * 0.0
    # the loss term enters J at every sample (J[i] += ...), P integrates it from t=0
    Pex = Pex + 0.0
    rel = np.max(np.abs(P[1:] - Pex[1:])) / np.max(np.abs(Pex))
    assert rel < 1e-4, rel
    assert np.all(np.diff(f) >= 0)


def test_plasma_zero_field(orc, capi):
    ppt = _const_rate_table(1e10, emin=1e6)
    h = _plain_desc(capi, n_to=4096, kerr=False, plasma=True, ppt=ppt, rho=1.0, ionpot=1e-18)
    st, P = orc.responses_real(h, np.zeros(4096))
    assert st == 0 and np.all(P == 0.0)
