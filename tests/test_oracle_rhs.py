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


def _plasma_closed_form(t, r0, E0, ip):
    """P(t) of PlasmaCumtrapz for a constant field E0 and a constant rate r0
    (preionfrac 0): F = r0 t, f = 1 - e^{-r0 t}, J = ∫ f e²/m E0 dt + Ip r0 (1-f)/E0,
    P = ∫ J dt, integrated exactly.  Returns (phase part, loss part)."""
    phase = E_RATIO * E0 * (t ** 2 / 2 - t / r0 + (1 - np.exp(-r0 * t)) / r0 ** 2)
    loss = ip / E0 * (1 - np.exp(-r0 * t))
    return phase, loss


def _plasma_case(orc, capi, r0, E0, ip, dt, n=4096):
    ppt = _const_rate_table(r0)
    h = _plain_desc(capi, n_to=n, kerr=False, plasma=True, ppt=ppt, rho=1.0, ionpot=ip)
    h.desc.dt_o = dt
    st, P = orc.responses_real(h, np.full(n, E0))
    assert st == 0
    return np.arange(n) * dt, P


def test_plasma_constant_rate_phase_term(orc, capi):
    # no ionisation loss: P is the free-electron current integrated twice.
    # The trapezoid error of the two outer integrals is O((r0 dt)²) ≈ 4e-6.
    t, P = _plasma_case(orc, capi, r0=2e13, E0=5e9, ip=0.0, dt=1e-16)
    phase, loss = _plasma_closed_form(t, 2e13, 5e9, 0.0)
    rel = np.max(np.abs(P[1:] - phase[1:])) / np.max(np.abs(phase))
    assert rel < 2e-5, rel


def test_plasma_constant_rate_loss_term(orc, capi):
    # A weak field (1e7 V/m) and a large Ip: the ionisation-loss current
    # Ip·rate·(1-f)/E (src/Nonlinear.jl:128-131) dominates P by ~600x, so
    # dropping it, doubling it or dividing by E² instead of E misses by far
    # more than the tolerance ((r0 dt)² = 1e-6 trapezoid error).
    r0, E0, ip, dt = 1e14, 1e7, 3.9e-18, 1e-17
    t, P = _plasma_case(orc, capi, r0=r0, E0=E0, ip=ip, dt=dt)
    phase, loss = _plasma_closed_form(t, r0, E0, ip)
    assert loss[-1] > 300 * abs(phase[-1])
    rel = np.max(np.abs(P[1:] - (phase + loss)[1:])) / np.max(np.abs(phase + loss))
    assert rel < 1e-5, rel
    for wrong in (phase, phase + 2 * loss, phase + loss / E0):
        assert np.max(np.abs(P[1:] - wrong[1:])) / np.max(np.abs(phase + loss)) > 0.1


def test_plasma_zero_field(orc, capi):
    ppt = _const_rate_table(1e10, emin=1e6)
    h = _plain_desc(capi, n_to=4096, kerr=False, plasma=True, ppt=ppt, rho=1.0, ionpot=1e-18)
    st, P = orc.responses_real(h, np.zeros(4096))
    assert st == 0 and np.all(P == 0.0)
