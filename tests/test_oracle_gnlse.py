"""The oracle's envelope (GNLSE) path against the reference's golden values.

test/test_gnlse.jl:33-56: soliton self-frequency shift of a fundamental soliton
with the SDO Raman response — peak frequency 1.3975031072069625e15 rad/s and
peak time 6.079566076444579e-12 s of the field saved at z = flength (rtol 1e-14,
i.e. the same grid bins; "checked with two independent GNLSE codes").
test/test_gnlse.jl:8-31: an N = 4 soliton over two periods restores its
spectrum (rtol 1e-2) and temporal intensity (rtol 1e-3).
"""
import math

import numpy as np


def _isapprox(a, b, rtol):
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


def test_soliton_self_frequency_shift_golden(orc, capi, pkg, refvals):
    from luna_jl_amd import interface
    gamma, beta2, N, tau0, fr = 0.1, -1e-26, 1.0, 10e-15, 0.18
    P0 = N ** 2 * abs(beta2) / ((1 - fr) * gamma * tau0 ** 2)
    flength = math.pi / 2 * tau0 ** 2 / abs(beta2) * 90
    Ew, grid, linop, tr, _, _ = interface.prop_gnlse_args(
        gamma, flength, [0.0, 0.0, beta2], lambda0=835e-9, tau_fwhm=1.763 * tau0, power=P0,
        pulseshape="sech", lambda_lims=(450e-9, 8000e-9), trange=13e-12, raman=True, shock=False,
        fr=fr, shotnoise=False, ramanmodel="sdo", tau1=12.2e-15, tau2=32e-15)
    h = tr.descriptor(linop)
    st, Ew_end, sv, zs, stats = orc.solve(h, Ew, capi.rk_opts(flength, save_n=2))
    assert st == 0 and zs[-1] == flength
    Eout = sv[-1]
    assert grid.omega[np.argmax(np.abs(Eout) ** 2)] == refvals["ssfs_peak_omega"]
    Et = np.fft.ifft(Eout)
    assert grid.t[np.argmax(np.abs(Et) ** 2)] == refvals["ssfs_peak_t"]


def test_n4_soliton(orc, capi, pkg):
    from luna_jl_amd import interface
    gamma, beta2, N, tau0, fr = 0.1, -1e-26, 4.0, 280e-15, 0.18
    tau_fwhm = 2 * math.log(1 + math.sqrt(2)) * tau0
    P0 = N ** 2 * abs(beta2) / ((1 - fr) * gamma * tau0 ** 2)
    flength = math.pi * tau0 ** 2 / abs(beta2)
    Ew, grid, linop, tr, _, _ = interface.prop_gnlse_args(
        gamma, flength, [0.0, 0.0, beta2], lambda0=835e-9, tau_fwhm=tau_fwhm, power=P0,
        pulseshape="sech", lambda_lims=(450e-9, 8000e-9), trange=4e-12, raman=False, shock=False,
        fr=fr, shotnoise=False)
    h = tr.descriptor(linop)
    st, Ew_end, sv, zs, stats = orc.solve(h, Ew, capi.rk_opts(flength, save_n=201))
    assert st == 0
    assert _isapprox(np.abs(sv[0]) ** 2, np.abs(sv[-1]) ** 2, rtol=1e-2)
    assert _isapprox(np.abs(np.fft.ifft(sv[0])) ** 2, np.abs(np.fft.ifft(sv[-1])) ** 2, rtol=1e-3)
