"""The oracle's FFTs against an independent FFT (numpy's pocketfft).

Pins FFTW's conventions used by the reference: unnormalised r2c/c2r/c2c
(src/NonlinearRHS.jl:16-53) and c2r ignoring the imaginary parts of the DC and
Nyquist bins.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [4096, 8192, 16384])
def test_rfft_matches_pocketfft(orc, n):
    x = np.random.default_rng(n).standard_normal(n)
    ref = np.fft.rfft(x)
    got = orc.rfft(x)
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-14


@pytest.mark.parametrize("n", [4096, 16384])
def test_irfft_ignores_dc_nyquist_imag(orc, n):
    rng = np.random.default_rng(1)
    X = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
    ref = np.fft.irfft(X, n) * n  # pocketfft also drops imag(X[0]), imag(X[n/2])
    got = orc.irfft_unnorm(X, n)
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-14
    X2 = X.copy()
    X2[0] = X2[0].real
    X2[-1] = X2[-1].real
    assert np.array_equal(orc.irfft_unnorm(X2, n), got)


@pytest.mark.parametrize("n", [8192, 32768])
def test_c2c_both_directions(orc, n):
    rng = np.random.default_rng(2)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = np.fft.fft(z)
    assert np.max(np.abs(orc.fft(z, -1) - f)) / np.max(np.abs(f)) < 1e-14
    assert np.max(np.abs(orc.fft(z, +1) - np.fft.ifft(z) * n)) / np.max(np.abs(f)) < 1e-14


def test_gaussian_closed_form(orc):
    # FT of a Gaussian is a Gaussian: exp(-t²/2) ↔ √(2π) exp(-ω²/2)
    n, dt = 4096, 0.01
    t = (np.arange(n) - n // 2) * dt
    x = np.exp(-t ** 2 / 2)
    X = orc.rfft(x) * dt * np.exp(1j * np.fft.rfftfreq(n, dt) * 2 * np.pi * (n // 2) * dt)
    w = 2 * np.pi * np.fft.rfftfreq(n, dt)
    ref = np.sqrt(2 * np.pi) * np.exp(-w ** 2 / 2)
    assert np.max(np.abs(X - ref)) < 1e-12
