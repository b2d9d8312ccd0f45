"""Pins the CPU restatement of the spectral surrogate (oracle/fno_oracle.py):
the truncated-DFT spectral convolution against numpy's FFT, and the analytic
gradient against central finite differences of the FP64 loss."""
import numpy as np
import pytest

from oracle import fno_oracle  # noqa: E402


def tiny(seed=0):
    o = fno_oracle.FNOOracle(n_x=64, cin=3, channels=6, modes=5, layers=2, hidden=8)
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(o.n_params) * 0.3
    x = rng.standard_normal((2, 64, 3))
    t = rng.standard_normal((2, 64))
    return o, p, x, t


def test_spectral_conv_matches_fft():
    # K(v) = irfft(R·rfft(v)[:modes], n): the oracle's matrices against pocketfft
    o = fno_oracle.FNOOracle(n_x=64, cin=1, channels=4, modes=5, layers=1, hidden=8)
    rng = np.random.default_rng(1)
    v = rng.standard_normal((64, 4))
    Rre = rng.standard_normal((5, 4, 4))
    Rim = rng.standard_normal((5, 4, 4))
    vhat = o.F @ v                                       # rows 2m: Re, 2m+1: Im
    V = np.fft.rfft(v, axis=0)[:5]
    np.testing.assert_allclose(vhat[0::2], V.real, atol=2e-6)
    np.testing.assert_allclose(vhat[1::2], V.imag, atol=2e-6)
    O = np.einsum("mi,mio->mo", V, Rre + 1j * Rim)
    full = np.zeros((33, 4), complex)
    full[:5] = O
    ref = np.fft.irfft(full, n=64, axis=0)
    oh = np.zeros((10, 4))
    oh[0::2], oh[1::2] = O.real, O.imag
    np.testing.assert_allclose(o.Gi @ oh, ref, atol=1e-6 * np.abs(ref).max())
    # the real block form of the complex channel mixing
    Rb = o.rblk(Rre[2], Rim[2])
    vb = np.concatenate([V[2].real, V[2].imag])
    ob = vb @ Rb
    np.testing.assert_allclose(ob[:4] + 1j * ob[4:], V[2] @ (Rre[2] + 1j * Rim[2]), rtol=1e-12)


def test_gradient_matches_finite_differences():
    o, p, x, t = tiny()
    wt = np.random.default_rng(4).uniform(0.2, 2.0, t.shape)
    loss, g, _ = o.loss_grad(p, x, t, wt=wt)
    rng = np.random.default_rng(2)
    names = list(o.lay)
    picks = []
    for name in names:  # a few entries of every tensor
        off, shape = o.lay[name]
        n = int(np.prod(shape))
        picks += list(off + rng.choice(n, size=min(3, n), replace=False))
    h = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (o.loss_grad(pp, x, t, wt=wt)[0] - o.loss_grad(pm, x, t, wt=wt)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_bf16_rounding():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 65504.0])
    r = fno_oracle.bf16_round(a)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7        # 1 + 0.75 ulp rounds up to 1 + 1 ulp
    assert abs(r[3] + 3.140625) < 1e-12
