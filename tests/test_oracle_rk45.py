"""The oracle's PreconStepper against the reference's own RK45 test.

test/test_rk45.jl:1-56, 118-135: an N = 5 NLS soliton over two soliton periods
with the interaction-picture stepper restores its initial spectrum (rtol 1e-3
at rtol = 1e-8), and a z-dependent-but-constant linear operator gives the same
answer as the constant array (here: the same linop passed twice).
"""
import numpy as np


def nls_setup():
    trange, wrange = 32, 160
    samples = 2 ** int(np.ceil(np.log2(trange * wrange / (2 * np.pi))))
    dt = trange / samples
    dw = 2 * np.pi / trange
    n = np.arange(samples)
    t = (n - samples / 2) * dt
    w = (n - samples / 2) * dw
    N = 5
    At = N / np.cosh(t)
    Aw = np.fft.fftshift(np.fft.fft(At))
    Lin = -1j / 2 * w ** 2
    zmax = np.pi / 2 * 2

    def fnl(Aw_, z):
        out = np.fft.ifft(Aw_)
        out = np.abs(out) ** 2 * out
        out = np.fft.fft(out)
        return 1j * out

    return t, w, zmax, Aw, fnl, Lin


def isapprox(a, b, rtol):
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


def test_soliton_restoration_precon(orc):
    t, w, zmax, Aw, fnl, Lin = nls_setup()
    st, tout, yout, steps = orc.solve_precon(fnl, Lin, Aw, 0.0, 1e-3, zmax, rtol=1e-8, outputN=501)
    assert st == 0
    assert tout[0] == 0.0 and tout[-1] == zmax
    assert isapprox(np.abs(yout[0]) ** 2, np.abs(yout[-1]) ** 2, rtol=1e-3)
    # one soliton period in: the N=5 breather is far from its input shape
    assert not isapprox(np.abs(yout[125]) ** 2, np.abs(yout[0]) ** 2, rtol=1e-1)
    # repeatability (the reference compares const vs function linop bitwise)
    st2, _, yout2, steps2 = orc.solve_precon(fnl, Lin.copy(), Aw, 0.0, 1e-3, zmax, rtol=1e-8, outputN=501)
    assert steps2 == steps and np.array_equal(np.abs(yout) ** 2, np.abs(yout2) ** 2)


def test_linear_propagation_exact(orc):
    # with the nonlinearity off the interaction picture is exact: y(z) = exp(L z) y0
    t, w, zmax, Aw, fnl, Lin = nls_setup()
    st, tout, yout, steps = orc.solve_precon(lambda y, z: np.zeros_like(y), Lin, Aw, 0.0, 1e-3, zmax,
                                             rtol=1e-8, outputN=11)
    assert st == 0
    for k in range(11):
        ref = Aw * np.exp(Lin * tout[k])
        assert np.max(np.abs(yout[k] - ref)) <= 1e-12 * np.max(np.abs(Aw))
