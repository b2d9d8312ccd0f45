"""Host-side set-up physics against the reference's golden values.

PPT/ADK: test/test_ionisation.jl:6-47.  Capillary loss: test/test_capillary.jl:9-17.
Density (REFPROP values): test/test_physdata.jl:62-75 (the reference uses
CoolProp; the host model here is a virial fit, so a loose tolerance).
Grids: sizes from src/Grid.jl:38-190 and its asserts.
"""
import math

import numpy as np
import pytest


def test_adk_golden(pkg):
    from luna_jl_amd import ionisation as ion
    assert ion.ionrate_adk("He", 1e10) == pytest.approx(1.2416371415312408e-18, rel=1e-6)
    assert ion.ionrate_adk("He", 2e10) == pytest.approx(1.0772390893742478, rel=1e-6)
    assert ion.ionrate_adk("Ar", 7e9) == pytest.approx(2.4422306306649472e-08, rel=1e-6)
    assert ion.ionrate_adk("Ar", 8e9) == pytest.approx(4.494711488416766e-05, rel=1e-6)
    E = np.linspace(1e9, 1e11, 32)
    assert np.array_equal(ion.ionrate_adk("He", E), ion.ionrate_adk("He", -E))


def test_ppt_golden(pkg):
    from luna_jl_amd import ionisation as ion
    r1 = ion.ionrate_ppt("He", 800e-9, 1e10, dipole_corr=False)[0]
    r2 = ion.ionrate_ppt("He", 800e-9, 1.3e10, dipole_corr=False)[0]
    assert r1 == pytest.approx(2 * 1.40432138471583e-5, rel=1e-3)
    assert r2 == pytest.approx(2 * 0.04517809797503506, rel=1e-3)


def test_ppt_vectorised_equals_scalar(pkg):
    from luna_jl_amd import ionisation as ion
    for gas in ("He", "Ar"):
        o = ion.ppt_options(gas)
        f = ion.ionrate_fun_ppt(o["ionpot"], 800e-9, o["Z"], o["l"], dalpha=o["dalpha"], alpha_ion=o["alpha_ion"])
        Emax = 2 * ion.barrier_suppression(o["ionpot"], o["Z"])
        E = np.linspace(Emax / 5000, Emax, 40)
        a = np.array([f(e) for e in E])
        b = ion.ionrate_ppt_vec(o["ionpot"], 800e-9, o["Z"], o["l"], E, dalpha=o["dalpha"], alpha_ion=o["alpha_ion"])
        np.testing.assert_allclose(b, a, rtol=1e-12)


def test_ppt_accel_vs_direct_and_oracle_lookup(pkg, orc, capi):
    # PPTaccel vs direct rate, test_ionisation.jl:36-47 (rtol 1e-2), and the
    # oracle's spline lookup reproduces the table at its nodes.
    from luna_jl_amd import ionisation as ion
    E, rate = ion.make_ppt_cache("He", 800e-9, N=2 ** 16)
    spl = ion.PPTSpline(E, rate)
    h = capi.DescHolder(ppt_n=len(spl.x), ppt_x=spl.x, ppt_y=spl.y, ppt_d=spl.d, plasma=1)
    Eq = np.linspace(1e9, 1e11, 2 ** 9)
    st, out = orc.ppt_rate(h, Eq)
    assert st == 0
    direct = ion.ionrate_ppt("He", 800e-9, Eq)
    np.testing.assert_allclose(out, direct, rtol=1e-2)
    st, outn = orc.ppt_rate(h, -Eq)
    assert np.array_equal(out, outn)
    st, at_nodes = orc.ppt_rate(h, spl.x[::997])
    np.testing.assert_allclose(at_nodes, rate[rate > 0][::997], rtol=1e-13)
    st, _ = orc.ppt_rate(h, np.array([spl.x[-1] * 1.001]))
    assert st == capi.LH_E_PPT_RANGE


def test_capillary_loss_golden(pkg):
    from luna_jl_amd import modes, physdata as pd
    w = np.array([pd.wlfreq(800e-9)])
    m = modes.MarcatiliMode(125e-6, "HeB", 1.0, model="reduced")
    assert m.losslength(w)[0] == pytest.approx(7.0593180702769, rel=1e-5)
    assert m.dB_per_m(w)[0] == pytest.approx(0.6152074146252722, rel=1e-5)
    m2 = modes.MarcatiliMode(250e-6, "HeB", 1.0, model="reduced")
    assert m.dB_per_m(w)[0] == pytest.approx(8 * m2.dB_per_m(w)[0], rel=1e-12)


def test_density_refprop(pkg):
    from luna_jl_amd import physdata as pd
    assert pd.density("Ar", 2.0, 294.0) == pytest.approx(4.933761614600933e25, rel=1e-4)
    assert pd.density("Ar", 40.0, 294.0) == pytest.approx(1.0101579129300146e27, rel=1e-4)
    assert pd.density("HeB", 2.0, 294.0) == pytest.approx(4.9224080868066745e25, rel=1e-4)
    assert pd.density("HeB", 40.0, 294.0) == pytest.approx(9.66755645602771e26, rel=1e-4)


def test_constants(pkg):
    from luna_jl_amd import physdata as pd
    assert pd.eps_0 == 8.854187817620389e-12
    assert pd.au_energy == pytest.approx(4.359744650e-18, rel=1e-9)


@pytest.mark.parametrize("trange,nt,nto,nw", [(1e-12, 8192, 16384, 4097), (0.5e-12, 4096, 8192, 2049)])
def test_real_grid_sizes(pkg, trange, nt, nto, nw):
    from luna_jl_amd.grid import RealGrid
    g = RealGrid(1.0, 800e-9, (150e-9, 4e-6), trange)
    assert (g.n_t, g.n_to, g.n_omega) == (nt, nto, nw)
    dt = g.t[1] - g.t[0]
    assert math.isclose(dt / g.dt_o, g.n_to / g.n_t, rel_tol=1e-12)
    assert math.isclose(dt, math.pi / g.omega.max(), rel_tol=1e-12)
    assert g.omega_win[g.sidx].max() == 1.0 and g.omega_win[~g.sidx].max() == 0.0


def test_env_grid_sizes(pkg):
    from luna_jl_amd.grid import EnvGrid
    g = EnvGrid(0.15, 835e-9, (400e-9, 2400e-9), 6e-12)
    assert g.n_t == g.n_to == g.n_omega == 8192
    g2 = EnvGrid(1.41, 835e-9, (450e-9, 8000e-9), 13e-12)
    assert g2.n_t == 16384


def test_input_energy(pkg):
    # PulseField energy normalisation, test/test_fields.jl:69-108 (rtol 1e-14)
    from luna_jl_amd import fields
    from luna_jl_amd.grid import RealGrid
    g = RealGrid(1.0, 800e-9, (150e-9, 4e-6), 1e-12)
    Ew = fields.GaussField(800e-9, 10e-15, energy=120e-6)(g)
    Et = np.fft.irfft(Ew, g.n_t)
    assert fields.energy_t(g, Et) == pytest.approx(120e-6, rel=1e-14)
