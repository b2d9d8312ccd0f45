"""Host side of the per-step statistics (no GPU): the record layout, the keys
Stats.default produces (src/Stats.jl:492-532) and the pressure constant."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def small(pkg):
    from luna_jl_amd import configs
    return configs.capillary_instance(energy=150e-6, pressure=2.0, trange=0.5e-12, ppt_n=2 ** 12)


def test_keys_follow_the_reference(small, capi):
    from luna_jl_amd import stats
    sf = stats.default(small.grid, small.Ew0, None, small.linop, small.transform, gas="He")
    rec = np.arange(2 * capi.LH_NSTAT, dtype=float).reshape(2, capi.LH_NSTAT)
    rec[:, capi.STAT_KEYS.index("density")] = small.transform.density  # PhysData.pressure of this density
    d = sf.to_dict(rec)
    for k in ("ω0", "energy", "peakpower", "fwhm_t_min", "fwhm_t_max", "density", "pressure",
              "peakintensity", "electrondensity", "peak_ionisation_rate", "z", "dz"):
        assert len(d[k]) == 2, k
    assert d["z"][1] == rec[1, 0] and d["dz"][0] == rec[0, 1]
    assert d["pressure"][0] == pytest.approx(2.0, rel=1e-12)


def test_no_plasma_no_ionisation_keys(pkg, capi):
    from luna_jl_amd import configs, stats
    cfg1 = configs.gnlse_instance(shotnoise=False)
    sf = stats.default(cfg1.grid, cfg1.Ew0, None, cfg1.linop, cfg1.transform)
    d = sf.to_dict(np.zeros((3, capi.LH_NSTAT)))
    assert "electrondensity" not in d and "peak_ionisation_rate" not in d
    assert "pressure" not in d


def test_unsupported_options_raise(small):
    from luna_jl_amd import stats
    with pytest.raises(NotImplementedError):
        stats.default(small.grid, small.Ew0, None, small.linop, small.transform, onaxis=True)
    with pytest.raises(ValueError):  # the record holds four windows
        stats.default(small.grid, small.Ew0, None, small.linop, small.transform,
                      windows=[(200e-9 + 1e-8 * i, 300e-9) for i in range(5)])


def test_window_keys_and_columns(small, capi):
    """energy_λ labels (src/Stats.jl:52-61): "energy_%.2fnm_%.2fnm" of the
    sorted limits; (λlims, winwidth) pairs accepted; the columns map in order."""
    from luna_jl_amd import stats
    sf = stats.default(small.grid, small.Ew0, None, small.linop, small.transform,
                       windows=[(300e-9, 200e-9), ((700e-9, 900e-9), 2e13)])
    rec = np.zeros((2, capi.LH_NSTAT))
    rec[:, capi.STAT_KEYS.index("window0")] = [1.0, 2.0]
    rec[:, capi.STAT_KEYS.index("window1")] = [3.0, 4.0]
    d = sf.to_dict(rec)
    assert list(d["energy_200.00nm_300.00nm"]) == [1.0, 2.0]
    assert list(d["energy_700.00nm_900.00nm"]) == [3.0, 4.0]
    assert not any(k.startswith("window") for k in d)
    desc = sf.desc()
    assert desc.n_windows == 2 and sf.window_w.shape == (2, len(small.grid.omega))


def test_planck_taper_shape():
    """Maths.planck_taper (src/Maths.jl:422-449): 1 between left1 and right1
    (ends included), 0 outside [left0, right0], 1/2 at the middle of a taper
    (z = 0 there), and the zero-width case is a rectangle with both ends."""
    from luna_jl_amd import stats
    x = np.linspace(0.0, 10.0, 1001)
    w = stats.planck_taper(x, 2.0, 3.0, 7.0, 9.0)
    assert np.all(w[(x >= 3.0) & (x <= 7.0)] == 1.0)
    assert np.all(w[(x <= 2.0) | (x >= 9.0)] == 0.0)
    assert w[np.argmin(abs(x - 2.5))] == pytest.approx(0.5, abs=1e-12)
    assert w[np.argmin(abs(x - 8.0))] == pytest.approx(0.5, abs=1e-12)
    assert np.all(np.diff(w[(x > 2.0) & (x < 3.0)]) >= 0)  # saturates to 0 and 1 at the ends
    r = stats.planck_taper(x, 3.0, 3.0, 7.0, 7.0)
    assert np.array_equal(r, ((x >= 3.0) & (x <= 7.0)).astype(float))


def test_pressure_inverts_density():
    from luna_jl_amd import physdata as pd
    for gas in ("He", "Ar", "Ne"):
        for P in (0.1, 1.0, 3.0, 20.0):
            assert pd.pressure(gas, pd.density(gas, P)) == pytest.approx(P, rel=1e-12)
    assert pd.pressure("He", 0.0) == 0.0


def test_energy_weights_envelope_is_rectangle_rule(pkg):
    from luna_jl_amd import configs, stats
    g = configs.gnlse_instance(shotnoise=False).grid
    w = stats.energy_weights(g)
    dt = g.t[1] - g.t[0]
    # Parseval for the unnormalised DFT: Σ|X|²·δt/N = Σ|x|²·δt
    assert np.allclose(w, dt / len(g.t), rtol=1e-12)


def test_zdw_reference_values(pkg):
    """Modes.zdw of Marcatili modes against test/test_capillary.jl:101-104, 109-119
    (the Ar case to 1e-5: the reference's 1e-7 rests on CoolProp's density)."""
    from luna_jl_amd import modes
    assert abs(1e9 * modes.zdw(modes.MarcatiliMode(125e-6, "HeB", 0.4, model="reduced")) - 379) < 1
    assert abs(1e9 * modes.zdw(modes.MarcatiliMode(75e-6, "HeB", 5.9, model="reduced")) - 562) < 1
    m = modes.MarcatiliMode(50e-6, "ArB", 2.0, model="reduced", T=294.0)
    assert modes.zdw(m) == pytest.approx(7.289431065526978e-07, rel=1e-5)
    from luna_jl_amd import physdata as pd
    assert modes.beta2(m, pd.wlfreq(800e-9)) == pytest.approx(-1.68187063100098e-29, rel=1e-4)
    # no zero crossing in the window → NaN, as the reference's missing
    assert np.isnan(modes.zdw(m, lam_min=900e-9, lam_max=1000e-9))


def test_stats_zdw_entries(small, pkg, capi):
    from luna_jl_amd import interface, stats
    Ew, grid, linop, tr, _, out = interface.prop_capillary_args(
        125e-6, 0.2, "He", 1.0, lambda0=800e-9, lambda_lims=(150e-9, 4e-6), trange=0.5e-12,
        tau_fwhm=10e-15, energy=100e-6, shotnoise=False, ppt_N=2 ** 12)
    rec = np.zeros((2, capi.LH_NSTAT))
    rec[:, capi.STAT_KEYS.index("density")] = tr.density
    d = out.statsfun.to_dict(rec)
    assert np.all(d["zdw"] == d["zdw"][0]) and 300e-9 < d["zdw"][0] < 700e-9
    Ew, grid, linop, tr, _, out = interface.prop_capillary_args(
        125e-6, 0.2, "He", (0.5, 3.0), lambda0=800e-9, lambda_lims=(150e-9, 4e-6), trange=0.5e-12,
        tau_fwhm=10e-15, energy=100e-6, shotnoise=False, ppt_N=2 ** 12)
    rec = np.zeros((2, capi.LH_NSTAT))
    rec[:, 0] = [0.0, 0.2]
    d = out.statsfun.to_dict(rec)
    assert d["zdw"][1] > d["zdw"][0]  # more gas at the end: normal dispersion moves the ZDW to the red


def test_onaxis_area(pkg):
    """Stats.default(...; onaxis=true) for one HE1m mode (src/Stats.jl:159-174,
    260-290): I(0) = P / (π a² J1(u)²).  Checked against the power integral of
    the mode's intensity profile I(r) ∝ J0(u r/a)², by quadrature: a wrong
    normalisation N or a dropped square fails."""
    from scipy import integrate, special
    from luna_jl_amd import modes
    for a, m in ((125e-6, 1), (50e-6, 2)):
        mode = modes.MarcatiliMode(a, "He", 1.0, m=m)
        u = special.jn_zeros(0, m)[-1]
        area, _ = integrate.quad(lambda r: 2 * np.pi * r * special.jv(0, u * r / a) ** 2, 0, a,
                                 epsabs=0, epsrel=1e-13, limit=200)
        assert mode.aeff_onaxis() == pytest.approx(area, rel=1e-10)
        assert mode.aeff_onaxis() != pytest.approx(mode.Aeff(), rel=1e-3)
    with pytest.raises(NotImplementedError):
        modes.MarcatiliMode(125e-6, "He", 1.0, kind="TE").aeff_onaxis()


def test_onaxis_stats_use_the_onaxis_area(pkg):
    from luna_jl_amd import interface, modes, stats
    kw = dict(lambda0=800e-9, lambda_lims=(150e-9, 4e-6), trange=0.5e-12, tau_fwhm=10e-15, energy=100e-6,
              shotnoise=False, ppt_N=2 ** 12)
    Ew, grid, linop, tr, _, out = interface.prop_capillary_args(125e-6, 0.2, "He", 1.0, **kw)
    mode = modes.MarcatiliMode(125e-6, "He", 1.0)
    sf = stats.default(grid, Ew, mode, linop, tr, onaxis=True)
    assert sf.aeff == mode.aeff_onaxis() and sf.desc().aeff == sf.aeff
    assert stats.default(grid, Ew, mode, linop, tr).aeff == tr.aeff
    Ew, grid, linop, tr, _, out = interface.prop_capillary_args(125e-6, 0.2, "He", (0.5, 1.5), **kw)
    with pytest.raises(NotImplementedError):
        stats.default(grid, Ew, mode, linop, tr, onaxis=True)
