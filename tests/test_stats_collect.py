"""Statistics evaluated on the host: ``collect_stats`` and its functions
(src/Stats.jl:18-199, 323-339, 411-493), ``Stats.default``'s userfuns
(:523-529) and the plumbing that hands them to the outputs.

* ω0, energy, peakpower and fwhm_t on the host (numpy) against the oracle's
  C restatement of the same statistics (itself pinned by closed forms in
  test_oracle_stats.py), on field-grid and envelope-grid states: 1e-12.
* plan_analytic's transform: for a field with no Nyquist content the analytic
  field equals the Hilbert transform of the real time field (Maths.hilbert)
  to 1e-12.
* Maths.fwhm: the reference's cases (test/test_maths.jl:121-137) and the
  no-crossing NaN of its try/catch.
* 2-D fields: per-column values equal the 1-D ones column by column.
* Outputs: one dict per step stacks into per-step arrays in MemoryOutput and
  HDF5Output; userfun values replace a device value of the same name.
"""
import math

import numpy as np
import pytest


def gauss_fwhm(x, fwhm, x0=0.0):
    return np.exp(-4 * math.log(2) * ((x - x0) / fwhm) ** 2)


@pytest.fixture(scope="module")
def small(pkg):
    from luna_jl_amd import configs
    return configs.capillary_instance(energy=150e-6, pressure=2.0, trange=0.5e-12)


def _evolved(inst, orc, capi, z=0.4):
    st, E, _, _, _ = orc.solve(inst.holder, inst.Ew0, capi.rk_opts(z, save_n=2), want_saves=False)
    assert st == 0
    return E


def _host_record(grid, Ew, z, dz):
    from luna_jl_amd import stats
    from luna_jl_amd.fields import energyfuncs
    _, ew = energyfuncs(grid)
    hs = stats.collect_stats(grid, Ew, stats.omega0(grid), stats.energy(grid, ew), stats.peakpower(grid),
                             stats.fwhm_t(grid))
    return hs(Ew, z, dz)


@pytest.mark.parametrize("evolved", [False, True])
def test_host_stats_match_oracle_field_grid(small, orc, capi, evolved):
    g = small.grid
    Ew = _evolved(small, orc, capi) if evolved else small.Ew0
    st, rec = orc.step_stats(small.holder, g.t, small.transform.aeff, Ew, 0.4, 2e-3)
    assert st == 0
    r = dict(zip(capi.STAT_KEYS, rec))
    d = _host_record(g, Ew, 0.4, 2e-3)
    assert d["z"] == 0.4 and d["dz"] == 2e-3
    for k in ("ω0", "energy", "peakpower", "fwhm_t_min", "fwhm_t_max"):
        assert d[k] == pytest.approx(r[k], rel=1e-12, abs=0), k


def test_host_stats_match_oracle_envelope_grid(pkg, orc, capi):
    from luna_jl_amd import configs
    inst = configs.gnlse_instance(flength=0.02, shotnoise=False)
    g = inst.grid
    Ew = _evolved(inst, orc, capi, z=0.02)
    st, rec = orc.step_stats(inst.holder, g.t, 1.0, Ew, 0.02, 1e-4)
    assert st == 0
    r = dict(zip(capi.STAT_KEYS, rec))
    d = _host_record(g, Ew, 0.02, 1e-4)
    for k in ("ω0", "energy", "peakpower", "fwhm_t_min", "fwhm_t_max"):
        assert d[k] == pytest.approx(r[k], rel=1e-12, abs=0), k


def test_analytic_is_hilbert_of_real_field(small):
    from luna_jl_amd import stats
    from luna_jl_amd.maths import hilbert
    g = small.grid
    Ew = small.Ew0.copy()
    Ew[-1] = 0.0  # no Nyquist content
    n = 2 * (len(Ew) - 1)
    Et_real = np.fft.irfft(Ew, n)
    a = stats.analytic(g, Ew)
    h = hilbert(Et_real)
    assert np.linalg.norm(a - h) <= 1e-12 * np.linalg.norm(h)
    # a dropped factor 2 (copyto_fft!) or a forward transform fails
    assert np.linalg.norm(0.5 * a - h) > 0.1 * np.linalg.norm(h)


def test_fwhm_host(pkg):
    from luna_jl_amd.maths import fwhm
    x = np.linspace(-1, 1, 2048)
    assert fwhm(x, gauss_fwhm(x, 0.2)) == pytest.approx(0.2, rel=1e-4)
    assert fwhm(x, gauss_fwhm(x, 0.2), "max") == pytest.approx(0.2, rel=1e-4)
    fw, sep = 0.1, 0.5
    y2 = gauss_fwhm(x, fw, -sep / 2) + gauss_fwhm(x, fw, sep / 2)
    assert fwhm(x, y2, "min") == pytest.approx(fw, rel=1e-4)
    assert fwhm(x, y2, "max") == pytest.approx(fw + sep, rel=1e-4)
    assert math.isnan(fwhm(x, np.ones(2048)))
    assert math.isnan(fwhm(x, np.where(x > 0.5, 1.0, 0.0), "max"))


def test_two_dimensional_fields(small):
    from luna_jl_amd import stats
    from luna_jl_amd.fields import energyfuncs
    g = small.grid
    E2 = np.stack([small.Ew0, 0.5 * np.roll(small.Ew0, 40)], axis=1)
    _, ew = energyfuncs(g)
    hs = stats.collect_stats(g, E2, stats.omega0(g), stats.energy(g, ew), stats.peakpower(g), stats.fwhm_t(g))
    d = hs(E2, 0.1, 1e-3)
    for i in range(2):
        d1 = _host_record(g, E2[:, i], 0.1, 1e-3)
        for k in ("ω0", "energy", "peakpower", "fwhm_t_min", "fwhm_t_max"):
            assert d[k][i] == pytest.approx(d1[k], rel=1e-14), k
    Et = stats.analytic(g, E2)
    assert d["peakpower_allmodes"] == pytest.approx(np.max(np.sum(np.abs(Et) ** 2, axis=1)), rel=1e-14)
    assert math.isfinite(d["fwhm_t_min_allmodes"])


def test_energy_lambda_window(small):
    from luna_jl_amd import stats
    from luna_jl_amd.fields import energyfuncs
    g = small.grid
    _, ew = energyfuncs(g)
    f = stats.energy_lambda(g, ew, (700e-9, 900e-9))
    d = {}
    f(d, small.Ew0, None, 0.0, 0.0)
    (k, v), = d.items()
    assert k == "energy_700.00nm_900.00nm"
    assert 0.5 * ew(small.Ew0) < v < ew(small.Ew0)


def test_collect_stats_in_memory_output(small):
    from luna_jl_amd import stats
    from luna_jl_amd.output import MemoryOutput
    g = small.grid
    hs = stats.collect_stats(g, small.Ew0, stats.omega0(g), stats.peakpower(g))
    o = MemoryOutput(0, 1.0, 3, hs)
    ys = [small.Ew0 * (1 + 0.1 * k) for k in range(4)]
    for k, y in enumerate(ys):
        rec = stats.StepRecords(None, [hs.host_step(y, 0.1 * (k + 1), 0.05)])
        nsv = 1 if k < 3 else 0  # the last step saves nothing
        o.on_step(k, np.repeat(y[:, None], nsv, axis=1), np.array([0.5 * k][:nsv]), 0.1 * (k + 1), 0.05, y, rec)
    st = o.data["stats"]
    assert set(st) == {"ω0", "peakpower", "z", "dz"}
    assert np.allclose(st["z"], [0.1, 0.2, 0.3, 0.4]) and st["z"].shape == (4,)
    pp = [float(np.max(np.abs(stats.analytic(g, y)) ** 2)) for y in ys]
    assert np.allclose(st["peakpower"], pp, rtol=1e-14)
    assert hs.names() == ["ω0", "peakpower", "z", "dz"]


def test_userfuns_merge_with_device_records(small, capi):
    from luna_jl_amd import stats
    sf = stats.StatsFun(small.grid, small.transform, userfuns=[
        lambda d, Ew, Et, z, dz: d.__setitem__("sum_abs2", float(np.sum(np.abs(Ew) ** 2))),
        lambda d, Ew, Et, z, dz: d.__setitem__("energy", -1.0)])
    assert sf.has_host
    rec = np.zeros((1, capi.LH_NSTAT))
    rec[0, capi.STAT_KEYS.index("energy")] = 7.0
    rec[0, capi.STAT_KEYS.index("z")] = 0.3
    hv = sf.host_step(small.Ew0, 0.3, 0.01)
    d = sf.to_dict(stats.StepRecords(rec, [hv]))
    assert d["z"][0] == 0.3
    assert d["sum_abs2"][0] == pytest.approx(np.sum(np.abs(small.Ew0) ** 2), rel=1e-15)
    assert d["energy"][0] == -1.0  # a later function of the same name wins
    assert "sum_abs2" in sf.names() and sf.names().count("energy") == 1


def test_collect_stats_in_hdf5_output(small, tmp_path):
    from luna_jl_amd import output as out
    from luna_jl_amd import stats
    try:
        out.h5lib()
    except out.HDF5Unavailable:
        pytest.skip("no HDF5 library")
    g = small.grid
    hs = stats.collect_stats(g, small.Ew0, stats.omega0(g))
    o = out.HDF5Output(str(tmp_path / "s.h5"), 0, 1.0, 2, hs)
    for k in range(3):
        y = small.Ew0 * (1 + k)
        rec = stats.StepRecords(None, [hs.host_step(y, 0.2 * (k + 1), 0.1)])
        nsv = 1 if k < 2 else 0
        o.on_step(k, np.repeat(y[:, None], nsv, axis=1), np.array([0.5 * k][:nsv]), 0.2 * (k + 1), 0.1, y,
                  rec)
    assert np.allclose(o["stats", "z"], [0.2, 0.4, 0.6])
    assert o["stats", "ω0"].shape == (3,)


@pytest.mark.parametrize("evolved", [False, True])
def test_host_record_matches_oracle(small, orc, capi, evolved):
    """StatsFun.host_record (the grids whose kernels record no statistics)
    against the oracle's record of the same state, energy windows included."""
    from luna_jl_amd import stats
    g = small.grid
    Ew = _evolved(small, orc, capi) if evolved else small.Ew0
    sf = stats.StatsFun(g, small.transform, capacity=16, windows=[(700e-9, 900e-9), ((300e-9, 500e-9), 1e14)])
    st, ref = orc.step_stats(small.holder, g.t, small.transform.aeff, Ew, 0.4, 2e-3,
                             windows=[w for _, w in sf.windows])
    assert st == 0
    rec = sf.host_record(Ew, 0.4, 2e-3)
    for k, name in enumerate(capi.STAT_KEYS):
        if np.isnan(ref[k]):
            assert np.isnan(rec[k]), name
        else:
            tol = 1e-10 if name in ("electrondensity", "peak_ionisation_rate") else 1e-12
            assert rec[k] == pytest.approx(ref[k], rel=tol, abs=0), name
    assert rec[capi.STAT_KEYS.index("electrondensity")] > 0


def test_host_record_envelope(pkg, orc, capi):
    from luna_jl_amd import configs, stats
    inst = configs.gnlse_instance(flength=0.02, shotnoise=False)
    Ew = _evolved(inst, orc, capi, z=0.02)
    sf = stats.StatsFun(inst.grid, inst.transform, capacity=16)
    st, ref = orc.step_stats(inst.holder, inst.grid.t, inst.transform.aeff, Ew, 0.02, 1e-4)
    assert st == 0
    rec = sf.host_record(Ew, 0.02, 1e-4)
    for k, name in enumerate(capi.STAT_KEYS):
        if np.isnan(ref[k]):
            assert np.isnan(rec[k]), name
        else:
            assert rec[k] == pytest.approx(ref[k], rel=1e-12, abs=0), name
