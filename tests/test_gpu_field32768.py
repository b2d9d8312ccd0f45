"""The Nto = 32768 field grid (prop_capillary with a 2 ps window over
150 nm-4 μm: Nt = 16384, Nω = 8193), whose c2r/r2c transforms are staged
through HBM and whose pointwise responses run over the series in two LDS
windows with the plasma integrals carried between them.

* TransModeAvg's RHS (src/NonlinearRHS.jl:323-373, Kerr + PPT plasma) at the
  input and at a chirped, stronger field: relative L2 ≤ 1e-12 against the
  oracle.
* A 5 cm propagation (Luna.run, src/Luna.jl:358-391): the same step counts as
  the oracle and ≤ 1e-6 relative spectral-energy error at every save.
* Per-step statistics on the device and z-dependent fills are refused on this
  grid (LH_E_ARG); prop_capillary's statistics are then evaluated on the host
  from every accepted step, streamed, with the trajectory unchanged.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def spec_err(a, b, s):
    pa, pb = np.abs(a[s]) ** 2, np.abs(b[s]) ** 2
    return np.linalg.norm(pa - pb) / np.linalg.norm(pb)


@pytest.fixture(scope="module")
def big(pkg):
    from luna_jl_amd import configs
    inst = configs.capillary_instance(energy=150e-6, pressure=2.0, trange=2e-12, flength=0.05)
    assert inst.grid.n_to == 32768 and inst.grid.n_omega == 8193
    return inst


def test_rhs_32768(big, orc):
    print("case: field grid Nto = 32768 rhs", flush=True)
    from luna_jl_amd import runtime
    ctx = runtime.Context(big.holder)
    rng = np.random.default_rng(5)
    chirped = 1.5 * big.Ew0 * np.exp(1j * rng.uniform(0, 0.3, big.Ew0.size))
    for E in (big.Ew0, chirped):
        nl = ctx.rhs(E)
        st, ref = orc.rhs(big.holder, E)
        assert st == 0 and np.linalg.norm(ref) > 0
        assert rel_l2(nl, ref) <= 1e-12, rel_l2(nl, ref)


def test_propagation_32768(big, orc, capi):
    print("case: field grid Nto = 32768, 5 cm propagation", flush=True)
    from luna_jl_amd import runtime
    ctx = runtime.Context(big.holder)
    opts = capi.rk_opts(big.zmax, save_n=6)
    Ew, sv, zs, stats = ctx.solve(big.Ew0, opts)
    st, Er, sv_r, zs_r, stats_r = orc.solve(big.holder, big.Ew0, opts)
    print("steps gpu/oracle", stats.steps, stats_r.steps, flush=True)
    assert stats.status == 0 and st == 0
    assert (stats.steps, stats.accepted, stats.rejected) == (stats_r.steps, stats_r.accepted, stats_r.rejected)
    for k in range(6):
        assert spec_err(sv[k], sv_r[k], big.sidx) <= 1e-6, k
    assert spec_err(Ew, Er, big.sidx) <= 1e-6


def test_refusals_32768(big, capi):
    print("case: field grid Nto = 32768 refusals", flush=True)
    from luna_jl_amd import runtime, stats
    ctx = runtime.Context(big.holder)
    with pytest.raises(capi.LunaHipError):
        ctx.set_step_stats(stats.StatsFun(big.grid, big.transform, capacity=64))


def test_prop_capillary_32768_statistics(pkg, capi):
    """prop_capillary on the 2 ps grid with its default statistics: the kernel
    records none on this grid, so Luna.run streams every accepted step and
    evaluates the same record on the host; the saves equal lh_solve's."""
    print("case: prop_capillary 5 cm on the Nto = 32768 grid, statistics on the host", flush=True)
    from luna_jl_amd import interface, runtime
    Ew0, grid, linop, tr, _, out = interface.prop_capillary_args(
        125e-6, 0.05, "He", 2.0, lambda0=800e-9, lambda_lims=(150e-9, 4e-6), trange=2e-12, tau_fwhm=10e-15,
        energy=150e-6, shotnoise=False, saveN=6)
    assert grid.n_to == 32768
    _, st = interface.run(Ew0, grid, linop, tr, None, out)
    assert st.status == 0
    s = out.data["stats"]
    assert len(s["z"]) == st.accepted and np.all(np.diff(s["z"]) > 0) and s["z"][-1] == st.z
    assert s["electrondensity"][-1] > 0 and np.all(np.isfinite(s["fwhm_t_min"]))
    e0 = out.statsfun.host_record(Ew0, 0.0, 0.0)[capi.STAT_KEYS.index("energy")]
    assert abs(s["energy"][0] - e0) <= 1e-2 * e0
    ctx = runtime.Context(tr.descriptor(linop))
    _, sv, _, st2 = ctx.solve(Ew0, capi.rk_opts(grid.zmax, save_n=6))
    assert st2.accepted == st.accepted
    assert np.array_equal(out.data["Eω"], sv.T)
