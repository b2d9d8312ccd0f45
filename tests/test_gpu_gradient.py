"""GPU parity of the gradient fill (z-dependent density, operator and
normalisation re-evaluated at every new z on the device) against the oracle.

Single RHS at several z: relative L2 ≤ 1e-12.  Propagations: the same step
counts and relative spectral-energy error ≤ 1e-9 (5 cm) and ≤ 1e-6 (the full
length); batch ≡ single; the statistics' density follows ρ(z).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW = dict(lambda0=800e-9, lambda_lims=(150e-9, 4e-6), tau_fwhm=10e-15, shotnoise=False, ppt_N=2 ** 12)


def setup(pressure, flength, trange=0.5e-12, energy=150e-6):
    from luna_jl_amd import interface
    Ew, grid, linop, tr, _, out = interface.prop_capillary_args(125e-6, flength, "He", pressure, trange=trange,
                                                                energy=energy, **KW)
    return Ew, grid, linop, tr, tr.descriptor(linop)


def spec_err(a, b, sidx):
    pa, pb = np.abs(a[sidx]) ** 2, np.abs(b[sidx]) ** 2
    return np.linalg.norm(pa - pb) / np.linalg.norm(pb)


@pytest.mark.parametrize("trange", [0.5e-12, 1e-12])
def test_rhs_at_z(orc, trange):
    print(f"case: gradient rhs, trange {trange}", flush=True)
    from luna_jl_amd import runtime
    Ew, grid, linop, tr, h = setup((0.5, 2.5), 0.2, trange=trange)
    ctx = runtime.Context(h)
    for z in (0.0, 0.05, 0.13, 0.2, 0.3):
        nl = ctx.rhs(Ew * 2.0, z)
        st, ref = orc.rhs_grad(h, tr.gradient, Ew * 2.0, z)
        assert st == 0
        assert np.linalg.norm(nl - ref) / np.linalg.norm(ref) <= 1e-12, z


def test_solve_short(orc, capi):
    print("case: gradient solve 5 cm", flush=True)
    from luna_jl_amd import runtime
    Ew, grid, linop, tr, h = setup((0.5, 2.5), 0.05)
    ctx = runtime.Context(h)
    opts = capi.rk_opts(0.05, save_n=11)
    Eg, sv, zs, stats = ctx.solve(Ew, opts)
    st, Er, sv_r, zs_r, stats_r = orc.solve_grad(h, tr.gradient, Ew, opts)
    assert st == 0 and stats.status == 0
    assert (stats.steps, stats.accepted, stats.rhs_evals) == (stats_r.steps, stats_r.accepted, stats_r.rhs_evals)
    assert spec_err(Eg, Er, grid.sidx) <= 1e-9
    for k in range(11):
        assert spec_err(sv[k], sv_r[k], grid.sidx) <= 1e-9, k


def test_solve_full_multipoint(orc, capi):
    print("case: multi-point gradient solve 0.5 m", flush=True)
    from luna_jl_amd import runtime
    Ew, grid, linop, tr, h = setup(([0.0, 0.2, 0.5], [0.5, 2.5, 1.0]), 0.5)
    ctx = runtime.Context(h)
    opts = capi.rk_opts(0.5, save_n=51)
    Eg, sv, zs, stats = ctx.solve(Ew, opts)
    st, Er, sv_r, zs_r, stats_r = orc.solve_grad(h, tr.gradient, Ew, opts)
    assert st == 0 and stats.status == 0 and stats.steps == stats_r.steps
    assert spec_err(Eg, Er, grid.sidx) <= 1e-6
    assert spec_err(sv[25], sv_r[25], grid.sidx) <= 1e-6


def test_batch_with_gradient_and_stats(capi):
    print("case: batch of gradient and constant fills, statistics", flush=True)
    from luna_jl_amd import runtime, stats
    insts = [setup((0.5, 2.5), 0.05), setup(1.5, 0.05), setup((2.0, 0.7), 0.05, trange=1e-12)]
    ctxs = [runtime.Context(i[4]) for i in insts]
    for (Ew, grid, linop, tr, h), ctx in zip(insts, ctxs):
        ctx.set_step_stats(stats.StatsFun(grid, tr, capacity=1024))
    opts = capi.rk_opts(0.05, save_n=3)
    single = []
    for (Ew, grid, linop, tr, h), ctx in zip(insts, ctxs):
        single.append((ctx.solve(Ew, opts)[0], ctx.step_stats()))
    outs, svs, zss, st = runtime.batch_solve(ctxs, [i[0] for i in insts], opts)
    for i, ctx in enumerate(ctxs):
        assert st[i].status == 0
        assert np.array_equal(outs[i], single[i][0])
        assert np.array_equal(ctx.step_stats(), single[i][1], equal_nan=True)
    rec = single[0][1]
    g = insts[0][3].gradient.gradient
    for r in rec:
        assert r[capi.STAT_KEYS.index("density")] == pytest.approx(g.density(r[0]), rel=1e-12)


def setup_taper(radius, pressure, flength, trange=0.5e-12, energy=150e-6):
    from luna_jl_amd import interface
    Ew, grid, linop, tr, _, out = interface.prop_capillary_args(radius, flength, "He", pressure, trange=trange,
                                                                energy=energy, **KW)
    return Ew, grid, linop, tr, tr.descriptor(linop)


def test_taper_rhs_solve_and_stats(orc, capi):
    """A linear taper a(z) with a pressure gradient: one RHS at several z
    (1e-12), a 5 cm propagation (same steps, 1e-9) and its per-step records
    (Aeff(z) in the peak intensity) against the oracle."""
    print("case: taper 125 -> 90 um with a gradient", flush=True)
    from luna_jl_amd import capillary, runtime, stats
    L = 0.05
    Ew, grid, linop, tr, h = setup_taper((125e-6, 90e-6), (0.5, 2.5), L)
    ctx = runtime.Context(h)
    for z in (0.0, 0.013, 0.04, 0.05, 0.08):
        nl = ctx.rhs(Ew * 2.0, z)
        st, ref = orc.rhs_grad(h, tr.gradient, Ew * 2.0, z)
        assert st == 0
        assert np.linalg.norm(nl - ref) / np.linalg.norm(ref) <= 1e-12, z
    sf = stats.StatsFun(grid, tr, capacity=2048)
    ctx.set_step_stats(sf)
    opts = capi.rk_opts(L, save_n=6)
    Eg, sv, zs, st_g = ctx.solve(Ew, opts)
    rec = ctx.step_stats()
    st, Er, st_r, rec_r = orc.solve_stats(h, Ew, opts, grid.t, tr.aeff, spec=tr.gradient)
    assert st == 0 and st_g.status == 0
    assert (st_g.steps, st_g.accepted) == (st_r.steps, st_r.accepted) and len(rec) == len(rec_r)
    assert spec_err(Eg, Er, grid.sidx) <= 1e-9
    k = {n: i for i, n in enumerate(capi.STAT_KEYS)}
    intg = tr.gradient.taper["sqrt_intg"] ** 2
    for r, q in zip(rec, rec_r):
        for key in ("energy", "peakpower", "peakintensity", "density"):
            assert r[k[key]] == pytest.approx(q[k[key]], rel=1e-7), key
        a = capillary.taper_radius((125e-6, 90e-6, L), r[k["z"]])
        assert r[k["peakintensity"]] == pytest.approx(r[k["peakpower"]] / (a * a * intg), rel=1e-12)


def test_piecewise_taper_rhs_and_solve(orc, capi):
    """A piecewise-linear radius through nodes: RHS at z in each segment, at a
    node and beyond the last one (1e-12), and a 5 cm solve with the same steps
    and a spectral difference within 1e-7 (measured 3.4e-8; the linear taper
    stays below 1e-9 — the operator has a kink at each node, where the steps'
    error estimates, and so the trajectories, separate more)."""
    print("case: piecewise-linear taper", flush=True)
    from luna_jl_amd import runtime
    L = 0.05
    Ew, grid, linop, tr, h = setup_taper(([0.0, 0.02, 0.05], [125e-6, 100e-6, 110e-6]), (0.8, 2.0), L)
    ctx = runtime.Context(h)
    for z in (0.0, 0.01, 0.02, 0.035, 0.05, 0.07):
        nl = ctx.rhs(Ew * 2.0, z)
        st, ref = orc.rhs_grad(h, tr.gradient, Ew * 2.0, z)
        assert st == 0
        assert np.linalg.norm(nl - ref) / np.linalg.norm(ref) <= 1e-12, z
    opts = capi.rk_opts(L, save_n=3)
    Eg, sv, zs, st_g = ctx.solve(Ew, opts)
    st, Er, sv_r, zs_r, st_r = orc.solve_grad(h, tr.gradient, Ew, opts)
    assert st == 0 and st_g.status == 0 and st_g.steps == st_r.steps
    assert spec_err(Eg, Er, grid.sidx) <= 1e-7


def test_invalid_taper_and_window_descriptors(orc, capi):
    """Descriptors the device path cannot honour are refused with LH_E_ARG and
    leave the context usable: a taper kind out of range, nodes that do not
    increase, a non-positive radius; more than four energy windows."""
    print("case: invalid taper and window descriptors", flush=True)
    import ctypes
    from luna_jl_amd import runtime, stats
    Ew, grid, linop, tr, h = setup_taper(([0.0, 0.02, 0.05], [125e-6, 100e-6, 110e-6]), 1.0, 0.05)
    ctx = runtime.Context(h)
    L = capi.lib()
    d = tr.gradient.desc
    zgood, agood = tr.gradient.taper_Z, tr.gradient.taper_A  # (ctypes pointer fields read as views)
    d.taper = 3
    assert L.lh_set_gradient(ctx.handle, ctypes.byref(d)) == capi.LH_E_ARG
    d.taper = 2
    d.taper_Z = None
    assert L.lh_set_gradient(ctx.handle, ctypes.byref(d)) == capi.LH_E_ARG
    zbad = np.array([0.0, 0.03, 0.02])
    d.taper_Z = capi._ptr(zbad)
    assert L.lh_set_gradient(ctx.handle, ctypes.byref(d)) == capi.LH_E_ARG
    abad = np.array([125e-6, 0.0, 110e-6])
    d.taper_Z, d.taper_A = capi._ptr(zgood), capi._ptr(abad)
    assert L.lh_set_gradient(ctx.handle, ctypes.byref(d)) == capi.LH_E_ARG
    d.taper_A = capi._ptr(agood)
    assert L.lh_set_gradient(ctx.handle, ctypes.byref(d)) == capi.LH_OK
    sf = stats.StatsFun(grid, tr, capacity=16)
    sd = sf.desc()
    sd.n_windows = 5
    assert L.lh_set_step_stats(ctx.handle, ctypes.byref(sd)) == capi.LH_E_ARG
    nl = ctx.rhs(Ew, 0.01)  # still usable
    st, ref = orc.rhs_grad(h, tr.gradient, Ew, 0.01)
    assert st == 0 and np.linalg.norm(nl - ref) / np.linalg.norm(ref) <= 1e-12
