"""GPU parity of the field-resolved mode-averaged path against the oracle.

Single RHS (TransModeAvg, src/NonlinearRHS.jl:362-373): relative L2 ≤ 1e-12.
Full propagations (Luna.run, src/Luna.jl:358-391): relative spectral-energy
error ≤ 1e-6 over sidx bins at rtol = 1e-6, same step counts.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def spec_err(a, b, sidx):
    pa, pb = np.abs(a[sidx]) ** 2, np.abs(b[sidx]) ** 2
    return np.linalg.norm(pa - pb) / np.linalg.norm(pb)


@pytest.fixture(scope="module")
def ctx0(cfg0):
    from luna_jl_amd import runtime
    return runtime.Context(cfg0.holder)


def test_rhs_initial_state(cfg0, ctx0, orc):
    print("case: cfg0 rhs initial", flush=True)
    nl = ctx0.rhs(cfg0.Ew0)
    st, ref = orc.rhs(cfg0.holder, cfg0.Ew0)
    assert st == 0
    assert rel_l2(nl, ref) <= 1e-12


def test_rhs_evolved_states(cfg0, ctx0, orc, capi):
    print("case: cfg0 rhs evolved", flush=True)
    opts = capi.rk_opts(0.3, save_n=4)
    st, Ew, sv, zs, stats = orc.solve(cfg0.holder, cfg0.Ew0, opts)
    assert st == 0
    for k in range(1, 4):
        nl = ctx0.rhs(sv[k])
        st, ref = orc.rhs(cfg0.holder, sv[k])
        assert rel_l2(nl, ref) <= 1e-12, (k, rel_l2(nl, ref))


def test_rhs_small_grid(cfg_small, orc):
    print("case: small-grid rhs", flush=True)
    from luna_jl_amd import runtime
    ctx = runtime.Context(cfg_small.holder)
    nl = ctx.rhs(cfg_small.Ew0)
    st, ref = orc.rhs(cfg_small.holder, cfg_small.Ew0)
    assert rel_l2(nl, ref) <= 1e-12


def test_rhs_ppt_range_error(capi, pkg):
    print("case: ppt range", flush=True)
    from luna_jl_amd import configs, runtime
    cfg = configs.capillary_instance(energy=120e-6, pressure=1.0, trange=0.5e-12)
    ctx = runtime.Context(cfg.holder)
    with pytest.raises(capi.LunaHipError) as e:
        ctx.rhs(cfg.Ew0 * 1e4)
    assert e.value.status == capi.LH_E_PPT_RANGE


def test_solve_short(cfg0, ctx0, orc, capi):
    print("case: cfg0 solve 5 cm", flush=True)
    opts = capi.rk_opts(0.05, save_n=11)
    Ew, sv, zs, stats = ctx0.solve(cfg0.Ew0, opts)
    st, Ew_r, sv_r, zs_r, stats_r = orc.solve(cfg0.holder, cfg0.Ew0, opts)
    assert stats.status == 0
    assert (stats.steps, stats.accepted, stats.rejected) == (stats_r.steps, stats_r.accepted, stats_r.rejected)
    assert stats.rhs_evals == stats_r.rhs_evals
    assert spec_err(Ew, Ew_r, cfg0.sidx) <= 1e-9
    for k in range(11):
        assert spec_err(sv[k], sv_r[k], cfg0.sidx) <= 1e-9, k
    assert np.array_equal(zs, zs_r)


def test_solve_full_cfg0(cfg0, ctx0, orc, capi):
    print("case: cfg0 solve 1 m", flush=True)
    opts = capi.rk_opts(cfg0.zmax, save_n=201)
    Ew, sv, zs, stats = ctx0.solve(cfg0.Ew0, opts)
    st, Ew_r, sv_r, zs_r, stats_r = orc.solve(cfg0.holder, cfg0.Ew0, opts)
    assert stats.status == 0 and stats.saved == 201
    assert stats.steps == stats_r.steps
    assert spec_err(Ew, Ew_r, cfg0.sidx) <= 1e-6
    assert spec_err(sv[-1], sv_r[-1], cfg0.sidx) <= 1e-6


def test_batch_equals_single(pkg, capi):
    print("case: batch of 4 mixed grids", flush=True)
    from luna_jl_amd import configs, runtime
    insts = [configs.capillary_instance(energy=e, pressure=p, trange=t, flength=0.1)
             for e, p, t in [(80e-6, 0.7, 0.5e-12), (150e-6, 2.5, 1e-12), (60e-6, 1.5, 1e-12), (190e-6, 2.9, 0.5e-12)]]
    ctxs = [runtime.Context(c.holder) for c in insts]
    opts = capi.rk_opts(0.1, save_n=6)
    outs, svs, zss, stats = runtime.batch_solve(ctxs, [c.Ew0 for c in insts], opts)
    for i, c in enumerate(insts):
        Ew1, sv1, zs1, st1 = ctxs[i].solve(c.Ew0, opts)
        assert stats[i].status == 0
        assert stats[i].steps == st1.steps
        assert np.array_equal(outs[i], Ew1)
        assert np.array_equal(svs[i], sv1)


def test_resident_batch_matches_host_batch(pkg, capi):
    print("case: resident batch", flush=True)
    from luna_jl_amd import configs, runtime
    insts = [configs.capillary_instance(energy=e, pressure=p, trange=t, flength=0.05)
             for e, p, t in [(120e-6, 1.2, 0.5e-12), (90e-6, 2.0, 1e-12)]]
    ctxs = [runtime.Context(c.holder) for c in insts]
    opts = capi.rk_opts(0.05, save_n=5)
    outs, svs, zss, st_host = runtime.batch_solve(ctxs, [c.Ew0 for c in insts], opts)
    rb = runtime.ResidentBatch(ctxs, [c.Ew0 for c in insts])
    for rep in range(2):  # inputs stay resident: a rerun gives the same answer
        st = rb.solve(opts)
        for i in range(2):
            assert st[i].status == 0 and st[i].steps == st_host[i].steps
            assert np.array_equal(rb.state(i), outs[i])
            sv, zs = rb.saves(i)
            assert np.array_equal(sv, svs[i])


def test_batch_into_pinned_host_buffers(pkg, capi):
    """batch_solve(..., out=HostBuffers): the same states, saves and save
    positions as with fresh host arrays, twice in a row into the same
    page-locked buffers."""
    print("case: batch into pinned host buffers", flush=True)
    from luna_jl_amd import configs, runtime
    insts = [configs.capillary_instance(energy=e, pressure=p, trange=t, flength=0.05)
             for e, p, t in [(120e-6, 1.2, 0.5e-12), (90e-6, 2.0, 1e-12), (150e-6, 0.8, 0.5e-12)]]
    ctxs = [runtime.Context(c.holder) for c in insts]
    opts = capi.rk_opts(0.05, save_n=5)
    outs, svs, zss, st_ref = runtime.batch_solve(ctxs, [c.Ew0 for c in insts], opts)
    hb = runtime.HostBuffers(ctxs, opts.save_n)
    try:
        for rep in range(2):
            o2, s2, z2, st2 = runtime.batch_solve(ctxs, [c.Ew0 for c in insts], opts, out=hb)
            for i in range(len(insts)):
                assert st2[i].status == 0 and st2[i].steps == st_ref[i].steps
                assert np.array_equal(o2[i], outs[i]) and np.array_equal(s2[i], svs[i])
                assert np.array_equal(z2[i], zss[i])
    finally:
        hb.close()
