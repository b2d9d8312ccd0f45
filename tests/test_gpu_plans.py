"""Plans of the radial RHS on the GPU (lh_autotune, lh_set_plan).

Every plan -- radix plan of the column transforms × FP64 MFMA tiles per wave
of the Hankel GEMMs -- is held to the single-RHS bar against the oracle
(relative L2 ≤ 1e-12) at the full Nr = 256 × Nto = 8192 size and at
Nto = 2048; a propagation under the non-default plans matches the oracle's
(≤ 1e-6 spectral error, same step counts), which also checks that a plan
change re-captures the step graph.  The autotuner's choice is one of the
agreeing plans and is recorded in the wisdom.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PLANS = [(f, t) for f in (0, 1) for t in (1, 2, 4)]


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def spec_err(a, b):
    pa, pb = np.abs(a) ** 2, np.abs(b) ** 2
    return np.linalg.norm(pa - pb) / np.linalg.norm(pb)


@pytest.fixture(scope="module")
def big(pkg, orc):
    from luna_jl_amd import configs, fields, runtime
    cfg = configs.radial_instance()
    Ew = fields.prop_radial(np.array(cfg.Ew0, dtype=complex), 0.152, cfg.grid, cfg.params["q"])
    st, ref = orc.rhs(cfg.holder, Ew)
    assert st == 0
    return cfg, runtime.Context(cfg.holder), Ew, ref


@pytest.mark.parametrize("plan", PLANS)
def test_every_plan_full_size_rhs(big, plan):
    print(f"case: radial plan {plan} rhs Nr=256 Nto=8192", flush=True)
    cfg, ctx, Ew, ref = big
    ctx.plan = plan
    assert ctx.plan == plan
    err = rel_l2(ctx.rhs(Ew), ref)
    print("rel_l2", err, flush=True)
    assert err <= 1e-12


def test_autotune_full_size(big):
    print("case: radial autotune Nr=256 Nto=8192", flush=True)
    from luna_jl_amd import runtime
    cfg, ctx, Ew, ref = big
    runtime.wisdom_forget()
    tried = ctx.autotune(Ew, reps=5)
    print("tried", tried, flush=True)
    assert [(f, t) for f, t, _, _ in tried][0] == (0, 2) and sorted((f, t) for f, t, _, _ in tried) == PLANS
    assert all(ok for *_, ok in tried) and all(us > 0 for _, _, us, _ in tried)
    best = min(tried, key=lambda r: r[2])
    assert ctx.plan == (best[0], best[1])
    assert f"radial 8192 256 {best[0]} {best[1]} " in runtime.wisdom_export()
    assert rel_l2(ctx.rhs(Ew), ref) <= 1e-12
    # a context created for this shape afterwards takes the recorded plan
    ctx2 = runtime.Context(cfg.holder)
    assert ctx2.plan == ctx.plan
    runtime.wisdom_forget()
    assert runtime.Context(cfg.holder).plan == (0, 2)


@pytest.mark.parametrize("plan", [(1, 1), (1, 4), (0, 4)])
def test_plan_propagation_vs_oracle(orc, capi, plan):
    print(f"case: radial propagation under plan {plan}, N=64 Nto=2048", flush=True)
    from luna_jl_amd import configs, runtime
    inst = configs.radial_instance(flength=0.1, energy=50e-6, propz=-0.05, R=1.5e-3, N=64, trange=0.3e-12)
    ctx = runtime.Context(inst.holder)
    opts = capi.rk_opts(inst.zmax, save_n=11)
    Ew0, _, _, st0 = ctx.solve(inst.Ew0, opts)  # graph captured under the default plan
    ctx.plan = plan
    Ew, sv, zs, stats = ctx.solve(inst.Ew0, opts)
    st, Ew_r, sv_r, zs_r, stats_r = orc.solve(inst.holder, inst.Ew0, opts)
    assert stats.status == 0 and st == 0
    assert (stats.steps, stats.accepted) == (stats_r.steps, stats_r.accepted)
    assert spec_err(Ew, Ew_r) <= 1e-6
    for k in range(11):
        assert spec_err(sv[k], sv_r[k]) <= 1e-6, k
    nl = ctx.rhs(inst.Ew0)
    _, ref = orc.rhs(inst.holder, inst.Ew0)
    assert rel_l2(nl, ref) <= 1e-12


def test_plan_errors(pkg):
    print("case: plan argument errors", flush=True)
    from luna_jl_amd import _capi, configs, runtime
    inst = configs.radial_instance(flength=0.1, energy=50e-6, propz=-0.05, R=1.5e-3, N=64, trange=0.3e-12)
    ctx = runtime.Context(inst.holder)
    for bad in [(2, 2), (0, 3), (-1, 1), (0, 8)]:
        with pytest.raises(_capi.LunaHipError):
            ctx.plan = bad
    assert ctx.plan == (0, 2)
    cap = configs.capillary_instance(trange=0.25e-12)
    cctx = runtime.Context(cap.holder)
    with pytest.raises(_capi.LunaHipError):
        cctx.plan
    with pytest.raises(_capi.LunaHipError):
        cctx.autotune()


def test_measure_mode_plans_new_shapes(pkg, tmp_path, monkeypatch):
    """set_fftw_mode(:measure): a shape without wisdom is measured when its
    context is created and the wisdom file written; a second context (and a
    process that loads the file) takes the recorded plan."""
    print("case: measure mode and the wisdom file", flush=True)
    from luna_jl_amd import configs, runtime
    monkeypatch.setenv("LUNA_CACHE_DIR", str(tmp_path))
    runtime.wisdom_forget()
    old = runtime.settings["fftw_flag"]
    try:
        runtime.set_fftw_mode(":measure")
        inst = configs.radial_instance(flength=0.1, energy=50e-6, propz=-0.05, R=1.5e-3, N=64, trange=0.3e-12)
        ctx = runtime.Context(inst.holder)
        text = (tmp_path / "plan_wisdom").read_text()
        assert text.startswith("luna-hip-wisdom 1\nradial 2048 64 ")
        plan = ctx.plan
        runtime.wisdom_forget()
        ctx2 = runtime.Context(inst.holder)  # loads the file before lh_create
        assert ctx2.plan == plan
    finally:
        runtime.settings["fftw_flag"] = old
        runtime.wisdom_forget()
