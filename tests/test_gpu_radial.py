"""GPU parity of the radial transform (TransRadial, src/NonlinearRHS.jl:466-474)
and its propagation (Luna.run, src/Luna.jl:358-391) against the oracle.

Single RHS: relative L2 ≤ 1e-12 at every compiled size (Nto = 2048 to
16384), including the full Nr = 256 × Nto = 8192 workload (the FP64 MFMA contraction against the oracle's
plain triple loop).  Propagations: relative spectral-energy error ≤ 1e-6 at
rtol = 1e-6, and the reference's paraxial-beam check (test/test_radial.jl).
"""
import numpy as np
import pytest

from test_radial_propagation import radial_beam_check, reference_radial_case

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def spec_err(a, b):
    pa, pb = np.abs(a) ** 2, np.abs(b) ** 2
    return np.linalg.norm(pa - pb) / np.linalg.norm(pb)


@pytest.fixture(scope="module")
def cfg3(pkg):
    from luna_jl_amd import configs
    return configs.radial_instance()


@pytest.fixture(scope="module")
def ctx3(cfg3):
    from luna_jl_amd import runtime
    return runtime.Context(cfg3.holder)


def test_radial_rhs_full_size(cfg3, ctx3, orc):
    print("case: radial Nr=256 Nto=8192 rhs at z=0", flush=True)
    nl = ctx3.rhs(cfg3.Ew0)
    st, ref = orc.rhs(cfg3.holder, cfg3.Ew0)
    assert st == 0
    err = rel_l2(nl, ref)
    print("rel_l2", err, flush=True)
    assert err <= 1e-12


def test_radial_rhs_near_focus(cfg3, ctx3, orc):
    # the same beam 5 mm before its focus: ~1e14 W/cm², plasma active
    print("case: radial Nr=256 Nto=8192 rhs near focus", flush=True)
    from luna_jl_amd import fields
    q = cfg3.params["q"]
    Ew = fields.prop_radial(np.array(cfg3.Ew0, dtype=complex), 0.152, cfg3.grid, q)
    nl = ctx3.rhs(Ew)
    st, ref = orc.rhs(cfg3.holder, Ew)
    assert st == 0
    err = rel_l2(nl, ref)
    print("rel_l2", err, flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("trange,N", [(0.3e-12, 64), (0.5e-12, 128), (2e-12, 64)])
def test_radial_rhs_small_sizes(orc, trange, N):
    print(f"case: radial rhs trange={trange} N={N}", flush=True)
    from luna_jl_amd import configs, runtime
    inst = configs.radial_instance(flength=0.1, energy=50e-6, propz=-0.004, R=1.5e-3, N=N, trange=trange)
    ctx = runtime.Context(inst.holder)
    nl = ctx.rhs(inst.Ew0)
    st, ref = orc.rhs(inst.holder, inst.Ew0)
    assert st == 0
    assert rel_l2(nl, ref) <= 1e-12


def test_radial_diffraction_gpu(orc, capi):
    # the reference's test_radial.jl case on a 0.3 ps window (Nto = 2048)
    print("case: radial paraxial beam check", flush=True)
    from luna_jl_amd import runtime
    inst = reference_radial_case(0.3e-12)
    ctx = runtime.Context(inst.holder)
    opts = capi.rk_opts(inst.zmax, save_n=201)
    Ew, sv, zs, stats = ctx.solve(inst.Ew0, opts)
    assert stats.status == 0 and stats.saved == 201
    err = radial_beam_check(inst, sv[-1])
    print("beam error", err, "steps", stats.steps, flush=True)
    assert err < 1.5e-4
    st, Ew_r, sv_r, zs_r, stats_r = orc.solve(inst.holder, inst.Ew0, opts)
    assert (stats.steps, stats.accepted) == (stats_r.steps, stats_r.accepted)
    assert spec_err(Ew, Ew_r) <= 1e-9


def test_radial_nonlinear_propagation_vs_oracle(orc, capi):
    # 50 μJ focused to 40 μm in 0.5 bar Ar: Kerr and ionisation both act
    print("case: radial nonlinear 10 cm N=64 Nto=2048", flush=True)
    from luna_jl_amd import configs, runtime
    inst = configs.radial_instance(flength=0.1, energy=50e-6, propz=-0.05, R=1.5e-3, N=64, trange=0.3e-12)
    ctx = runtime.Context(inst.holder)
    opts = capi.rk_opts(inst.zmax, save_n=11)
    Ew, sv, zs, stats = ctx.solve(inst.Ew0, opts)
    st, Ew_r, sv_r, zs_r, stats_r = orc.solve(inst.holder, inst.Ew0, opts)
    print("steps gpu/oracle", stats.steps, stats_r.steps, flush=True)
    assert stats.status == 0 and stats_r.status == 0
    assert spec_err(Ew, Ew_r) <= 1e-6
    for k in range(11):
        assert spec_err(sv[k], sv_r[k]) <= 1e-6, k
    assert np.array_equal(zs, zs_r)


def test_radial_16384_propagation_vs_oracle(orc, capi):
    # the largest radial class: Nto = 16384 (a 2 ps window), Nr = 64, 2 cm before the focus
    print("case: radial 2 cm N=64 Nto=16384", flush=True)
    from luna_jl_amd import configs, runtime
    inst = configs.radial_instance(flength=0.02, energy=50e-6, propz=-0.01, R=1.5e-3, N=64, trange=2e-12)
    assert inst.grid.n_to == 16384
    ctx = runtime.Context(inst.holder)
    opts = capi.rk_opts(inst.zmax, save_n=5)
    Ew, sv, zs, stats = ctx.solve(inst.Ew0, opts)
    st, Ew_r, sv_r, zs_r, stats_r = orc.solve(inst.holder, inst.Ew0, opts)
    assert stats.status == 0 and st == 0
    assert (stats.steps, stats.accepted) == (stats_r.steps, stats_r.accepted)
    for k in range(5):
        assert spec_err(sv[k], sv_r[k]) <= 1e-6, k


def test_radial_full_workload(cfg3, ctx3, capi):
    # Nr = 256 × Nt = 4096, 0.5 m through the focus: runs to the end, saves 201
    print("case: radial full workload", flush=True)
    import time
    from luna_jl_amd import fields
    opts = capi.rk_opts(cfg3.zmax, save_n=201)
    t0 = time.perf_counter()
    Ew, sv, zs, stats = ctx3.solve(cfg3.Ew0, opts)
    el = time.perf_counter() - t0
    print("radial full:", el, "s steps", stats.steps, "rejected", stats.rejected, flush=True)
    assert stats.status == 0 and stats.saved == 201
    q, g = cfg3.params["q"], cfg3.grid
    e = [fields.energy_omega_radial(g, q, sv[k].reshape(q.N, g.n_omega).T) for k in (0, 100, 200)]
    print("energy", e, flush=True)
    assert np.all(np.isfinite(e)) and e[2] <= e[0] * (1 + 1e-6)
