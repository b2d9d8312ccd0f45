"""GPU parity of TransRadial on the envelope grid (src/NonlinearRHS.jl:457-459,
466-474, Kerr_env) and of its propagation against the oracle (itself pinned in
test_oracle_radial_env.py), on the reference's free-space envelope example
(examples/low_level_interface/freespace/radial_env.jl: Ar at 1.2 bar, 20 fs at
800 nm, 40 μm waist 0.3 m after the input) with a 1 ps window.

* One RHS: relative L2 ≤ 1e-12 at Nto = 1024 (Nr = 128), 2048 (Nr = 64) and
  the example's own grid (0.2 ps: Nt = Nto = 256, Nr = 512), at the input and
  at the beam 1 cm before its focus (20 μJ: Kerr active).
* The example's grid propagated over 0.4 m at 20 μJ with Nr = 64: the
  oracle's step counts, final state within 1e-6.
* Propagation over 0.4 m at 20 μJ (19 % away from the linear propagation):
  the oracle's step counts, final state and saves within 1e-6 relative
  spectral-energy error.
* Streaming (lh_solve_stream) gives lh_solve's saves bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def spec_err(a, b):
    pa, pb = np.abs(a) ** 2, np.abs(b) ** 2
    return np.linalg.norm(pa - pb) / np.linalg.norm(pb)


@pytest.mark.parametrize("trange,N", [(1e-12, 128), (2e-12, 64), (0.2e-12, 512)])
def test_radial_env_rhs(pkg, orc, trange, N):
    from luna_jl_amd import configs, fields, runtime
    inst = configs.radial_env_instance(N=N, trange=trange, energy=20e-6)
    print(f"case: radial envelope rhs Nto={inst.grid.n_to} Nr={N}", flush=True)
    ctx = runtime.Context(inst.holder)
    q = inst.params["q"]
    near = fields.prop_radial(np.array(inst.Ew0, dtype=complex), 0.29, inst.grid, q)
    for Ew in (inst.Ew0, near):
        nl = ctx.rhs(Ew)
        st, ref = orc.rhs(inst.holder, Ew)
        assert st == 0
        err = rel_l2(nl, ref)
        print("rel_l2", err, flush=True)
        assert err <= 1e-12


@pytest.fixture(scope="module")
def prop_case(pkg, orc, capi):
    from luna_jl_amd import configs
    inst = configs.radial_env_instance(N=64, energy=20e-6, flength=0.4)
    opts = capi.rk_opts(0.4, save_n=5)
    st, E1, sv, zs, stats = orc.solve(inst.holder, inst.Ew0, opts)
    assert st == 0
    return inst, opts, (E1, sv, zs, stats)


def test_radial_env_propagation(prop_case):
    print("case: radial envelope propagation Nto=1024 Nr=64, 0.4 m", flush=True)
    from luna_jl_amd import runtime
    inst, opts, (E1, sv, zs, st_o) = prop_case
    ctx = runtime.Context(inst.holder)
    Eg, svg, zsg, st = ctx.solve(inst.Ew0, opts)
    assert st.status == 0
    assert (st.steps, st.accepted, st.rejected) == (st_o.steps, st_o.accepted, st_o.rejected)
    assert np.array_equal(zsg, zs)
    assert spec_err(Eg, E1) <= 1e-6
    for k in range(1, len(zs)):
        assert spec_err(svg[k], sv[k]) <= 1e-6, k


def test_radial_env_stream(prop_case, capi):
    print("case: radial envelope stream vs lh_solve", flush=True)
    from luna_jl_amd import runtime
    inst, opts, _ = prop_case
    ctx = runtime.Context(inst.holder)
    Ea, sva, _, sta = ctx.solve(inst.Ew0, opts)
    got = []
    Eb, stb = ctx.solve_stream(inst.Ew0, opts, lambda first, saves, zs, z, dz, y, rec: got.append(saves.copy()))
    assert stb.status == 0 and stb.accepted == sta.accepted
    assert np.array_equal(Ea, Eb)
    assert np.array_equal(np.concatenate(got, axis=1).T, sva)


def test_radial_env_example_grid_propagation(pkg, orc, capi):
    print("case: radial envelope propagation Nto=256 Nr=64, 0.4 m", flush=True)
    from luna_jl_amd import configs, runtime
    inst = configs.radial_env_instance(N=64, trange=0.2e-12, energy=20e-6, flength=0.4)
    assert inst.grid.n_to == 256
    opts = capi.rk_opts(0.4, save_n=3)
    st, E1, _, _, st_o = orc.solve(inst.holder, inst.Ew0, opts, want_saves=False)
    assert st == 0
    Eg, _, _, stg = runtime.Context(inst.holder).solve(inst.Ew0, opts, want_saves=False)
    assert stg.status == 0 and (stg.steps, stg.accepted, stg.rejected) == (st_o.steps, st_o.accepted, st_o.rejected)
    assert spec_err(Eg, E1) <= 1e-6
