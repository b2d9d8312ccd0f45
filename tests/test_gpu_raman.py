"""GPU parity of the field-grid Raman response (RamanPolarField, molecular gases).

prop_capillary in N2 adds RamanPolarField with N2's rotational and vibrational
response to Kerr and PPT plasma (src/Interface.jl:534-558,
src/Nonlinear.jl:183-247, 284-333).  On the device the E² convolution over
the zero-padded 2·Nto grid runs through HBM-staged transforms for every field
grid class.  Single RHS relative L2 ≤ 1e-12 against the oracle on the three
field grids (Nto = 4096, 8192, 16384); propagations with the same step counts
and ≤ 1e-6 relative spectral-energy error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def spec_err(a, b, sidx):
    pa, pb = np.abs(a[sidx]) ** 2, np.abs(b[sidx]) ** 2
    return np.linalg.norm(pa - pb) / np.linalg.norm(pb)


def _n2(trange, flength=0.1, energy=120e-6, pressure=2.0, kerr=True, plasma=True):
    from luna_jl_amd import interface
    return interface.prop_capillary_args(
        125e-6, flength, "N2", pressure, lambda0=800e-9, lambda_lims=(150e-9, 4e-6), trange=trange,
        tau_fwhm=10e-15, energy=energy, shotnoise=False, ppt_N=2 ** 12, kerr=kerr, plasma=plasma)


@pytest.mark.parametrize("trange,n_to", [(1e-12, 16384), (0.5e-12, 8192), (0.25e-12, 4096)])
def test_n2_rhs_with_raman(pkg, orc, trange, n_to):
    print(f"case: N2 rhs with Raman, Nto = {n_to}", flush=True)
    from luna_jl_amd import runtime
    from luna_jl_amd.nonlinear import RamanPolarField
    Ew, grid, linop, tr, _, _ = _n2(trange)
    assert grid.n_to == n_to and any(isinstance(r, RamanPolarField) for r in tr.responses)
    h = tr.descriptor(linop)
    ctx = runtime.Context(h)
    rng = np.random.default_rng(3)
    chirped = Ew * np.exp(1j * rng.uniform(0, 0.2, Ew.size))
    for E in (Ew, 1.6 * chirped):
        nl = ctx.rhs(E)
        st, ref = orc.rhs(h, E)
        assert st == 0
        assert rel_l2(nl, ref) <= 1e-12, rel_l2(nl, ref)


def test_n2_rhs_raman_only(pkg, orc):
    """The Raman polarisation alone (no Kerr, no plasma) at Nto = 16384."""
    print("case: N2 rhs, Raman only, Nto = 16384", flush=True)
    from luna_jl_amd import runtime
    Ew, grid, linop, tr, _, _ = _n2(1e-12, kerr=False, plasma=False)
    h = tr.descriptor(linop)
    ctx = runtime.Context(h)
    nl = ctx.rhs(3.0 * Ew)
    st, ref = orc.rhs(h, 3.0 * Ew)
    assert st == 0 and np.linalg.norm(ref) > 0
    assert rel_l2(nl, ref) <= 1e-12, rel_l2(nl, ref)


@pytest.mark.parametrize("trange", [0.5e-12, 1e-12])
def test_n2_propagation(pkg, orc, capi, trange):
    print(f"case: N2 propagation with Raman, trange {trange}", flush=True)
    from luna_jl_amd import runtime
    Ew, grid, linop, tr, _, _ = _n2(trange, flength=0.1, energy=150e-6, pressure=3.0)
    h = tr.descriptor(linop)
    ctx = runtime.Context(h)
    opts = capi.rk_opts(grid.zmax, save_n=11)
    Eg, sv, zs, stats = ctx.solve(Ew, opts)
    st, Er, sv_r, zs_r, stats_r = orc.solve(h, Ew, opts)
    assert stats.status == 0 and st == 0
    assert stats.steps == stats_r.steps and stats.accepted == stats_r.accepted
    assert spec_err(Eg, Er, grid.sidx) <= 1e-6
    for k in range(11):
        assert spec_err(sv[k], sv_r[k], grid.sidx) <= 1e-6
