"""Size-independent properties of the GPU path at the full workload sizes
(Nto = 16384 and 8192 capillary grids, the Nt = 8192 GNLSE grid), and the
reference's own golden values computed by the GPU solver.

* Kerr-only RHS (Kerr_field, src/Nonlinear.jl:8-25, through TransModeAvg,
  src/NonlinearRHS.jl:362-373) is cubic in the field: every step of the RHS
  (c2r, E³, windows, r2c, output factor) scales exactly under a power-of-two
  scaling of its input, so nl(2E) = 8·nl(E) bit for bit; nl(3E) = 27·nl(E) to
  rounding.
* A zero field gives a zero RHS with plasma on (PlasmaCumtrapz with no rate:
  src/Nonlinear.jl:183-236).
* With no nonlinear response the solver reduces to the linear propagation
  exp(L z)·E0 (test/test_linearprop.jl:27-34 states it to 1e-2; here the
  windows are ≈ 1 where the field lives and the bound is set by rounding of
  the per-step exponentials: 3e-9, the oracle itself reaching 3.8e-10).

The reference's soliton self-frequency-shift golden values (test/test_gnlse.jl:
33-56) need an Nt = 16384 envelope grid, whose transforms are staged through
HBM; the GPU solver reproduces them in test_gpu_env.py.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("trange", [1e-12, 0.5e-12])
def test_kerr_rhs_is_cubic(pkg, trange):
    print(f"case: Kerr-only rhs homogeneity, trange {trange}", flush=True)
    from luna_jl_amd import configs, runtime
    cfg = configs.capillary_instance(energy=150e-6, pressure=2.0, trange=trange, plasma=False)
    ctx = runtime.Context(cfg.holder)
    nl1 = ctx.rhs(cfg.Ew0)
    nl2 = ctx.rhs(cfg.Ew0 * 2.0)
    nl3 = ctx.rhs(cfg.Ew0 * 3.0)
    assert np.linalg.norm(nl1) > 0
    assert np.array_equal(nl2, 8.0 * nl1)
    assert np.linalg.norm(nl3 - 27.0 * nl1) <= 1e-14 * np.linalg.norm(27.0 * nl1)


def test_zero_field_zero_rhs(pkg):
    print("case: zero field with plasma, Nto = 16384", flush=True)
    from luna_jl_amd import configs, runtime
    cfg = configs.capillary_instance(trange=1e-12)
    ctx = runtime.Context(cfg.holder)
    nl = ctx.rhs(np.zeros_like(cfg.Ew0))
    assert np.all(nl == 0)


@pytest.mark.parametrize("trange", [1e-12, 0.5e-12])
def test_linear_propagation_closed_form(pkg, capi, trange):
    print(f"case: linear propagation 1 m, trange {trange}", flush=True)
    from luna_jl_amd import configs, runtime
    cfg = configs.capillary_instance(energy=150e-6, pressure=2.0, trange=trange, plasma=False, kerr=False)
    ctx = runtime.Context(cfg.holder)
    opts = capi.rk_opts(cfg.zmax, save_n=11)
    Ew, sv, zs, stats = ctx.solve(cfg.Ew0, opts)
    assert stats.status == 0 and stats.saved == 11
    L = np.asarray(cfg.linop)
    s = cfg.sidx
    for k in (5, 10):
        ref = cfg.Ew0 * np.exp(L * zs[k])
        err = np.linalg.norm(sv[k][s] - ref[s]) / np.linalg.norm(ref[s])
        assert err <= 3e-9, (k, err)  # the oracle: 3.8e-10 (rounding of 22 step exponentials)
