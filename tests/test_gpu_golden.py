"""The GPU path against the committed golden fixtures (tests/golden/).

* cfg0_rhs.npz: TransModeAvg's RHS (src/NonlinearRHS.jl:323-373) of the
  single-point capillary instance at its input and at the state the oracle
  reached at z = 0.25 m: relative L2 ≤ 1e-12.  The device solve from the input
  to z = 0.25 m takes the fixture's step counts and lands within 1e-6 relative
  spectral-energy error of the fixture state.
* ppt_he_800nm.npz, ppt_ar_800nm.npz: the device PPT rate table at N = 2^16
  (makePPTcache, src/Ionisation.jl:185-200) within 1e-11 relative of the
  fixture wherever the rate is above 1e-300.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def g0(pkg):
    from luna_jl_amd import configs, runtime
    d = np.load(os.path.join(GOLD, "cfg0_rhs.npz"))
    p = json.loads(str(d["params"]))
    inst = configs.capillary_instance(energy=p["energy"], pressure=p["pressure"], trange=p["trange"],
                                      flength=p["flength"], gas=p["gas"], radius=p["radius"])
    return d, inst, runtime.Context(inst.holder)


def test_golden_rhs(g0):
    print("case: golden cfg0 rhs (input and z = 0.25 m state)", flush=True)
    d, inst, ctx = g0
    assert rel_l2(ctx.rhs(d["Ew0"]), d["nl0"]) <= 1e-12
    assert rel_l2(ctx.rhs(d["Ew1"], float(d["z1"])), d["nl1"]) <= 1e-12


def test_golden_propagation(g0, capi):
    print("case: golden cfg0 propagation to z = 0.25 m", flush=True)
    d, inst, ctx = g0
    E1, _, _, stats = ctx.solve(d["Ew0"], capi.rk_opts(float(d["z1"]), save_n=2), want_saves=False)
    assert stats.status == 0
    assert (stats.steps, stats.accepted, stats.rejected) == (int(d["steps"]), int(d["accepted"]),
                                                             int(d["rejected"]))
    s = inst.sidx
    pa, pb = np.abs(E1[s]) ** 2, np.abs(d["Ew1"][s]) ** 2
    assert np.linalg.norm(pa - pb) / np.linalg.norm(pb) <= 1e-6


@pytest.mark.parametrize("gas", ["He", "Ar"])
def test_golden_ppt_table(gas, pkg):
    print(f"case: golden ppt table {gas} 800 nm, N = 2^16", flush=True)
    from luna_jl_amd import ionisation
    d = np.load(os.path.join(GOLD, f"ppt_{gas.lower()}_800nm.npz"))
    E, rate = ionisation.make_ppt_cache(gas, 800e-9, N=int(d["N"]), device=True)
    assert E[0] == d["Emin"] and E[-1] == d["Emax"]
    m = d["rate"] > 1e-300
    assert np.all(np.isfinite(rate))
    assert np.max(np.abs(rate[m] - d["rate"][m]) / d["rate"][m]) <= 1e-11
    assert np.all(rate[~m] < 1e-290)
