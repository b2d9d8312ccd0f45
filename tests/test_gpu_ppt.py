"""GPU PPT table (lh_ppt_table) against the host restatement of
ionrate_fun_PPT (src/Ionisation.jl:317-390), itself pinned to the reference's
PPT golden values (test/test_ionisation.jl:15-16) in test_setup_physics.py.

He exercises φ_0 (Dawson's integral), Ar the m = ±1 terms (φ_1, including the
large-argument branch the reference integrates numerically).  Rates agree to
1e-11 relative wherever the host rate is above 1e-300, and the tables give the
same log-rate spline nodes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("gas,lam0", [("He", 800e-9), ("Ar", 800e-9), ("Ar", 1800e-9), ("Xe", 400e-9)])
def test_table_matches_host(gas, lam0, pkg):
    print(f"case: ppt table {gas} {lam0}", flush=True)
    from luna_jl_amd import ionisation
    E, r_host = ionisation.make_ppt_cache(gas, lam0, N=2 ** 12)
    E2, r_dev = ionisation.make_ppt_cache(gas, lam0, N=2 ** 12, device=True)
    assert np.array_equal(E, E2)
    m = r_host > 1e-300
    assert m.sum() > 1000
    assert np.all(np.isfinite(r_dev))
    rel = np.abs(r_dev[m] - r_host[m]) / r_host[m]
    assert rel.max() <= 1e-11, (rel.max(), E[m][np.argmax(rel)])
    assert np.all(r_dev[~m] < 1e-290)


def test_spline_from_device_table(pkg):
    print("case: ppt spline from device table", flush=True)
    from luna_jl_amd import ionisation
    a = ionisation.PPTSpline(*ionisation.make_ppt_cache("He", 800e-9, N=2 ** 12))
    b = ionisation.PPTSpline(*ionisation.make_ppt_cache("He", 800e-9, N=2 ** 12, device=True))
    assert np.array_equal(a.x, b.x)
    assert np.max(np.abs(a.y - b.y)) < 1e-10
