"""Error paths of the device solver, against the oracle.

* Repeated rejection: RK45.solve raises after repeat_limit consecutive
  rejections (src/RK45.jl:71-77); the device returns LH_E_REPEAT_LIMIT after
  the same 11 attempts.
* The step-attempt cap (no counterpart in the reference) has its own status.
* A batch whose instance hits the PPT table's upper end mid-run
  (src/Ionisation.jl:262-265) reports LH_E_PPT_RANGE for that instance while
  the others finish bit-identical to their single runs, as Scans.runscan
  isolates each point's failure (src/Scans.jl:279-285).
* A run stopped after an accepted step (LH_E_STOPPED) and resumed from its
  state, z, dz and save count -- the cached continuation of HDF5Output
  (src/Output.jl:473-487, src/Luna.jl:374-379, test/test_output.jl:203-276).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def spec_err(a, b, sidx):
    pa, pb = np.abs(a[sidx]) ** 2, np.abs(b[sidx]) ** 2
    return np.linalg.norm(pa - pb) / np.linalg.norm(pb)


@pytest.fixture(scope="module")
def small(pkg):
    from luna_jl_amd import configs
    return configs.capillary_instance(energy=120e-6, pressure=1.5, trange=0.5e-12, flength=0.3)


def test_repeat_limit(small, orc, capi):
    print("case: repeat limit (rtol 1e-200)", flush=True)
    from luna_jl_amd import runtime
    ctx = runtime.Context(small.holder)
    opts = capi.rk_opts(0.05, rtol=1e-200, save_n=3)
    import ctypes as C
    Ew = np.ascontiguousarray(small.Ew0).copy()
    stats = capi.LhStats()
    rc = capi.lib().lh_solve(ctx.handle, Ew.ctypes.data_as(C.POINTER(C.c_double)), C.byref(opts), None, None,
                             C.byref(stats))
    assert rc == capi.LH_E_REPEAT_LIMIT and stats.status == capi.LH_E_REPEAT_LIMIT
    assert (stats.steps, stats.accepted, stats.rejected) == (11, 0, 11)
    r, Ew_r, sv_r, zs_r, st_r = orc.solve(small.holder, small.Ew0, opts, want_saves=False)
    assert r == capi.LH_E_REPEAT_LIMIT
    assert (st_r.steps, st_r.accepted, st_r.rejected) == (11, 0, 11)
    assert stats.dz == st_r.dz


def test_step_cap(small, orc, capi):
    print("case: step cap", flush=True)
    import ctypes as C
    from luna_jl_amd import runtime
    ctx = runtime.Context(small.holder)
    opts = capi.rk_opts(0.3, save_n=5, max_steps=7)
    Ew = np.ascontiguousarray(small.Ew0).copy()
    stats = capi.LhStats()
    rc = capi.lib().lh_solve(ctx.handle, Ew.ctypes.data_as(C.POINTER(C.c_double)), C.byref(opts), None, None,
                             C.byref(stats))
    assert rc == capi.LH_E_MAXSTEPS and stats.steps == 7
    assert capi.lib().lh_strerror(capi.LH_E_MAXSTEPS).decode() != capi.lib().lh_strerror(
        capi.LH_E_REPEAT_LIMIT).decode()
    r, Ew_r, _, _, st_r = orc.solve(small.holder, small.Ew0, opts, want_saves=False)
    assert r == capi.LH_E_MAXSTEPS and st_r.steps == 7 and st_r.accepted == stats.accepted
    # a mid-run state (no window filter has settled the far tails yet): well
    # inside the 1e-6 propagation bar
    assert spec_err(Ew, Ew_r, small.sidx) <= 1e-8


def _failing_instance(capi, pkg, gain=4.0, margin=1.3):
    """A capillary point whose field grows (a real gain added to the linear
    operator) past the end of a PPT table cut at `margin` times the initial
    peak field: the range error comes after several accepted steps."""
    from luna_jl_amd import configs, nonlinearrhs
    inst = configs.capillary_instance(energy=100e-6, pressure=1.0, trange=0.5e-12, flength=0.3)
    tr = inst.transform
    Et = np.fft.irfft(inst.Ew0, inst.grid.n_t) / (nonlinearrhs.NLSCALE * np.sqrt(tr.aeff))
    peak = np.max(np.abs(Et))
    spl = [r for r in tr.responses if type(r).__name__ == "PlasmaCumtrapz"][0].ratefunc
    m = int(np.searchsorted(spl.x, margin * peak)) + 1
    assert 8 < m < len(spl.x)
    holder = tr.descriptor(inst.linop + gain)
    x, y, d = (np.ascontiguousarray(a[:m]) for a in (spl.x, spl.y, spl.d))
    holder._keep += [x, y, d]
    holder.desc.ppt_n = m
    holder.desc.ppt_x = capi._ptr(x)
    holder.desc.ppt_y = capi._ptr(y)
    holder.desc.ppt_d = capi._ptr(d)
    return inst, holder


def test_batch_isolates_a_failing_instance(pkg, orc, capi):
    print("case: batch with one PPT-range failure mid-run", flush=True)
    from luna_jl_amd import configs, runtime
    bad_inst, bad_holder = _failing_instance(capi, pkg)
    good = [configs.capillary_instance(energy=e, pressure=p, trange=0.5e-12, flength=0.3)
            for e, p in [(80e-6, 0.7), (150e-6, 2.5)]]
    ctxs = [runtime.Context(good[0].holder), runtime.Context(bad_holder), runtime.Context(good[1].holder)]
    Ews = [good[0].Ew0, bad_inst.Ew0, good[1].Ew0]
    opts = capi.rk_opts(0.3, save_n=7)
    outs, svs, zss, stats = runtime.batch_solve(ctxs, Ews, opts)
    assert stats[1].status == capi.LH_E_PPT_RANGE
    assert stats[1].accepted >= 3, stats[1].accepted  # mid-run, not at the first evaluation
    r, Ew_r, _, _, st_r = orc.solve(bad_holder, bad_inst.Ew0, opts, want_saves=False)
    assert r == capi.LH_E_PPT_RANGE
    assert (stats[1].steps, stats[1].accepted) == (st_r.steps, st_r.accepted)
    for i in (0, 2):
        Ew1, sv1, zs1, st1 = ctxs[i].solve(Ews[i], opts)
        assert stats[i].status == 0 and st1.status == 0
        assert stats[i].steps == st1.steps
        assert np.array_equal(outs[i], Ew1) and np.array_equal(svs[i], sv1)


def test_stop_and_resume(pkg, orc, capi):
    """The configuration of test/test_output.jl:203-276 (13 μm Ar capillary at
    5 bar, no loss, Kerr only, 1 μJ 30 fs at 800 nm, 5 cm, 51 saves): stop after
    the first accepted step beyond 3 cm, resume from the returned state (a
    fresh stepper at that z and dz, saves continuing at the returned count),
    and compare with the uninterrupted run as that test does (|Eω|² of every
    save within 1e-7 over the ωwin == 1 band, same save positions), and each
    leg with the oracle's."""
    print("case: stop at 3 cm and resume", flush=True)
    import ctypes as C
    from luna_jl_amd import interface, runtime
    Ew0, grid, linop, tr, _, _ = interface.prop_capillary_args(
        13e-6, 5e-2, "Ar", 5.0, lambda0=800e-9, lambda_lims=(160e-9, 3000e-9), trange=1e-12,
        tau_fwhm=30e-15, energy=1e-6, shotnoise=False, plasma=False, loss=False)
    holder = tr.descriptor(linop)
    sidx = grid.sidx
    ctx = runtime.Context(holder)
    save_n = 51
    full = capi.rk_opts(0.05, save_n=save_n)
    Ew_f, sv_f, zs_f, st_f = ctx.solve(Ew0, full)
    assert st_f.status == 0
    # leg 1: interrupted
    o1 = capi.rk_opts(0.05, save_n=save_n, zstop=0.03)
    Ew1 = np.ascontiguousarray(Ew0).copy()
    sv1 = np.zeros((save_n, grid.n_omega), complex)
    zs1 = np.zeros(save_n)
    st1 = capi.LhStats()
    dp = C.POINTER(C.c_double)
    rc = capi.lib().lh_solve(ctx.handle, Ew1.ctypes.data_as(dp), C.byref(o1), sv1.ctypes.data_as(dp),
                             zs1.ctypes.data_as(dp), C.byref(st1))
    assert rc == capi.LH_E_STOPPED and st1.status == capi.LH_E_STOPPED
    assert 0.03 < st1.z < 0.05 and 0 < st1.saved < save_n
    r, Ew1_r, sv1_r, _, st1_r = orc.solve(holder, Ew0, o1)
    assert r == capi.LH_E_STOPPED and (st1_r.steps, st1_r.saved) == (st1.steps, st1.saved)
    assert spec_err(Ew1, Ew1_r, sidx) <= 1e-8
    # leg 2: resumed from the cache (y, t, dt, saved)
    o2 = capi.rk_opts(0.05, z0=st1.z, init_dz=st1.dz, save_n=save_n, save_first=st1.saved)
    Ew2, sv2, zs2, st2 = ctx.solve(Ew1, o2)
    assert st2.status == 0 and st2.saved == save_n
    r, Ew2_r, sv2_r, zs2_r, st2_r = orc.solve(holder, Ew1, o2)
    assert r == 0 and st2_r.steps == st2.steps
    assert spec_err(Ew2, Ew2_r, sidx) <= 1e-8
    k = st1.saved
    sv = np.vstack([sv1[:k], sv2[k:]])
    zs = np.concatenate([zs1[:k], zs2[k:]])
    assert np.array_equal(zs, zs_f)
    band = grid.omega_win == 1
    Iw, Iwm = np.abs(sv[:, band]) ** 2, np.abs(sv_f[:, band]) ** 2
    assert np.linalg.norm(Iw - Iwm) / np.linalg.norm(Iw) < 1e-7
    assert np.allclose(sv[:, band], sv_f[:, band], atol=1e-4 * np.max(np.abs(sv_f[:, band])), rtol=0)
