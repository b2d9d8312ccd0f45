"""Every accepted step handed to the host (lh_solve_stream with every_step),
the reference's output call at each step (src/Luna.jl:366-372,
src/Output.jl:65-81), and streaming of radial propagations.

* Mode-averaged: one call per accepted step, each with that step's record;
  saves, final state and records equal lh_solve's bit for bit; the state of
  each call is the step's filtered state (its energy is the step's recorded
  energy, 1e-12); a callback that asks to stop gets no further call.
* Stats.default with userfuns through Luna.run: one value per accepted step,
  a userfun's energy of the handed state equals the device's record (1e-12),
  and the saves equal a run without them.
* Radial (TransRadial): the streamed saves and final state equal lh_solve's
  bit for bit, with and without every_step; collect_stats through Luna.run
  gives one value per accepted step (ω0 and peak power per k column), the
  first energy within 1e-3 of the input's (one step of propagation and
  ionisation loss), and saves equal to a run without statistics.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(pkg):
    from luna_jl_amd import configs
    return configs.capillary_instance(energy=150e-6, pressure=2.0, trange=0.5e-12, flength=0.2)


def test_every_step_stream(small, capi):
    print("case: every-step stream vs lh_solve, Nto = 8192, 0.2 m", flush=True)
    from luna_jl_amd import runtime, stats
    ctx = runtime.Context(small.holder)
    sf = stats.StatsFun(small.grid, small.transform, capacity=4096)
    ctx.set_step_stats(sf)
    opts = capi.rk_opts(small.zmax, save_n=21)
    Ew_a, sv_a, zs_a, st_a = ctx.solve(small.Ew0, opts)
    rec_a = ctx.step_stats()
    calls = []

    def on_step(first, saves, zs, z, dz, y, rec):
        calls.append((first, saves.copy(), zs.copy(), z, dz, y.copy(), rec.copy()))
        return False

    opts_e = capi.rk_opts(small.zmax, save_n=21, every_step=True)
    Ew_b, st_b = ctx.solve_stream(small.Ew0, opts_e, on_step)
    assert st_a.status == st_b.status == 0
    assert (st_a.steps, st_a.accepted, st_a.saved) == (st_b.steps, st_b.accepted, st_b.saved)
    assert np.array_equal(Ew_a, Ew_b)
    assert len(calls) == st_a.accepted
    saves = np.concatenate([c[1] for c in calls], axis=1).T
    assert np.array_equal(saves, sv_a) and np.array_equal(np.concatenate([c[2] for c in calls]), zs_a[:21])
    recs = np.concatenate([c[6] for c in calls])
    assert np.array_equal(recs, rec_a, equal_nan=True)
    kz, ke = capi.STAT_KEYS.index("z"), capi.STAT_KEYS.index("energy")
    for first, _, _, z, dz, y, rec in calls:
        assert rec.shape == (1, capi.LH_NSTAT) and rec[0, kz] == z
        e = np.sum(sf.energy_w * np.abs(y) ** 2)
        assert abs(e - rec[0, ke]) <= 1e-12 * rec[0, ke]
    zc = [c[3] for c in calls]
    assert all(b > a for a, b in zip(zc, zc[1:]))

    n = []

    def stop_after_three(*a):
        n.append(1)
        return len(n) == 3

    _, st_c = ctx.solve_stream(small.Ew0, opts_e, stop_after_three)
    assert st_c.status == capi.LH_E_STOPPED and len(n) == 3


def test_run_with_userfuns(small, pkg):
    print("case: Luna.run with Stats.default userfuns", flush=True)
    from luna_jl_amd import interface, stats
    from luna_jl_amd.fields import energyfuncs
    from luna_jl_amd.output import MemoryOutput
    g = small.grid
    _, ew = energyfuncs(g)
    uf = lambda d, Ew, Et, z, dz: d.__setitem__("energy_host", ew(Ew))  # noqa: E731
    sf = stats.default(g, small.Ew0, None, small.linop, small.transform, userfuns=[uf], capacity=4096)
    out = MemoryOutput(0, g.zmax, 21, sf)
    _, st = interface.run(small.Ew0, g, small.linop, small.transform, None, out)
    ref = MemoryOutput(0, g.zmax, 21, stats.default(g, small.Ew0, None, small.linop, small.transform,
                                                    capacity=4096))
    _, st_r = interface.run(small.Ew0, g, small.linop, small.transform, None, ref)
    assert st.accepted == st_r.accepted
    s = out.data["stats"]
    assert len(s["energy_host"]) == len(s["energy"]) == st.accepted
    assert np.max(np.abs(s["energy_host"] - s["energy"]) / s["energy"]) <= 1e-12
    assert np.array_equal(out.data["Eω"], ref.data["Eω"])
    assert np.array_equal(s["energy"], ref.data["stats"]["energy"])


@pytest.fixture(scope="module")
def rad(pkg):
    from luna_jl_amd import configs
    return configs.radial_instance(flength=0.1, energy=50e-6, propz=-0.05, R=1.5e-3, N=64, trange=0.3e-12)


def test_radial_stream(rad, capi):
    print("case: radial stream vs lh_solve, N = 64, Nto = 2048", flush=True)
    from luna_jl_amd import runtime
    ctx = runtime.Context(rad.holder)
    opts = capi.rk_opts(rad.zmax, save_n=11)
    Ew_a, sv_a, zs_a, st_a = ctx.solve(rad.Ew0, opts)
    for every in (False, True):
        got = {"sv": [], "calls": 0}

        def on_step(first, saves, zs, z, dz, y, rec):
            got["sv"].append(saves.copy())
            got["calls"] += 1
            assert rec.shape[0] == 0 and y.shape == (ctx.n_state,)
            return False

        Ew_b, st_b = ctx.solve_stream(rad.Ew0, capi.rk_opts(rad.zmax, save_n=11, every_step=every), on_step)
        assert st_b.status == 0 and (st_a.steps, st_a.accepted) == (st_b.steps, st_b.accepted)
        assert np.array_equal(Ew_a, Ew_b)
        assert np.array_equal(np.concatenate(got["sv"], axis=1).T, sv_a)
        if every:
            assert got["calls"] == st_a.accepted
        else:
            assert 2 <= got["calls"] <= 11


def test_radial_run_collect_stats(rad, capi):
    print("case: radial Luna.run with collect_stats", flush=True)
    from luna_jl_amd import interface, stats
    from luna_jl_amd.fields import energyfuncs
    from luna_jl_amd.output import MemoryOutput
    g, q = rad.grid, rad.params["q"]
    _, ew = energyfuncs(g, q)
    tot = lambda d, Ew, Et, z, dz: d.__setitem__("energy_total", ew(Ew))  # noqa: E731
    hs = stats.collect_stats(g, rad.Ew0, stats.omega0(g), stats.peakpower(g), tot)
    out = MemoryOutput(0, g.zmax, 11, hs)
    Ew, st = interface.run(rad.Ew0, g, rad.linop, rad.transform, None, out)
    assert st.status == 0
    s = out.data["stats"]
    assert len(s["z"]) == len(s["energy_total"]) == st.accepted
    assert s["ω0"].shape == (q.N, st.accepted) and s["peakpower"].shape == (q.N, st.accepted)
    e0 = ew(np.asarray(rad.Ew0))
    assert abs(s["energy_total"][0] - e0) <= 1e-3 * e0
    plain = MemoryOutput(0, g.zmax, 11)
    interface.run(rad.Ew0, g, rad.linop, rad.transform, None, plain)
    assert np.array_equal(out.data["Eω"], plain.data["Eω"])


def test_run_grows_statistics(small, pkg):
    """Luna.run keeps a record of every accepted step when the run outgrows
    the device capacity (the reference grows its arrays, src/Stats.jl:495-532):
    the records and saves equal those of a run with room for all of them."""
    print("case: Luna.run statistics beyond the capacity", flush=True)
    from luna_jl_amd import interface, stats
    from luna_jl_amd.output import MemoryOutput
    g = small.grid
    sf = stats.default(g, small.Ew0, None, small.linop, small.transform, capacity=8)
    out = MemoryOutput(0, g.zmax, 21, sf)
    _, st = interface.run(small.Ew0, g, small.linop, small.transform, None, out)
    ref = MemoryOutput(0, g.zmax, 21, stats.default(g, small.Ew0, None, small.linop, small.transform,
                                                    capacity=4096))
    _, st_r = interface.run(small.Ew0, g, small.linop, small.transform, None, ref)
    assert st.accepted == st_r.accepted > 8
    assert sf.capacity == st.accepted
    s, r = out.data["stats"], ref.data["stats"]
    assert len(s["energy"]) == st.accepted
    for k in ("z", "energy", "peakpower", "ω0"):
        assert np.array_equal(s[k], r[k], equal_nan=True), k
    assert np.array_equal(out.data["Eω"], ref.data["Eω"])
