"""Writes the golden fixtures of this directory.

Run from the repository root (CPU only, about a minute):

    python tests/golden/make_golden.py

Each fixture holds inputs and expected outputs; the tests in ``tests/``
compare the host set-up, the oracle and the GPU path against them.

``cfg0_rhs.npz``
    The prop_capillary instance of the single-point workload (125 μm He
    capillary at 1 bar, 120 μJ, 10 fs at 800 nm, λ 150 nm–4 μm, 1 ps window:
    Nto = 16384, Nω = 4097; ``Interface.prop_capillary`` src/Interface.jl:
    345-401, Kerr + PPT plasma): the input spectrum, the state the oracle
    reaches at z = 0.25 m (with its step counts), and the nonlinear RHS of both
    (``TransModeAvg`` src/NonlinearRHS.jl:323-373).
``ppt_he_800nm.npz``, ``ppt_ar_800nm.npz``
    ``makePPTcache`` (src/Ionisation.jl:185-200) rate tables at 800 nm, N = 2^16,
    with the default PPT options (Stark shift and dipole correction), and the
    log-rate spline derivatives (``Maths.CSpline``, src/Maths.jl:721-848) of the
    N = 2^12 table.
``cfg1_shotnoise.npz``
    The shot-noise phases of the GNLSE supercontinuum workload (8192 uniform
    [0, 1) values, seed 0; they stand in for Julia's GLOBAL_RNG draws of
    src/Fields.jl:313-320) and the resulting input spectrum.
``scan_seed1.json``
    The 256 points (energy, pressure, time window, grid sizes) of the batched
    capillary sweep drawn from seed 1.
``qdht_256.npz``
    The QDHT of order 0 at N = 256 computed with mpmath at 40 significant
    digits from Guizar-Sicairos & Gutiérrez-Vega, J. Opt. Soc. Am. A 21, 53
    (2004): the first 257 zeros j_n of J0, J1(j_n)², and T_nm =
    (2/S) J0(j_n j_m/S)/J1(j_m)² with S = j_257, rounded to float64.
``reference_values.json``
    Values the reference's own tests hold, with their file:line.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402

CFG0 = dict(energy=120e-6, pressure=1.0, trange=1e-12, flength=1.0, gas="He", radius=125e-6)
Z1 = 0.25


def cfg0(orc, capi):
    from luna_jl_amd import configs
    inst = configs.capillary_instance(energy=CFG0["energy"], pressure=CFG0["pressure"], trange=CFG0["trange"],
                                      flength=CFG0["flength"], gas=CFG0["gas"], radius=CFG0["radius"])
    opts = capi.rk_opts(Z1, save_n=2)
    st, E1, _, zs, stats = orc.solve(inst.holder, inst.Ew0, opts)
    assert st == 0 and stats.status == 0
    st0, nl0 = orc.rhs(inst.holder, inst.Ew0, 0.0)
    st1, nl1 = orc.rhs(inst.holder, E1, Z1)
    assert st0 == 0 and st1 == 0
    np.savez(os.path.join(HERE, "cfg0_rhs.npz"), Ew0=inst.Ew0, nl0=nl0, Ew1=E1, nl1=nl1, z1=Z1,
             steps=stats.steps, accepted=stats.accepted, rejected=stats.rejected,
             params=json.dumps(CFG0))


def ppt():
    from luna_jl_amd import ionisation
    for gas in ("He", "Ar"):
        E, rate = ionisation.make_ppt_cache(gas, 800e-9, N=2 ** 16)
        spl = ionisation.PPTSpline(*ionisation.make_ppt_cache(gas, 800e-9, N=2 ** 12))
        np.savez(os.path.join(HERE, f"ppt_{gas.lower()}_800nm.npz"), Emin=E[0], Emax=E[-1], N=len(E),
                 rate=rate, spline_x=spl.x, spline_y=spl.y, spline_d=spl.d)


def shotnoise():
    from luna_jl_amd import configs
    inst = configs.gnlse_instance(noise_seed=0)
    u = np.random.default_rng(0).random(inst.grid.omega.size)
    np.savez(os.path.join(HERE, "cfg1_shotnoise.npz"), u=u, Ew0=inst.Ew0)


def scan():
    from luna_jl_amd import configs
    from luna_jl_amd.grid import RealGrid
    E, P, T = configs.scan_points(256, 1)
    pts = []
    for e, p, t in zip(E, P, T):
        g = RealGrid(1.0, 800e-9, (150e-9, 4e-6), float(t))
        pts.append(dict(energy=float(e), pressure=float(p), trange=float(t), n_to=int(g.n_to),
                        n_omega=int(g.n_omega)))
    with open(os.path.join(HERE, "scan_seed1.json"), "w") as f:
        json.dump(dict(seed=1, points=pts), f, indent=0)


def qdht():
    import mpmath
    mpmath.mp.dps = 40
    n = 256
    j = [mpmath.besseljzero(0, k) for k in range(1, n + 2)]
    S = j[n]
    J1sq = [mpmath.besselj(1, j[m]) ** 2 for m in range(n)]
    T = np.empty((n, n))
    for a in range(n):
        for b in range(a, n):
            v = mpmath.besselj(0, j[a] * j[b] / S)
            T[a, b] = float(2 * v / (S * J1sq[b]))
            T[b, a] = float(2 * v / (S * J1sq[a]))
    np.savez(os.path.join(HERE, "qdht_256.npz"), j=np.array([float(x) for x in j]),
             J1sq=np.array([float(x) for x in J1sq]), T=T)


def reference_values():
    vals = {
        "ssfs_peak_omega": {"value": 1.3975031072069625e15, "source": "test/test_gnlse.jl:51"},
        "ssfs_peak_t": {"value": 6.079566076444579e-12, "source": "test/test_gnlse.jl:55"},
    }
    with open(os.path.join(HERE, "reference_values.json"), "w") as f:
        json.dump(vals, f, indent=1)


def main():
    ge.load_package()
    from luna_jl_amd import _capi
    orc = ge.load_oracle().Oracle(_capi)
    cfg0(orc, _capi)
    ppt()
    shotnoise()
    scan()
    qdht()
    reference_values()


if __name__ == "__main__":
    main()
