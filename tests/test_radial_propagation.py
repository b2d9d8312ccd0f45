"""Radial free-space propagation against the paraxial Gaussian-beam solution,
as the reference's test/test_radial.jl does (Ar 1.2 bar, Kerr + PPT plasma at
negligible energy, 40 μm waist focused half-way along 0.6 m, N = 128, R = 4 mm):
the normalised I(r) at ω0 after the run matches the Gaussian of waist
w1 = w0 √(1 + (L/2 · λ/(π w0²))²) to 1.5e-4 (max |ΔI| / ‖I‖).

Here the oracle runs it on the CPU; the GPU runs the same case in
test_gpu_radial.py.
"""
import numpy as np


def radial_beam_check(inst, saves_last):
    grid, q = inst.grid, inst.params["q"]
    Eout = saves_last.reshape(q.N, grid.n_omega).T           # (Nω, Nr) at z = L
    Er = q.inverse(Eout)
    Iwr = np.abs(Er) ** 2
    c = 299792458.0
    lam0 = 800e-9
    w0idx = int(np.argmin(np.abs(grid.omega - 2 * np.pi * c / lam0)))
    lam = 2 * np.pi * c / grid.omega[w0idx]
    w0, L = 40e-6, inst.zmax
    w1 = w0 * np.sqrt(1 + (L / 2 * lam / (np.pi * w0 ** 2)) ** 2)
    Ira = np.exp(-0.5 * (q.r / (w1 / 2)) ** 2) * (w0 / w1) ** 2
    Ir = Iwr[w0idx] / Iwr[w0idx].max()
    Ira = Ira / Ira.max()
    return float(np.max(np.abs(Ir - Ira) / np.linalg.norm(Ir)))


def reference_radial_case(trange):
    from luna_jl_amd import configs
    return configs.radial_instance(gas="Ar", pressure=1.2, flength=0.6, tau_fwhm=20e-15, energy=1e-12,
                                   w0=40e-6, propz=-0.3, R=4e-3, N=128, trange=trange)


def test_oracle_radial_diffraction(orc, capi):
    inst = reference_radial_case(0.2e-12)
    opts = capi.rk_opts(inst.zmax, save_n=201)
    st, Ew, saves, zs, stats = orc.solve(inst.holder, inst.Ew0, opts)
    assert st == 0 and stats.status == 0
    err = radial_beam_check(inst, saves[-1])
    assert err < 1.5e-4, err
