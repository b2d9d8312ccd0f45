"""The C-ABI library loads and exports every entry point include/*.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


HIP_HEADERS = ("luna_hip.h", "luna_fno.h")   # libluna_hip.so
H5_HEADERS = ("luna_h5.h",)                  # libluna_h5.so (built where HDF5 is installed)
HEADERS = HIP_HEADERS + H5_HEADERS


def declared_symbols(headers=HIP_HEADERS):
    syms = set()
    for h in headers:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms |= set(re.findall(r"\b(lh_[a-z_0-9]+)\s*\(", text))
    return sorted(syms)


def test_every_header_is_listed():
    assert sorted(os.listdir(os.path.join(ROOT, "include"))) == sorted(HEADERS)


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("lh_create", "lh_destroy", "lh_rhs", "lh_rhs_device", "lh_solve", "lh_batch_solve",
              "lh_strerror", "lh_fno_create", "lh_fno_train", "lh_fno_forward"):
        assert s in syms


def test_library_exports_all_declared(capi):
    L = capi.lib()
    for s in declared_symbols():
        assert hasattr(L, s), s
    assert set(capi.EXPORTED) == set(declared_symbols())


def test_h5_library_exports_all_declared():
    from luna_jl_amd import output
    if not os.path.isfile(output.H5_LIB_PATH):
        pytest.skip("libluna_h5.so not built (no HDF5 library)")
    L = output.h5lib()
    for s in declared_symbols(H5_HEADERS):
        assert hasattr(L, s), s


def test_strerror_and_arg_checks(capi):
    L = capi.lib()
    assert L.lh_strerror(capi.LH_E_PPT_RANGE).decode().startswith("field strength")
    h = ctypes.c_void_p()
    assert L.lh_create(ctypes.byref(h), None) == capi.LH_E_ARG
    d = capi.LhDesc()
    d.grid_kind = 7  # unsupported
    assert L.lh_create(ctypes.byref(h), ctypes.byref(d)) == capi.LH_E_ARG


def test_struct_sizes_match_header(capi):
    # layout of the descriptor as a C compiler sees it
    import subprocess
    import tempfile
    src = '#include <stdio.h>\n#include <stddef.h>\n#include "luna_hip.h"\nint main(){printf("%zu %zu %zu %zu %zu %zu %zu %d",' \
          'sizeof(lh_desc), sizeof(lh_rk_opts), sizeof(lh_stats), offsetof(lh_desc, qdht_inv), offsetof(lh_desc, ppt_x),' \
          'sizeof(lh_stats_desc), offsetof(lh_stats_desc, aeff), (int)LH_NSTAT);}'
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(td, "s")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, c])
        out = subprocess.check_output([exe]).decode().split()
    sizes = [int(x) for x in out]
    assert sizes == [ctypes.sizeof(capi.LhDesc), ctypes.sizeof(capi.LhRkOpts), ctypes.sizeof(capi.LhStats),
                     capi.LhDesc.qdht_inv.offset, capi.LhDesc.ppt_x.offset,
                     ctypes.sizeof(capi.LhStatsDesc), capi.LhStatsDesc.aeff.offset, capi.LH_NSTAT]


def test_plan_struct_matches_header(capi):
    import subprocess
    import tempfile
    src = '#include <stdio.h>\n#include <stddef.h>\n#include "luna_hip.h"\nint main(){printf("%zu %zu %zu",' \
          'sizeof(lh_plan), offsetof(lh_plan, rhs_us), offsetof(lh_plan, agrees));}'
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(td, "s")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, c])
        out = [int(x) for x in subprocess.check_output([exe]).decode().split()]
    assert out == [ctypes.sizeof(capi.LhPlan), capi.LhPlan.rhs_us.offset, capi.LhPlan.agrees.offset]


def test_wisdom_text_round_trip(capi, pkg):
    """The wisdom is host state: import, export and forget need no GPU."""
    from luna_jl_amd import runtime
    L = capi.lib()
    runtime.wisdom_forget()
    assert runtime.wisdom_export() == "luna-hip-wisdom 1\n"
    good = "luna-hip-wisdom 1\nradial 8192 256 1 4 123.500\nradial 2048 64 0 1 7.250\n"
    runtime.wisdom_import(good)
    assert runtime.wisdom_export() == good
    # a later import of the same shape replaces its entry
    runtime.wisdom_import("luna-hip-wisdom 1\nradial 2048 64 1 2 6.000\n")
    assert runtime.wisdom_export() == "luna-hip-wisdom 1\nradial 8192 256 1 4 123.500\nradial 2048 64 1 2 6.000\n"
    # any malformed line refuses the whole text, merging nothing
    for bad in ["", "luna-hip-wisdom 2\n", "luna-hip-wisdom 1\nradial 8192 256 2 4 1.0\n",
                "luna-hip-wisdom 1\nradial 8192 256 0 3 1.0\n", "luna-hip-wisdom 1\nradial 1000 256 0 2 1.0\n",
                "luna-hip-wisdom 1\nradial 8192 100 0 2 1.0\n", "luna-hip-wisdom 1\nmodeavg 8192 1 0 2 1.0\n",
                "luna-hip-wisdom 1\nradial 4096 64 0 2 1.0\nradial 8192 256 0 2 1.0 extra\n"]:
        assert L.lh_wisdom_import(bad.encode()) == capi.LH_E_ARG, bad
    assert "radial 4096 64" not in runtime.wisdom_export()
    assert L.lh_wisdom_import(None) == capi.LH_E_ARG
    n = ctypes.c_size_t(0)
    assert L.lh_wisdom_export(None, 0, ctypes.byref(n)) == capi.LH_OK and n.value == len(runtime.wisdom_export()) + 1
    small = ctypes.create_string_buffer(4)
    assert L.lh_wisdom_export(small, 4, ctypes.byref(n)) == capi.LH_E_ARG
    runtime.wisdom_forget()
    assert runtime.wisdom_export() == "luna-hip-wisdom 1\n"
    assert L.lh_get_plan(None, None) == capi.LH_E_ARG and L.lh_set_plan(None, None) == capi.LH_E_ARG
    assert L.lh_autotune(None, None, 1, None, 0, None) == capi.LH_E_ARG


def test_planning_mode_mirror(pkg):
    from luna_jl_amd import runtime
    old = runtime.settings["fftw_flag"]
    try:
        assert runtime.set_fftw_mode(":patient") == "patient" and runtime.settings["fftw_flag"] == "patient"
        assert runtime.set_fftw_mode("ESTIMATE") == "estimate"
        with pytest.raises(ValueError):
            runtime.set_fftw_mode(":wild")
    finally:
        runtime.settings["fftw_flag"] = old


def test_stat_keys_follow_header_order(capi):
    import re
    text = open(os.path.join(ROOT, "include", "luna_hip.h")).read()
    names = re.findall(r"LH_STAT_([A-Z0-9_]+)\s*=\s*(\d+)", text)
    assert [int(v) for _, v in names] == list(range(capi.LH_NSTAT))
    expect = {"OMEGA0": "ω0"}
    assert [expect.get(n, n.lower()) for n, _ in names] == list(capi.STAT_KEYS)


def test_fno_structs_match_header(capi):
    import subprocess
    import tempfile
    src = '#include <stdio.h>\n#include "luna_fno.h"\nint main(){printf("%zu %zu", sizeof(lh_fno_cfg), sizeof(lh_adam));}'
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(td, "s")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, c])
        out = [int(x) for x in subprocess.check_output([exe]).decode().split()]
    assert out == [ctypes.sizeof(capi.LhFnoCfg), ctypes.sizeof(capi.LhAdam)]


def test_fno_arg_checks(capi):
    L = capi.lib()
    h = ctypes.c_void_p()
    cfg = capi.LhFnoCfg(n_x=100, cin=4, channels=64, modes=32, layers=4, hidden=128, max_batch=4)
    assert L.lh_fno_create(ctypes.byref(h), ctypes.byref(cfg)) == capi.LH_E_ARG  # n_x not a multiple of 64


def test_grids_beyond_largest_classes_rejected(capi, pkg):
    """An envelope grid of Nt = 32768 (trange 26 ps over 450-8000 nm) and a
    field grid of Nto = 65536 (trange 4 ps over 150 nm-4 μm) lie above the
    largest compiled classes (envelope Nt = 16384, field Nto = 32768);
    lh_create refuses them with LH_E_ARG before touching a device.  The Nt =
    16384 envelope grid of the soliton self-frequency-shift case
    (test/test_gnlse.jl:33-56) and the Nto = 32768 field grid of a 2 ps
    window are supported classes."""
    from luna_jl_amd import interface
    gamma, beta2, tau0, fr = 0.1, -1e-26, 10e-15, 0.18
    P0 = abs(beta2) / ((1 - fr) * gamma * tau0 ** 2)
    sizes = {}
    for trange in (13e-12, 26e-12):
        Ew, grid, linop, tr, _, _ = interface.prop_gnlse_args(
            gamma, 1.0, [0.0, 0.0, beta2], lambda0=835e-9, tau_fwhm=1.763 * tau0, power=P0,
            pulseshape="sech", lambda_lims=(450e-9, 8000e-9), trange=trange, raman=True, shock=False,
            fr=fr, shotnoise=False, ramanmodel="sdo", tau1=12.2e-15, tau2=32e-15)
        sizes[trange] = (len(grid.t), tr.descriptor(linop))
    assert sizes[13e-12][0] == 16384 and sizes[26e-12][0] == 32768
    h = ctypes.c_void_p()
    assert capi.lib().lh_create(ctypes.byref(h), sizes[26e-12][1].ref) == capi.LH_E_ARG
    Ew, grid, linop, tr, _, _ = interface.prop_capillary_args(
        125e-6, 1.0, "He", 1.0, lambda0=800e-9, lambda_lims=(150e-9, 4e-6), trange=4e-12,
        tau_fwhm=10e-15, energy=120e-6, shotnoise=False, ppt_N=2 ** 10)
    assert grid.n_to == 65536
    assert capi.lib().lh_create(ctypes.byref(h), tr.descriptor(linop).ref) == capi.LH_E_ARG


def test_host_alloc_argument_checks(capi):
    """lh_host_alloc refuses a zero size or a missing result pointer before
    touching a device; lh_host_free(NULL) is a no-op."""
    L = capi.lib()
    p = ctypes.c_void_p()
    assert L.lh_host_alloc(0, ctypes.byref(p)) == capi.LH_E_ARG and not p.value
    assert L.lh_host_alloc(64, None) == capi.LH_E_ARG
    assert L.lh_host_free(None) == capi.LH_OK


@pytest.mark.parametrize("header", HEADERS)
@pytest.mark.parametrize("lang", ["c99", "c11", "c++17"])
def test_header_compiles_on_its_own(header, lang):
    """Each public header compiles as the only include of a C or C++ file."""
    import subprocess
    import tempfile
    cpp = lang.startswith("c++")
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "h.cpp" if cpp else "h.c")
        open(src, "w").write(f'#include "{header}"\nint main(void) {{ return 0; }}\n')
        cc = "g++" if cpp else "gcc"
        res = subprocess.run([cc, f"-std={lang}", "-Wall", "-Wextra", "-Werror", "-pedantic",
                              "-I", os.path.join(ROOT, "include"), "-o", os.path.join(td, "h"), src],
                             capture_output=True, text=True)
        assert res.returncode == 0, res.stderr


def test_radial_envelope_limits(capi, pkg):
    """Envelope-grid radial contexts: Nt = Nto from 256 to 8192 points and Nr a
    multiple of 64; a window of 0.1 ps (Nt = 128) or Nr = 100 is refused with
    LH_E_ARG before a device is touched (a supported descriptor gets as far as
    the device query)."""
    from luna_jl_amd import configs
    h = ctypes.c_void_p()
    small = configs.radial_env_instance(N=64, trange=0.1e-12)
    assert small.grid.n_t == 128
    assert capi.lib().lh_create(ctypes.byref(h), small.holder.ref) == capi.LH_E_ARG
    odd = configs.radial_env_instance(N=100, trange=0.2e-12)
    assert capi.lib().lh_create(ctypes.byref(h), odd.holder.ref) == capi.LH_E_ARG
    ok = configs.radial_env_instance(N=64, trange=0.2e-12)
    st = capi.lib().lh_create(ctypes.byref(h), ok.holder.ref)
    if st == capi.LH_OK:
        capi.lib().lh_destroy(h)
    assert st != capi.LH_E_ARG
