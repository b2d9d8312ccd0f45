"""The allocation-free solve, lh_solve_noalloc, against lh_solve.

lh_solve_noalloc runs the same device solve as lh_solve, with the solver's
working memory in the library's fixed per-GPU buffer instead of buffers it
allocates, on device pointers the caller owns.  Both run the same kernel on the
same inputs, so the final state, every save, the save positions, the run
counters and the per-step statistics records must agree bit for bit.  Two
calls in a row on one GPU, the second smaller (a coarser grid class and fewer
saves), show that the shared buffer carries nothing from one call into the
next; an envelope-grid case with Raman puts the Raman stash in the buffer too.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """The caller's own device memory, from the HIP runtime the library uses
    (the soname it links; torch bundles a runtime of its own, whose pointers
    this process's other runtime cannot take)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, dev, a):
        assert self.L.hipMemcpy(dev, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice

    def get(self, a, dev):
        assert self.L.hipMemcpy(a.ctypes.data, dev, a.nbytes, 2) == 0  # hipMemcpyDeviceToHost

    def free(self, p):
        self.L.hipFree(p)


def _noalloc(capi, ctx, Ew0, opts, want_saves=True):
    hip = _Hip()
    n = Ew0.size
    host = np.ascontiguousarray(Ew0, dtype=np.complex128).copy()
    ew = hip.alloc(host.nbytes)
    hip.put(ew, host)
    sv = hip.alloc(16 * n * max(opts.save_n, 1)) if want_saves else None
    zs = np.zeros(max(opts.save_n, 1))
    st = capi.LhStats()
    try:
        rc = capi.lib().lh_solve_noalloc(ctx.handle, ew, C.byref(opts), sv, zs.ctypes.data_as(C.POINTER(C.c_double)),
                                         C.byref(st), None)
        out = np.empty_like(host)
        hip.get(out, ew)
        saves = None
        if want_saves:
            saves = np.empty((max(opts.save_n, 1), n), dtype=np.complex128)
            hip.get(saves, sv)
            saves = saves[:opts.save_n]
    finally:
        hip.free(ew)
        if sv is not None:
            hip.free(sv)
    return rc, out, saves, zs[:opts.save_n], st


def _same(capi, ctx, Ew0, opts, records=False):
    Ew_a, sv_a, zs_a, st_a = ctx.solve(Ew0, opts)
    rec_a = ctx.step_stats() if records else None
    rc, Ew_b, sv_b, zs_b, st_b = _noalloc(capi, ctx, Ew0, opts)
    assert rc == st_a.status == st_b.status == 0, (rc, st_a.status, st_b.status)
    assert (st_a.steps, st_a.accepted, st_a.rejected, st_a.rhs_evals, st_a.saved) == \
        (st_b.steps, st_b.accepted, st_b.rejected, st_b.rhs_evals, st_b.saved)
    assert st_a.z == st_b.z and st_a.dz == st_b.dz
    assert np.array_equal(Ew_a, Ew_b)
    assert np.array_equal(sv_a, sv_b)
    assert np.array_equal(zs_a, zs_b)
    if records:
        rec_b = ctx.step_stats()
        assert rec_a.shape == rec_b.shape and rec_a.shape[0] == st_a.accepted
        assert np.array_equal(rec_a, rec_b, equal_nan=True)
    return st_b


def test_noalloc_matches_solve_twice(pkg, capi):
    from luna_jl_amd import configs, runtime, stats
    big = configs.capillary_instance(energy=150e-6, pressure=2.0, trange=1e-12, flength=0.2)
    small = configs.capillary_instance(energy=90e-6, pressure=1.0, trange=0.5e-12, flength=0.1)
    print("case: noalloc Nto 16384, 41 saves, statistics on", flush=True)
    ctx_big = runtime.Context(big.holder)
    ctx_big.set_step_stats(stats.StatsFun(big.grid, big.transform, capacity=4096))
    st1 = _same(capi, ctx_big, big.Ew0, capi.rk_opts(big.zmax, save_n=41), records=True)
    print("case: noalloc Nto 8192, 5 saves (second, smaller call)", flush=True)
    ctx_small = runtime.Context(small.holder)
    st2 = _same(capi, ctx_small, small.Ew0, capi.rk_opts(small.zmax, save_n=5))
    assert st1.accepted > 0 and st2.accepted > 0


def test_noalloc_envelope_raman(pkg, capi):
    from luna_jl_amd import configs, runtime
    print("case: noalloc envelope Nt 8192 with Raman, 2 cm", flush=True)
    g = configs.gnlse_instance(flength=0.02)
    ctx = runtime.Context(g.holder)
    _same(capi, ctx, g.Ew0, capi.rk_opts(g.zmax, save_n=11))


def test_noalloc_rejects_oversize(pkg, capi):
    from luna_jl_amd import configs, runtime
    print("case: noalloc save count beyond the buffer", flush=True)
    inst = configs.capillary_instance(energy=90e-6, pressure=1.0, trange=0.5e-12, flength=0.02)
    ctx = runtime.Context(inst.holder)
    too_many = capi.LH_NOALLOC_SCRATCH_BYTES // 8 + 1
    opts = capi.rk_opts(inst.zmax, save_n=too_many)
    st = capi.LhStats()
    Ew = np.ascontiguousarray(inst.Ew0).copy()
    # rejected on the host before any device pointer is touched
    rc = capi.lib().lh_solve_noalloc(ctx.handle, C.c_void_p(Ew.ctypes.data), C.byref(opts), None, None,
                                     C.byref(st), None)
    assert rc == capi.LH_E_TOO_LARGE
    assert "noalloc" in capi.lib().lh_strerror(rc).decode()
    # the context is still usable
    _same(capi, ctx, inst.Ew0, capi.rk_opts(inst.zmax, save_n=3))
