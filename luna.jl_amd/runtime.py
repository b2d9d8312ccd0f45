"""Host runtime over the C ABI: device contexts, single and batched solves."""
import ctypes as C

import numpy as np

from . import _capi


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """Owns one ``lh_ctx`` (device copies of a transform's constants)."""

    def __init__(self, holder):
        self.holder = holder
        d = holder.desc
        self.n_omega, self.n_r = d.n_omega, d.n_r
        self.n_state = d.n_omega * d.n_r
        L = _capi.lib()
        radial = d.transform == _capi.LH_TRANS_RADIAL
        if radial and settings["fftw_flag"] in _REPS:
            load_wisdom()  # before lh_create, which takes a recorded plan for its shape
        h = C.c_void_p()
        _capi.check(L.lh_create(C.byref(h), holder.ref), "lh_create")
        self.handle = h
        grad = getattr(holder, "gradient", None)
        if grad is not None:  # z-dependent fill (pressure gradient)
            _capi.check(L.lh_set_gradient(h, grad.ref), "lh_set_gradient")
        self.statsfun = None
        if radial:
            plan_radial(self)

    # -- plans of the radial RHS (lh_get_plan / lh_set_plan / lh_autotune) ---
    @property
    def plan(self):
        """(fft_radix_plan, gemm_tiles) of a radial context."""
        p = _capi.LhPlan()
        _capi.check(_capi.lib().lh_get_plan(self.handle, C.byref(p)), "lh_get_plan")
        return p.fft_radix_plan, p.gemm_tiles

    @plan.setter
    def plan(self, value):
        p = _capi.LhPlan()
        p.fft_radix_plan, p.gemm_tiles = value
        _capi.check(_capi.lib().lh_set_plan(self.handle, C.byref(p)), "lh_set_plan")

    def autotune(self, Ew=None, reps=10):
        """Times every plan on one RHS of ``Ew`` (zero field if None) and
        installs the fastest; returns [(fft_radix_plan, gemm_tiles, µs per RHS,
        agrees), ...] with the default plan first."""
        ptr = None
        if Ew is not None:
            Ew = np.ascontiguousarray(np.asarray(Ew, dtype=np.complex128).reshape(-1, order="F"))
            if Ew.size != self.n_state:
                raise ValueError("Eω has the wrong size")
            ptr = _p(Ew)
        tried = (_capi.LhPlan * 8)()
        n = C.c_int(0)
        _capi.check(_capi.lib().lh_autotune(self.handle, ptr, int(reps), tried, 8, C.byref(n)), "lh_autotune")
        return [(t.fft_radix_plan, t.gemm_tiles, t.rhs_us, bool(t.agrees)) for t in tried[:n.value]]

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _capi.lib().lh_destroy(h)
            except Exception:
                pass
            self.handle = None

    def rhs(self, Ew, z=0.0):
        Ew = np.ascontiguousarray(np.asarray(Ew, dtype=np.complex128).reshape(-1, order="F"))
        if Ew.size != self.n_state:
            raise ValueError("Eω has the wrong size")
        nl = np.empty_like(Ew)
        _capi.check(_capi.lib().lh_rhs(self.handle, _p(nl), _p(Ew), z), "lh_rhs")
        return nl

    def solve(self, Ew0, opts, want_saves=True):
        Ew = np.ascontiguousarray(np.asarray(Ew0, dtype=np.complex128).reshape(-1, order="F")).copy()
        saves = np.zeros((opts.save_n, self.n_state), dtype=np.complex128) if want_saves else None
        zs = np.zeros(max(opts.save_n, 1))
        stats = _capi.LhStats()
        st = _capi.lib().lh_solve(self.handle, _p(Ew), C.byref(opts),
                                  _p(saves) if want_saves else None, _p(zs), C.byref(stats))
        _capi.check(st, "lh_solve")
        return Ew, saves, zs, stats


    def solve_stream(self, Ew0, opts, on_step):
        """``lh_solve_stream``: the device solve with every accepted step that
        saves (every accepted step with ``opts.every_step``) handed to
        ``on_step(first_save, saves, zs, z, dz, y, records)`` while the kernel
        runs (saves: (n_state, n) column per save; y: the filtered state at z;
        records: [n][LH_NSTAT] statistics of the steps since the previous
        call).  ``on_step`` returning True stops the run.  An exception raised
        by ``on_step`` stops it too and is re-raised."""
        Ew = np.ascontiguousarray(np.asarray(Ew0, dtype=np.complex128).reshape(-1, order="F")).copy()
        ns = self.n_state
        failure = []

        def cb(user, out):
            try:
                o = out.contents
                if o.n_saves > 0:
                    sv = np.ctypeslib.as_array(o.saves, shape=(o.n_saves * ns * 2,)).view(np.complex128)
                    saves = sv.reshape(o.n_saves, ns).T.copy()
                    zs = np.ctypeslib.as_array(o.zsaves, shape=(o.n_saves,)).copy()
                else:  # an accepted step that saves nothing (every_step)
                    saves = np.zeros((ns, 0), dtype=np.complex128)
                    zs = np.zeros(0)
                y = np.ctypeslib.as_array(o.y, shape=(2 * ns,)).view(np.complex128).copy()
                if o.n_records > 0:
                    rec = np.ctypeslib.as_array(o.records, shape=(o.n_records * _capi.LH_NSTAT,))
                    rec = rec.reshape(o.n_records, _capi.LH_NSTAT).copy()
                else:
                    rec = np.zeros((0, _capi.LH_NSTAT))
                return 1 if on_step(o.first_save, saves, zs, o.z, o.dz, y, rec) else 0
            except BaseException as e:  # noqa: BLE001 -- handed back to the caller below
                failure.append(e)
                return 1

        stats = _capi.LhStats()
        st = _capi.lib().lh_solve_stream(self.handle, _p(Ew), C.byref(opts), _capi.LhOutputCb(cb), None,
                                         C.byref(stats))
        if failure:
            raise failure[0]
        if st not in (_capi.LH_OK, _capi.LH_E_STOPPED):
            _capi.check(st, "lh_solve_stream")
        return Ew, stats

    # -- per-step statistics (Stats.default, src/Stats.jl:476-532) ------------
    def set_step_stats(self, statsfun):
        """Record ``statsfun``'s values at every accepted step of later solves
        (None: off)."""
        L = _capi.lib()
        if statsfun is None:
            _capi.check(L.lh_set_step_stats(self.handle, None), "lh_set_step_stats")
            self.statsfun = None
            return
        d = statsfun.desc()
        _capi.check(L.lh_set_step_stats(self.handle, C.byref(d)), "lh_set_step_stats")
        self.statsfun = statsfun

    def stats_eval(self, Ew, z=0.0, dz=0.0):
        """One record (``LH_NSTAT`` values) of the state Ew."""
        Ew = np.ascontiguousarray(np.asarray(Ew, dtype=np.complex128).reshape(-1))
        if Ew.size != self.n_state:
            raise ValueError("Eω has the wrong size")
        rec = np.empty(_capi.LH_NSTAT)
        _capi.check(_capi.lib().lh_stats_eval(self.handle, _p(Ew), z, dz, _p(rec)), "lh_stats_eval")
        return rec

    def step_stats(self, capacity=None):
        """Records [n][LH_NSTAT] of the last solve."""
        cap = capacity if capacity is not None else self.statsfun.capacity
        rec = np.empty((max(cap, 1), _capi.LH_NSTAT))
        n = C.c_int(0)
        _capi.check(_capi.lib().lh_get_step_stats(self.handle, _p(rec), cap, C.byref(n)), "lh_get_step_stats")
        return rec[:n.value].copy()


# ---------------------------------------------------------------- planning
# ``Luna.settings`` and ``set_fftw_mode`` (src/Luna.jl:13-36): the planning
# mode of the transforms.  "estimate" takes the recorded wisdom or the
# default plan without measuring; "measure", "patient" and "exhaustive" time
# the radial plans (3, 10 and 30 launches each) for a shape with no wisdom.
# The default here is "estimate" (the reference's is FFTW.PATIENT): a
# measured choice can change with the machine's load, and with it the last
# bits of a result.
settings = {"fftw_flag": "estimate", "fftw_threads": 0}
_REPS = {"measure": 3, "patient": 10, "exhaustive": 30}


def set_fftw_mode(mode):
    m = str(mode).lstrip(":").lower()
    if m not in ("estimate", "measure", "patient", "exhaustive"):
        raise ValueError(f"unknown planning mode {mode}")
    settings["fftw_flag"] = m
    return m


def plan_radial(ctx):
    """Plan a new radial context as ``Luna.setup`` plans its FFTs
    (src/Luna.jl:239-258): a shape in the wisdom (loaded before the context
    was created) keeps its plan; another is measured if the mode asks, and
    the wisdom file rewritten."""
    reps = _REPS.get(settings["fftw_flag"])
    if reps is None:
        return
    known = f"\nradial {ctx.holder.desc.n_to} {ctx.n_r} " in wisdom_export()
    if not known:
        ctx.autotune(None, reps)
        save_wisdom()


def wisdom_export():
    """The wisdom as text (``lh_wisdom_export``)."""
    L = _capi.lib()
    n = C.c_size_t(0)
    _capi.check(L.lh_wisdom_export(None, 0, C.byref(n)), "lh_wisdom_export")
    buf = C.create_string_buffer(n.value)
    _capi.check(L.lh_wisdom_export(buf, n.value, C.byref(n)), "lh_wisdom_export")
    return buf.value.decode()


def wisdom_import(text):
    _capi.check(_capi.lib().lh_wisdom_import(text.encode()), "lh_wisdom_import")


def wisdom_forget():
    _capi.check(_capi.lib().lh_wisdom_forget(), "lh_wisdom_forget")


def _wisdom_path():
    import os
    d = os.environ.get("LUNA_CACHE_DIR") or os.path.join(os.path.expanduser("~"), ".cache", "luna_jl_amd")
    return d, os.path.join(d, "plan_wisdom")


def load_wisdom():
    """``Utils.loadFFTwisdom`` (src/Utils.jl:84-97): merge the cached wisdom
    file, under a file lock, if there is one."""
    import fcntl
    import os
    d, path = _wisdom_path()
    if not os.path.isfile(path):
        return False
    with open(path + ".lock", "a") as lk:
        fcntl.flock(lk, fcntl.LOCK_SH)
        try:
            with open(path) as f:
                text = f.read()
        finally:
            fcntl.flock(lk, fcntl.LOCK_UN)
    try:
        wisdom_import(text)
    except _capi.LunaHipError:
        return False  # a malformed file is ignored, as FFTW ignores unusable wisdom
    return True


def save_wisdom():
    """``Utils.saveFFTwisdom`` (src/Utils.jl:99-108)."""
    import fcntl
    import os
    d, path = _wisdom_path()
    os.makedirs(d, exist_ok=True)
    text = wisdom_export()
    with open(path + ".lock", "a") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        try:
            tmp = f"{path}.tmp{os.getpid()}"
            with open(tmp, "w") as f:
                f.write(text)
            os.replace(tmp, path)
        finally:
            fcntl.flock(lk, fcntl.LOCK_UN)
    return path


class HostBuffers:
    """Page-locked host arrays (``lh_host_alloc``) for the final states, saves
    and save positions of a sweep: allocated once and handed to every
    ``batch_solve(..., out=...)`` of the sweep, so the copies back run at DMA
    speed into memory that is already mapped.  The arrays are views of the
    pinned block and are invalid after ``close()``."""

    def __init__(self, contexts, save_n):
        n = len(contexts)
        ns = [c.n_state for c in contexts]
        nz = max(save_n, 1)
        nbytes = sum(16 * m + 16 * save_n * m + 8 * nz for m in ns)
        ptr = C.c_void_p()
        _capi.check(_capi.lib().lh_host_alloc(nbytes, C.byref(ptr)), "lh_host_alloc")
        self._ptr = ptr
        raw = np.frombuffer((C.c_char * nbytes).from_address(ptr.value), dtype=np.uint8)
        self.states, self.saves, self.zs = [], [], []
        off = 0
        for m in ns:
            self.states.append(raw[off:off + 16 * m].view(np.complex128))
            off += 16 * m
            self.saves.append(raw[off:off + 16 * save_n * m].view(np.complex128).reshape(save_n, m))
            off += 16 * save_n * m
            self.zs.append(raw[off:off + 8 * nz].view(np.float64))
            off += 8 * nz
        self.save_n = save_n
        self.n = n

    def close(self):
        if self._ptr is not None and self._ptr.value:
            self.states = self.saves = self.zs = None
            _capi.lib().lh_host_free(self._ptr)
        self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_solve(contexts, Ews, opts, want_saves=True, out=None):
    """Run every instance in one persistent launch (``lh_batch_solve``).
    ``out``: HostBuffers for this sweep's contexts (states, saves and save
    positions are written there) or None (fresh arrays)."""
    n = len(contexts)
    handles = (C.c_void_p * n)(*[c.handle for c in contexts])
    if out is not None:
        if out.n != n or out.save_n != opts.save_n:
            raise ValueError("HostBuffers made for another sweep shape")
        ew_arrs = out.states
        for a, e in zip(ew_arrs, Ews):
            a[:] = np.asarray(e, dtype=np.complex128).reshape(-1, order="F")
    else:
        ew_arrs = [np.ascontiguousarray(np.asarray(e, dtype=np.complex128).reshape(-1, order="F")).copy()
                   for e in Ews]
    ew_ptrs = (C.POINTER(C.c_double) * n)(*[_p(a) for a in ew_arrs])
    if want_saves:
        sv = out.saves if out is not None else [np.zeros((opts.save_n, c.n_state), dtype=np.complex128)
                                                for c in contexts]
        sv_ptrs = (C.POINTER(C.c_double) * n)(*[_p(a) for a in sv])
    else:
        sv = [None] * n
        sv_ptrs = None
    zs = out.zs if out is not None else [np.zeros(max(opts.save_n, 1)) for _ in range(n)]
    zs_ptrs = (C.POINTER(C.c_double) * n)(*[_p(a) for a in zs])
    stats = (_capi.LhStats * n)()
    st = _capi.lib().lh_batch_solve(handles, n, ew_ptrs, C.byref(opts), sv_ptrs, zs_ptrs, stats)
    _capi.check(st, "lh_batch_solve")
    return ew_arrs, sv, zs, list(stats)


def batch_solve_opts(contexts, Ews, opts_list, want_saves=True):
    """``lh_batch_solve_opts``: one batch, each instance with its own options
    (save grid, z range, tolerances)."""
    n = len(contexts)
    handles = (C.c_void_p * n)(*[c.handle for c in contexts])
    ew_arrs = [np.ascontiguousarray(np.asarray(e, dtype=np.complex128).reshape(-1, order="F")).copy() for e in Ews]
    ew_ptrs = (C.POINTER(C.c_double) * n)(*[_p(a) for a in ew_arrs])
    sv = [np.zeros((o.save_n, c.n_state), dtype=np.complex128) for c, o in zip(contexts, opts_list)] \
        if want_saves else [None] * n
    sv_ptrs = (C.POINTER(C.c_double) * n)(*[_p(a) for a in sv]) if want_saves else None
    zs = [np.zeros(max(o.save_n, 1)) for o in opts_list]
    zs_ptrs = (C.POINTER(C.c_double) * n)(*[_p(a) for a in zs])
    opts = (_capi.LhRkOpts * n)(*opts_list)
    stats = (_capi.LhStats * n)()
    _capi.check(_capi.lib().lh_batch_solve_opts(handles, n, ew_ptrs, opts, sv_ptrs, zs_ptrs, stats),
                "lh_batch_solve_opts")
    return ew_arrs, sv, zs, list(stats)


class ResidentBatch:
    """A sweep whose initial fields stay in HBM: upload once, solve many times.

    ``solve`` runs every instance from its resident initial field; final states
    and saves remain on the device until ``state(i)`` / ``saves(i)`` copy them.
    """

    def __init__(self, contexts, Ews):
        self.contexts = list(contexts)
        L = _capi.lib()
        for c, e in zip(self.contexts, Ews):
            e = np.ascontiguousarray(np.asarray(e, dtype=np.complex128).reshape(-1, order="F"))
            _capi.check(L.lh_set_state(c.handle, _p(e)), "lh_set_state")
        n = len(self.contexts)
        self._handles = (C.c_void_p * n)(*[c.handle for c in self.contexts])

    def solve(self, opts, want_saves=True):
        n = len(self.contexts)
        stats = (_capi.LhStats * n)()
        _capi.check(_capi.lib().lh_batch_solve_resident(self._handles, n, C.byref(opts), int(want_saves), stats),
                    "lh_batch_solve_resident")
        self.save_n = opts.save_n
        return list(stats)

    def state(self, i):
        c = self.contexts[i]
        out = np.empty(c.n_state, dtype=np.complex128)
        _capi.check(_capi.lib().lh_get_state(c.handle, _p(out)), "lh_get_state")
        return out

    def saves(self, i):
        c = self.contexts[i]
        sv = np.empty((self.save_n, c.n_state), dtype=np.complex128)
        zs = np.empty(self.save_n)
        _capi.check(_capi.lib().lh_get_saves(c.handle, _p(sv), _p(zs)), "lh_get_saves")
        return sv, zs


def last_launch_times():
    """{n_to: (ms, instances)} of the persistent launches of the last batch."""
    k = len(_capi.CLASS_NAMES)
    ms = (C.c_double * k)()
    n = (C.c_int * k)()
    _capi.check(_capi.lib().lh_last_launch_times(ms, n, k), "lh_last_launch_times")
    return {_capi.CLASS_NAMES[c]: (ms[c], n[c]) for c in range(k) if n[c] > 0}
