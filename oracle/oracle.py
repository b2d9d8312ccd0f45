"""ctypes wrapper of the C oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke`` and bench.py's ``cpu_baseline`` leg load
this module, and only as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libluna_oracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib(capi):
    """Load the oracle; ``capi`` is the package's ``_capi`` module (struct layouts)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(LIB):
        build()
    L = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    D = C.POINTER(capi.LhDesc)
    L.orc_fft_c2c.argtypes = [C.c_int, dp, C.c_int]
    L.orc_rfft.argtypes = [C.c_int, dp, dp]
    L.orc_irfft_unnorm.argtypes = [C.c_int, dp, dp]
    L.orc_cumtrapz.argtypes = [C.c_int, dp, C.c_double, dp]
    L.orc_cspline_eval.argtypes = [C.c_int, dp, dp, dp, C.c_int, dp, dp]
    L.orc_ppt_rate.argtypes = [D, C.c_int, dp, dp]
    L.orc_responses_real.argtypes = [D, dp, dp]
    L.orc_rhs.argtypes = [D, dp, dp, C.c_double]
    L.orc_solve.argtypes = [D, dp, C.POINTER(capi.LhRkOpts), dp, dp, C.POINTER(capi.LhStats)]
    L.orc_step_stats.argtypes = [D, dp, C.c_double, C.c_int, dp, dp, C.c_double, C.c_double, dp]
    L.orc_solve_stats.argtypes = [D, dp, C.POINTER(capi.LhRkOpts), dp, dp, C.POINTER(capi.LhStats), dp,
                                  C.c_double, C.c_int, dp, dp, C.c_int, C.POINTER(C.c_int)]
    L.orc_fwhm.argtypes = [C.c_int, dp, dp, C.c_int, dp]
    L.orc_solve_grad_stats.argtypes = [D, C.POINTER(capi.LhGradientDesc), dp, C.POINTER(capi.LhRkOpts), dp, dp,
                                       C.POINTER(capi.LhStats), dp, C.c_double, C.c_int, dp, dp, C.c_int,
                                       C.POINTER(C.c_int)]
    G = C.POINTER(capi.LhGradientDesc)
    L.orc_rhs_grad.argtypes = [D, G, dp, dp, C.c_double]
    L.orc_solve_grad.argtypes = [D, G, dp, C.POINTER(capi.LhRkOpts), dp, dp, C.POINTER(capi.LhStats)]
    L.orc_grad_constants.argtypes = [D, G, C.c_double, dp, dp, dp]
    L.ORC_RHS_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, dp, dp, C.c_double)
    L.orc_solve_precon.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, L.ORC_RHS_CB, C.c_void_p, C.c_int, dp, dp,
                                   C.POINTER(C.c_int64)]
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    def __init__(self, capi):
        self.capi = capi
        self.L = lib(capi)

    def fft(self, x, sign=-1):
        y = np.ascontiguousarray(x, dtype=np.complex128).copy()
        self.L.orc_fft_c2c(len(y), _p(y), sign)
        return y

    def rfft(self, x):
        x = np.ascontiguousarray(x, dtype=float)
        out = np.zeros(len(x) // 2 + 1, dtype=np.complex128)
        self.L.orc_rfft(len(x), _p(x), _p(out))
        return out

    def irfft_unnorm(self, X, n):
        X = np.ascontiguousarray(X, dtype=np.complex128)
        out = np.zeros(n)
        self.L.orc_irfft_unnorm(n, _p(X), _p(out))
        return out

    def cumtrapz(self, y, dx):
        y = np.ascontiguousarray(y, dtype=float)
        out = np.zeros_like(y)
        self.L.orc_cumtrapz(len(y), _p(y), dx, _p(out))
        return out

    def cspline(self, x, y, D, x0):
        x0 = np.ascontiguousarray(x0, dtype=float)
        out = np.zeros_like(x0)
        self.L.orc_cspline_eval(len(x), _p(np.ascontiguousarray(x)), _p(np.ascontiguousarray(y)),
                                _p(np.ascontiguousarray(D)), len(x0), _p(x0), _p(out))
        return out

    def ppt_rate(self, holder, E):
        E = np.ascontiguousarray(E, dtype=float)
        out = np.zeros_like(E)
        st = self.L.orc_ppt_rate(holder.ref, len(E), _p(E), _p(out))
        return st, out

    def responses_real(self, holder, E):
        E = np.ascontiguousarray(E, dtype=float)
        out = np.zeros_like(E)
        st = self.L.orc_responses_real(holder.ref, _p(E), _p(out))
        return st, out

    def rhs(self, holder, Ew, z=0.0):
        # multi-column states are flattened column-major (ω fastest)
        Ew = np.ascontiguousarray(np.asarray(Ew, dtype=np.complex128).reshape(-1, order="F"))
        nl = np.zeros_like(Ew)
        st = self.L.orc_rhs(holder.ref, _p(nl), _p(Ew), z)
        return st, nl

    def solve(self, holder, Ew0, opts, want_saves=True):
        d = holder.desc
        n = d.n_omega * d.n_r
        Ew = np.ascontiguousarray(np.asarray(Ew0, dtype=np.complex128).reshape(-1, order="F")).copy()
        saves = np.zeros((opts.save_n, n), dtype=np.complex128) if want_saves else None
        zs = np.zeros(opts.save_n)
        stats = self.capi.LhStats()
        st = self.L.orc_solve(holder.ref, _p(Ew), C.byref(opts), _p(saves) if want_saves else None,
                              _p(zs), C.byref(stats))
        return st, Ew, saves, zs, stats

    # -- gradient fill ------------------------------------------------------
    def grad_constants(self, holder, spec, z):
        """(operator L(ω, z), norm factor, ρ(z)) of a gradient fill."""
        n = holder.desc.n_omega
        Lz = np.zeros(n, dtype=np.complex128)
        nf = np.zeros(n, dtype=np.complex128)
        rho = C.c_double(0.0)
        self.L.orc_grad_constants(holder.ref, spec.ref, z, _p(Lz), _p(nf), C.byref(rho))
        return Lz, nf, rho.value

    def rhs_grad(self, holder, spec, Ew, z):
        Ew = np.ascontiguousarray(Ew, dtype=np.complex128)
        nl = np.zeros_like(Ew)
        st = self.L.orc_rhs_grad(holder.ref, spec.ref, _p(nl), _p(Ew), z)
        return st, nl

    def solve_grad(self, holder, spec, Ew0, opts, want_saves=True):
        n = holder.desc.n_omega
        Ew = np.ascontiguousarray(Ew0, dtype=np.complex128).copy()
        saves = np.zeros((opts.save_n, n), dtype=np.complex128) if want_saves else None
        zs = np.zeros(max(opts.save_n, 1))
        stats = self.capi.LhStats()
        st = self.L.orc_solve_grad(holder.ref, spec.ref, _p(Ew), C.byref(opts), _p(saves) if want_saves else None,
                                   _p(zs), C.byref(stats))
        return st, Ew, saves, zs, stats

    def fwhm(self, x, y, minmax="min"):
        x = np.ascontiguousarray(x, dtype=float)
        y = np.ascontiguousarray(y, dtype=float)
        out = C.c_double(0.0)
        self.L.orc_fwhm(len(x), _p(x), _p(y), int(minmax == "max"), C.byref(out))
        return out.value

    @staticmethod
    def _windows(windows):
        """Window values [nwin][Nω] (or None) → (nwin, contiguous array or None)."""
        if windows is None or len(windows) == 0:
            return 0, None
        w = np.ascontiguousarray(np.atleast_2d(np.asarray(windows, dtype=float)))
        return w.shape[0], w

    def step_stats(self, holder, t, aeff, Ew, z=0.0, dz=0.0, windows=None):
        """One statistics record (``capi.LH_NSTAT`` values) of the state Ew;
        ``windows``: window values [nwin][Nω] of the energy windows."""
        Ew = np.ascontiguousarray(Ew, dtype=np.complex128)
        t = np.ascontiguousarray(t, dtype=float)
        rec = np.zeros(self.capi.LH_NSTAT)
        nw, w = self._windows(windows)
        st = self.L.orc_step_stats(holder.ref, _p(t), float(aeff), nw, _p(w) if nw else None, _p(Ew), z, dz,
                                   _p(rec))
        return st, rec

    def solve_stats(self, holder, Ew0, opts, t, aeff, cap=4096, want_saves=False, windows=None, spec=None):
        """``solve`` that also returns the per-step statistics records [n][LH_NSTAT]."""
        d = holder.desc
        n = d.n_omega * d.n_r
        Ew = np.ascontiguousarray(np.asarray(Ew0, dtype=np.complex128).reshape(-1, order="F")).copy()
        saves = np.zeros((opts.save_n, n), dtype=np.complex128) if want_saves else None
        zs = np.zeros(max(opts.save_n, 1))
        stats = self.capi.LhStats()
        t = np.ascontiguousarray(t, dtype=float)
        rec = np.zeros((cap, self.capi.LH_NSTAT))
        nrec = C.c_int(0)
        nw, w = self._windows(windows)
        if spec is not None:  # a gradient fill (GradientSpec)
            st = self.L.orc_solve_grad_stats(holder.ref, spec.ref, _p(Ew), C.byref(opts),
                                             _p(saves) if want_saves else None, _p(zs), C.byref(stats), _p(t),
                                             float(aeff), nw, _p(w) if nw else None, _p(rec), cap, C.byref(nrec))
        else:
            st = self.L.orc_solve_stats(holder.ref, _p(Ew), C.byref(opts), _p(saves) if want_saves else None,
                                        _p(zs), C.byref(stats), _p(t), float(aeff), nw, _p(w) if nw else None,
                                        _p(rec), cap, C.byref(nrec))
        return st, Ew, stats, rec[:nrec.value]

    def solve_precon(self, f, linop, y0, t0, dt0, tmax, rtol=1e-6, atol=1e-10, outputN=201):
        n = len(y0)
        linop = np.ascontiguousarray(linop, dtype=np.complex128)
        y0 = np.ascontiguousarray(y0, dtype=np.complex128)
        tout = np.zeros(outputN)
        yout = np.zeros((outputN, n), dtype=np.complex128)

        def cb(user, out, y, t):
            yv = np.ctypeslib.as_array(y, shape=(2 * n,)).view(np.complex128)
            ov = np.ctypeslib.as_array(out, shape=(2 * n,)).view(np.complex128)
            ov[:] = f(yv.copy(), t)
            return 0

        cfun = self.L.ORC_RHS_CB(cb)
        steps = C.c_int64(0)
        st = self.L.orc_solve_precon(n, _p(linop), _p(y0), t0, dt0, tmax, rtol, atol, cfun, None,
                                     outputN, _p(tout), _p(yout), C.byref(steps))
        return st, tout, yout, steps.value
