"""Spectral surrogate for capillary scans (no counterpart in the reference).

A 1-D Fourier neural operator (include/luna_fno.h; kernels in
csrc/lh_fno.hip) maps (energy, pressure, input spectrum on 1024 bins) to the
log exit spectrum on the same bins.  It is trained on exit spectra that the
batched FP64 solver (lh_batch_solve) produces for the capillary scan, and it
screens new scan points: a point whose predicted error is above the bound is
run on the solver, which stays the judge.

Predicted error: the surrogate's error is measured against the solver on a
held-out calibration set; the predicted error of a new point is the largest
calibration error among its k nearest calibration points in the normalised
(energy, pressure) plane, times a safety factor.
"""
import ctypes as C
import math

import numpy as np

from . import _capi
from . import physdata as pd

N_BINS = 1024
LAMBDA_LIMS = (150e-9, 4e-6)
FLOOR = 1e-6  # relative floor of the log spectrum


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def layout(n_x=N_BINS, cin=4, channels=64, modes=32, layers=4, hidden=128):
    """Parameter vector layout (name → (offset, shape)), as in include/luna_fno.h."""
    out, off = {}, 0

    def take(name, shape):
        nonlocal off
        out[name] = (off, shape)
        off += int(np.prod(shape))

    take("P", (channels, cin))
    take("bP", (channels,))
    for l in range(layers):
        take(f"W{l}", (channels, channels))
        take(f"b{l}", (channels,))
        take(f"Rre{l}", (modes, channels, channels))
        take(f"Rim{l}", (modes, channels, channels))
    take("Q1", (hidden, channels))
    take("bq1", (hidden,))
    take("Q2", (hidden,))
    take("bq2", (1,))
    return out, off


class FNO:
    """One network on the GPU (BF16 matrix-core products, FP32 master weights)."""

    def __init__(self, n_x=N_BINS, cin=4, channels=64, modes=32, layers=4, hidden=128, max_batch=64):
        self.dims = dict(n_x=n_x, cin=cin, channels=channels, modes=modes, layers=layers, hidden=hidden)
        self.max_batch = max_batch
        cfg = _capi.LhFnoCfg(n_x=n_x, cin=cin, channels=channels, modes=modes, layers=layers, hidden=hidden,
                             max_batch=max_batch)
        L = _capi.lib()
        h = C.c_void_p()
        _capi.check(L.lh_fno_create(C.byref(h), C.byref(cfg)), "lh_fno_create")
        self.handle = h
        n = C.c_int64()
        _capi.check(L.lh_fno_param_count(h, C.byref(n)), "lh_fno_param_count")
        self.n_params = n.value
        self.layout, total = layout(**self.dims)
        assert total == self.n_params

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _capi.lib().lh_fno_destroy(h)
            except Exception:
                pass
            self.handle = None

    def init_params(self, seed=0):
        """Linear layers U(±1/√fan_in); spectral weights U(0, 1/C²) (FNO reference initialisation)."""
        rng = np.random.default_rng(seed)
        p = np.zeros(self.n_params, dtype=np.float32)
        C_ = self.dims["channels"]
        fan = {"P": self.dims["cin"], "bP": self.dims["cin"], "Q1": C_, "bq1": C_, "Q2": self.dims["hidden"],
               "bq2": self.dims["hidden"]}
        for name, (o, shape) in self.layout.items():
            n = int(np.prod(shape))
            if name.startswith("R"):
                v = rng.uniform(0.0, 1.0 / (C_ * C_), n)
            else:
                fi = fan.get(name, C_)
                v = rng.uniform(-1.0, 1.0, n) / math.sqrt(fi)
            p[o:o + n] = v
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        _capi.check(_capi.lib().lh_fno_set_params(self.handle, _fp(p)), "lh_fno_set_params")

    def get_params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        _capi.check(_capi.lib().lh_fno_get_params(self.handle, _fp(p)), "lh_fno_get_params")
        return p

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(x.shape[:2], dtype=np.float32)
        _capi.check(_capi.lib().lh_fno_forward(self.handle, _fp(x), x.shape[0], _fp(y)), "lh_fno_forward")
        return y

    def loss_grad(self, x, t, w=None):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32)
        _capi.check(_capi.lib().lh_fno_loss_grad(self.handle, _fp(x), _fp(t), None if w is None else _fp(w),
                                                 x.shape[0], _fp(loss), _fp(g)), "lh_fno_loss_grad")
        return float(loss[0]), g

    def set_data(self, x, t, w=None):
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        self._t = np.ascontiguousarray(t, dtype=np.float32)
        self._w = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        _capi.check(_capi.lib().lh_fno_set_data(self.handle, _fp(self._x), _fp(self._t),
                                                None if self._w is None else _fp(self._w), self._x.shape[0]),
                    "lh_fno_set_data")
        self.n_data = self._x.shape[0]

    def train(self, steps, batch=32, lr=1e-3, lr_min=1e-5, weight_decay=1e-4, seed=0, chunk=500):
        """AdamW with a cosine learning-rate schedule over shuffled epochs; returns per-step losses."""
        rng = np.random.default_rng(seed)
        n = self.n_data
        idx = []
        while len(idx) < steps * batch:
            idx.extend(rng.permutation(n).tolist())
        idx = np.asarray(idx[:steps * batch], dtype=np.int32)
        sched = (lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * np.arange(steps) / steps))).astype(np.float32)
        opt = _capi.LhAdam(lr=lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=weight_decay)
        losses = np.empty(steps, dtype=np.float32)
        for s0 in range(0, steps, chunk):
            k = min(chunk, steps - s0)
            ib = np.ascontiguousarray(idx[s0 * batch:(s0 + k) * batch])
            sc = np.ascontiguousarray(sched[s0:s0 + k])
            out = np.empty(k, dtype=np.float32)
            _capi.check(_capi.lib().lh_fno_train(self.handle, ib.ctypes.data_as(C.POINTER(C.c_int32)), k, batch,
                                                  C.byref(opt), _fp(sc), _fp(out)), "lh_fno_train")
            losses[s0:s0 + k] = out
        return losses


# ---------------------------------------------------------------------------
# Features and targets
# ---------------------------------------------------------------------------
def spectral_axis(n=N_BINS, lambda_lims=LAMBDA_LIMS):
    """Uniform angular-frequency bins spanning the grid's wavelength limits."""
    return np.linspace(pd.wlfreq(lambda_lims[1]), pd.wlfreq(lambda_lims[0]), n)


def resample(omega, Ew, axis):
    """|Eω|² (the continuous-FT spectral density up to a constant: δt is the
    same for every grid of the scan) interpolated onto the common axis."""
    return np.interp(axis, omega, np.abs(Ew) ** 2)


class Scaler:
    """log10(S/S_ref + floor), standardised; and its inverse."""

    def __init__(self, S_ref, mu=0.0, sigma=1.0):
        self.S_ref, self.mu, self.sigma = float(S_ref), float(mu), float(sigma)

    def to_log(self, S):
        return np.log10(np.maximum(S, 0.0) / self.S_ref + FLOOR)

    def fit(self, S):
        L = self.to_log(S)
        self.mu, self.sigma = float(L.mean()), float(L.std())
        return self

    def encode(self, S):
        return (self.to_log(S) - self.mu) / self.sigma

    def decode(self, y):
        return self.S_ref * (10.0 ** (np.asarray(y, dtype=np.float64) * self.sigma + self.mu) - FLOOR)


PARAM_RANGES = {"energy": (50e-6, 200e-6), "pressure": (0.5, 3.0)}


def normalise(v, lo, hi):
    return 2.0 * (np.asarray(v) - lo) / (hi - lo) - 1.0


def features(S_in, energy, pressure, xs_scaler):
    """Input features per bin: [encoded input spectrum, energy, pressure, bin coordinate]."""
    n = S_in.shape[0]
    x = np.empty((n, N_BINS, 4), dtype=np.float32)
    x[:, :, 0] = xs_scaler.encode(S_in)
    x[:, :, 1] = normalise(energy, *PARAM_RANGES["energy"])[:, None]
    x[:, :, 2] = normalise(pressure, *PARAM_RANGES["pressure"])[:, None]
    x[:, :, 3] = np.linspace(-1.0, 1.0, N_BINS)[None, :]
    return x


def loss_weights(S, floor=1e-3):
    """Per-bin loss weights ∝ S/max(S) + floor, mean 1 per sample: the
    weighted log error then tracks the linear spectral error."""
    w = S / S.max(axis=-1, keepdims=True) + floor
    return (w / w.mean(axis=-1, keepdims=True)).astype(np.float32)


def spectral_error(S_pred, S_ref):
    """‖S_pred − S_ref‖₂ / ‖S_ref‖₂ per row (linear spectral energy density)."""
    return np.linalg.norm(S_pred - S_ref, axis=-1) / np.linalg.norm(S_ref, axis=-1)


# ---------------------------------------------------------------------------
# Data from the batched solver
# ---------------------------------------------------------------------------
def scan_params(n, seed):
    """Energy U[50, 200] μJ, pressure U[0.5, 3] bar; alternating 0.5 / 1 ps windows."""
    rng = np.random.default_rng(seed)
    E = rng.uniform(*PARAM_RANGES["energy"], n)
    P = rng.uniform(*PARAM_RANGES["pressure"], n)
    T = np.where(np.arange(n) % 2 == 0, 0.5e-12, 1e-12)
    return E, P, T


def solve_spectra(E, P, T, flength=1.0, chunk=512, log=None):
    """Exit and input spectra on the common axis from the batched FP64 solver.
    Returns (S_in, S_out, status)."""
    from . import configs, runtime
    axis = spectral_axis()
    n = len(E)
    S_in = np.zeros((n, N_BINS))
    S_out = np.zeros((n, N_BINS))
    status = np.zeros(n, dtype=int)
    opts = _capi.rk_opts(flength, save_n=0)
    for c0 in range(0, n, chunk):
        sl = range(c0, min(n, c0 + chunk))
        insts = [configs.capillary_instance(energy=float(E[i]), pressure=float(P[i]), trange=float(T[i]),
                                            flength=flength) for i in sl]
        ctxs = [runtime.Context(c.holder) for c in insts]
        batch = runtime.ResidentBatch(ctxs, [c.Ew0 for c in insts])
        stats = batch.solve(opts, want_saves=False)
        for j, i in enumerate(sl):
            g = insts[j].grid
            S_in[i] = resample(g.omega, insts[j].Ew0, axis)
            S_out[i] = resample(g.omega, batch.state(j), axis)
            status[i] = stats[j].status
        del batch, ctxs
        if log:
            log(f"surrogate data: {sl.stop}/{n} solved")
    return S_in, S_out, status


class Screen:
    """Predicted-error model from a held-out calibration set: safety × the
    largest calibration error among the k nearest calibration points in the
    normalised (energy, pressure) plane.

    With safety=None the factor is calibrated by leave-one-out on the
    calibration set: the smallest factor on a 0.1 grid for which at most a
    fraction `miss` of the calibration points that would be accepted have an
    error above the bound, plus one grid step of margin.  (Requiring zero
    misses lets one isolated outlier set the factor for the whole plane.)"""

    def __init__(self, E_cal, P_cal, err_cal, k=8, safety=None, bound=1e-2, miss=0.01):
        self.pts = np.stack([normalise(E_cal, *PARAM_RANGES["energy"]),
                             normalise(P_cal, *PARAM_RANGES["pressure"])], axis=1)
        self.err = np.asarray(err_cal, dtype=float)
        self.k = k
        self.miss = miss
        self.safety = self.calibrate(bound) if safety is None else float(safety)

    def _nearest(self, q, exclude_self=False):
        d = ((q[:, None, :] - self.pts[None, :, :]) ** 2).sum(-1)
        if exclude_self:
            np.fill_diagonal(d, np.inf)
        return np.argsort(d, axis=1)[:, :self.k]

    def calibrate(self, bound, grid=np.round(np.arange(1.0, 3.01, 0.1), 1)):
        local = self.err[self._nearest(self.pts, exclude_self=True)].max(axis=1)
        for s in grid:
            acc = s * local <= bound
            if np.sum(self.err[acc] > bound) <= self.miss * max(1, int(acc.sum())):
                return float(s) + 0.1
        return float(grid[-1]) + 0.1

    def predicted_error(self, E, P):
        q = np.stack([normalise(E, *PARAM_RANGES["energy"]), normalise(P, *PARAM_RANGES["pressure"])], axis=1)
        return self.safety * self.err[self._nearest(q)].max(axis=1)
