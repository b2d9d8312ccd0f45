"""ctypes mirror of ``include/luna_hip.h`` and the loader of ``libluna_hip.so``.

The shared library is built in-tree (``luna.jl_amd/_build/libluna_hip.so``) by
``luna.jl_amd/build.py``.  There is no CPU fallback: if the library is missing
every entry point raises.
"""
import ctypes as C
import os

import numpy as np

(LH_OK, LH_E_ARG, LH_E_HIP, LH_E_PPT_RANGE, LH_E_REPEAT_LIMIT, LH_E_NONFINITE, LH_E_NOMEM, LH_E_NODEVICE,
 LH_E_MAXSTEPS, LH_E_STOPPED, LH_E_TOO_LARGE, LH_E_HOST_WAIT) = range(12)
LH_NOALLOC_SCRATCH_BYTES = 1536 * 1024
LH_GRID_REAL, LH_GRID_ENV = 0, 1
LH_TRANS_MODEAVG, LH_TRANS_RADIAL = 0, 1
LH_KERR_NONE, LH_KERR_FIELD, LH_KERR_ENV = 0, 1, 2
# per-step statistics record: field order of LH_STAT_* (keys as Stats.jl names them)
STAT_KEYS = ("z", "dz", "ω0", "energy", "peakpower", "peakintensity", "fwhm_t_min", "fwhm_t_max",
             "electrondensity", "peak_ionisation_rate", "density", "window0", "window1", "window2", "window3")
LH_NSTAT = len(STAT_KEYS)
LH_STAT_MAX_WINDOWS = 4

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


class LhDesc(C.Structure):
    _fields_ = [
        ("grid_kind", C.c_int32), ("transform", C.c_int32),
        ("n_omega", C.c_int32), ("n_t", C.c_int32), ("n_to", C.c_int32), ("n_r", C.c_int32),
        ("dt_o", C.c_double),
        ("omega", _dp), ("sidx", _u8p), ("omega_win", _dp), ("t_win", _dp), ("to_win", _dp),
        ("linop", _dp), ("norm", _dp),
        ("field_div", C.c_double), ("density", C.c_double),
        ("kerr", C.c_int32), ("plasma", C.c_int32), ("gamma3", C.c_double),
        ("ppt_n", C.c_int32), ("raman", C.c_int32),
        ("ppt_x", _dp), ("ppt_y", _dp), ("ppt_d", _dp),
        ("ionpot", C.c_double), ("preionfrac", C.c_double),
        ("raman_h", _dp), ("qdht", _dp), ("qdht_fwd", C.c_double), ("qdht_inv", C.c_double),
    ]


class LhRkOpts(C.Structure):
    _fields_ = [
        ("z0", C.c_double), ("zmax", C.c_double), ("rtol", C.c_double), ("atol", C.c_double),
        ("safety", C.c_double), ("max_dz", C.c_double), ("min_dz", C.c_double),
        ("init_dz", C.c_double), ("repeat_limit", C.c_int32), ("save_n", C.c_int32),
        ("save_z0", C.c_double), ("save_first", C.c_int32), ("every_step", C.c_int32),
        ("zstop", C.c_double), ("max_steps", C.c_int64),
    ]


class LhStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("accepted", C.c_int64), ("rejected", C.c_int64),
        ("rhs_evals", C.c_int64), ("status", C.c_int32), ("saved", C.c_int32),
        ("z", C.c_double), ("dz", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LhStepOut(C.Structure):
    _fields_ = [("first_save", C.c_int32), ("n_saves", C.c_int32), ("saves", _dp), ("zsaves", _dp),
                ("z", C.c_double), ("dz", C.c_double), ("y", _dp), ("accepted", C.c_int64),
                ("first_record", C.c_int32), ("n_records", C.c_int32), ("records", _dp)]


LhOutputCb = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(LhStepOut))


class LhGradientDesc(C.Structure):
    """``lh_gradient_desc`` of include/luna_hip.h."""
    _fields_ = [("n_points", C.c_int32), ("two_point", C.c_int32), ("Z", _dp), ("P", _dp),
                ("dens_n", C.c_int32), ("model", C.c_int32), ("dens_p", _dp), ("dens_rho", _dp),
                ("dens_d", _dp), ("loss", C.c_int32), ("taper", C.c_int32), ("gamma", _dp),
                ("nwg", _dp), ("pre", _dp), ("sqrt_aeff", C.c_double), ("fd_h", C.c_double), ("fd", _dp),
                ("taper_a0", C.c_double), ("taper_a1", C.c_double), ("taper_L", C.c_double),
                ("sqrt_aeff_intg", C.c_double), ("wgU", _dp), ("wgV", _dp), ("fdt", _dp),
                ("taper_n", C.c_int32), ("taper_reserved", C.c_int32), ("taper_Z", _dp), ("taper_A", _dp)]


class LhPptDesc(C.Structure):
    """``lh_ppt_desc`` of include/luna_hip.h."""
    _fields_ = [("ionpot", C.c_double), ("lambda0", C.c_double), ("Z", C.c_double), ("l", C.c_int32),
                ("msum", C.c_int32), ("dalpha", C.c_double), ("alpha_ion_au", C.c_double),
                ("sum_tol", C.c_double), ("occupancy", C.c_double), ("cycle_average", C.c_int32),
                ("reserved", C.c_int32), ("cnl", C.c_double), ("c", C.c_double), ("au_energy", C.c_double),
                ("au_time", C.c_double), ("au_efield", C.c_double)]


class LhStatsDesc(C.Structure):
    """``lh_stats_desc`` of include/luna_hip.h."""
    _fields_ = [("capacity", C.c_int32), ("n_windows", C.c_int32), ("energy_w", C.POINTER(C.c_double)),
                ("t", C.POINTER(C.c_double)), ("aeff", C.c_double), ("window_w", C.POINTER(C.c_double))]


class LhFnoCfg(C.Structure):
    """``lh_fno_cfg`` of include/luna_fno.h."""
    _fields_ = [("n_x", C.c_int32), ("cin", C.c_int32), ("channels", C.c_int32), ("modes", C.c_int32),
                ("layers", C.c_int32), ("hidden", C.c_int32), ("max_batch", C.c_int32)]


class LhPlan(C.Structure):
    """``lh_plan`` of include/luna_hip.h: one variant of the radial RHS."""
    _fields_ = [("fft_radix_plan", C.c_int32), ("gemm_tiles", C.c_int32), ("rhs_us", C.c_double),
                ("agrees", C.c_int32)]


class LhAdam(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("weight_decay", C.c_float)]


def _ptr(a, ctype=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class DescHolder:
    """Owns the numpy arrays an ``LhDesc`` points into (keeps them alive)."""

    def __init__(self, **arrays_and_scalars):
        self._keep = []
        d = LhDesc()
        for name, val in arrays_and_scalars.items():
            if isinstance(val, np.ndarray):
                if name == "sidx":
                    arr = np.ascontiguousarray(val, dtype=np.uint8)
                    setattr(d, name, _ptr(arr, C.c_uint8))
                else:
                    if np.iscomplexobj(val):
                        arr = np.ascontiguousarray(val, dtype=np.complex128)
                    else:
                        arr = np.ascontiguousarray(val, dtype=np.float64)
                    setattr(d, name, _ptr(arr))
                self._keep.append(arr)
            elif val is not None:
                setattr(d, name, val)
        self.desc = d

    @property
    def ref(self):
        return C.byref(self.desc)


def rk_opts(zmax, z0=0.0, rtol=1e-6, atol=1e-10, safety=0.9, max_dz=None, min_dz=0.0,
            init_dz=1e-4, repeat_limit=10, save_n=201, save_z0=0.0, save_first=0, zstop=0.0, max_steps=0,
            every_step=False):
    """``Luna.run`` defaults, ``src/Luna.jl:358-362``; the save grid
    ``range(save_z0, zmax, save_n)`` is MemoryOutput's/HDF5Output's
    ``GridCondition(0, zmax, saveN)``; ``save_first``/``zstop`` resume and stop
    a run (HDF5Output's cache, ``src/Output.jl:473-487``); ``every_step``:
    ``lh_solve_stream`` hands over every accepted step, not only those that
    save."""
    o = LhRkOpts()
    o.z0, o.zmax, o.rtol, o.atol, o.safety = z0, zmax, rtol, atol, safety
    o.max_dz = zmax / 2 if max_dz is None else max_dz
    o.min_dz, o.init_dz, o.repeat_limit, o.save_n = min_dz, init_dz, repeat_limit, save_n
    o.save_z0, o.save_first, o.zstop, o.max_steps = save_z0, save_first, zstop, max_steps
    o.every_step = 1 if every_step else 0
    return o


_HERE = os.path.dirname(os.path.abspath(__file__))
# LUNA_HIP_LIB selects another build of the same library (A/B timing of kernel variants)
LIB_PATH = os.environ.get("LUNA_HIP_LIB") or os.path.join(_HERE, "_build", "libluna_hip.so")
_lib = None


class LunaHipError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = lib().lh_strerror(status).decode() if _lib is not None else str(status)
        super().__init__(f"{where}: {msg} (status {status})")


def lib():
    """Load libluna_hip.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(LIB_PATH):
        raise RuntimeError(f"libluna_hip.so not built: expected {LIB_PATH} "
                           "(run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    L.lh_strerror.restype = C.c_char_p
    L.lh_strerror.argtypes = [C.c_int]
    L.lh_device_count.argtypes = [C.POINTER(C.c_int)]
    L.lh_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(LhDesc)]
    L.lh_destroy.argtypes = [C.c_void_p]
    L.lh_rhs.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
    L.lh_rhs_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
    L.lh_ctx_status.argtypes = [C.c_void_p]
    L.lh_solve.argtypes = [C.c_void_p, _dp, C.POINTER(LhRkOpts), _dp, _dp, C.POINTER(LhStats)]
    L.lh_solve_stream.argtypes = [C.c_void_p, _dp, C.POINTER(LhRkOpts), LhOutputCb, C.c_void_p,
                                  C.POINTER(LhStats)]
    L.lh_batch_solve.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(_dp), C.POINTER(LhRkOpts),
                                 C.POINTER(_dp), C.POINTER(_dp), C.POINTER(LhStats)]
    L.lh_batch_solve_opts.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(_dp), C.POINTER(LhRkOpts),
                                      C.POINTER(_dp), C.POINTER(_dp), C.POINTER(LhStats)]
    L.lh_set_state.argtypes = [C.c_void_p, _dp]
    L.lh_get_state.argtypes = [C.c_void_p, _dp]
    L.lh_get_saves.argtypes = [C.c_void_p, _dp, _dp]
    L.lh_batch_solve_resident.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(LhRkOpts), C.c_int,
                                          C.POINTER(LhStats)]
    L.lh_last_launch_times.argtypes = [_dp, C.POINTER(C.c_int), C.c_int]
    L.lh_host_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
    L.lh_host_free.argtypes = [C.c_void_p]
    L.lh_solve_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(LhRkOpts), C.c_void_p, _dp,
                                   C.POINTER(LhStats), C.c_void_p]
    L.lh_get_plan.argtypes = [C.c_void_p, C.POINTER(LhPlan)]
    L.lh_set_plan.argtypes = [C.c_void_p, C.POINTER(LhPlan)]
    L.lh_autotune.argtypes = [C.c_void_p, _dp, C.c_int, C.POINTER(LhPlan), C.c_int, C.POINTER(C.c_int)]
    L.lh_wisdom_export.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.lh_wisdom_import.argtypes = [C.c_char_p]
    L.lh_wisdom_forget.argtypes = []
    L.lh_set_step_stats.argtypes = [C.c_void_p, C.POINTER(LhStatsDesc)]
    L.lh_set_gradient.argtypes = [C.c_void_p, C.POINTER(LhGradientDesc)]
    L.lh_ppt_table.argtypes = [C.POINTER(LhPptDesc), C.c_int, _dp, _dp, _dp]
    L.lh_stats_eval.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, _dp]
    L.lh_get_step_stats.argtypes = [C.c_void_p, _dp, C.c_int, C.POINTER(C.c_int)]
    _fp = C.POINTER(C.c_float)
    L.lh_fno_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(LhFnoCfg)]
    L.lh_fno_destroy.argtypes = [C.c_void_p]
    L.lh_fno_param_count.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    L.lh_fno_set_params.argtypes = [C.c_void_p, _fp]
    L.lh_fno_get_params.argtypes = [C.c_void_p, _fp]
    L.lh_fno_forward.argtypes = [C.c_void_p, _fp, C.c_int, _fp]
    L.lh_fno_loss_grad.argtypes = [C.c_void_p, _fp, _fp, _fp, C.c_int, _fp, _fp]
    L.lh_fno_set_data.argtypes = [C.c_void_p, _fp, _fp, _fp, C.c_int]
    L.lh_fno_train.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int, C.c_int, C.POINTER(LhAdam), _fp, _fp]
    _lib = L
    return L


EXPORTED = ["lh_strerror", "lh_device_count", "lh_create", "lh_destroy", "lh_rhs",
            "lh_rhs_device", "lh_ctx_status", "lh_solve", "lh_solve_stream", "lh_batch_solve", "lh_batch_solve_opts", "lh_set_state",
            "lh_batch_solve_resident", "lh_get_state", "lh_get_saves", "lh_last_launch_times",
            "lh_host_alloc", "lh_host_free", "lh_solve_noalloc",
            "lh_get_plan", "lh_set_plan", "lh_autotune", "lh_wisdom_export", "lh_wisdom_import", "lh_wisdom_forget",
            "lh_set_step_stats", "lh_stats_eval", "lh_get_step_stats", "lh_set_gradient", "lh_ppt_table",
            "lh_fno_create", "lh_fno_destroy", "lh_fno_param_count", "lh_fno_set_params",
            "lh_fno_get_params", "lh_fno_forward", "lh_fno_loss_grad", "lh_fno_set_data", "lh_fno_train"]

CLASS_NAMES = (16384, 8192, 4096, "env8192", "env4096", "radial", "env16384", 32768)


def check(status, where=""):
    if status != LH_OK:
        raise LunaHipError(status, where)
