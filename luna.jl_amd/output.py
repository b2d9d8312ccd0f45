"""Output handlers: ``MemoryOutput``, ``HDF5Output`` and the scan variants.

Mirrors ``src/Output.jl``: ``MemoryOutput`` (``:15-141``), ``HDF5Output``
(``:146-419``), the continuation cache and ``check_cache`` (``:462-487``),
``GridCondition`` (``:490-502``), ``ScanHDF5Output``/``ScanMemoryOutput``/
``savescan`` (``:513-583``).  HDF5 files are written through
``libluna_h5.so`` (``include/luna_h5.h``) in HDF5.jl's layout: column-major
dimensions stored reversed, complex values as the compound ``{r, i}``, so
``Eω`` of a mode-averaged run is an (Nω, saveN) array here and a
[saveN][Nω] dataset in the file.

The propagation calls an output once per accepted step that writes a save
(``lh_solve_stream``): the new saves, the filtered state with its z and dz
(the cache) and the statistics records of the steps since the last call.
"""
import ctypes as C
import os
import subprocess
import warnings

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
H5_LIB_PATH = os.path.join(_HERE, "_build", "libluna_h5.so")

(H5_F64, H5_C128, H5_I64, H5_U64, H5_U8, H5_STR) = range(6)
_H5_ERR = {-1: "HDF5 library error", -2: "invalid argument", -3: "no such object",
           -4: "type or shape mismatch", -5: "buffer too small"}
_h5lib = None


class HDF5Unavailable(RuntimeError):
    pass


def h5lib():
    """Load libluna_h5.so (built when the HDF5 library was found at build time)."""
    global _h5lib
    if _h5lib is not None:
        return _h5lib
    if not os.path.isfile(H5_LIB_PATH):
        raise HDF5Unavailable(f"HDF5 output needs {H5_LIB_PATH}, which is built only where the HDF5 "
                              "C library is installed (LUNA_HDF5_ROOT, default /opt/conda)")
    L = C.CDLL(H5_LIB_PATH)
    i64p = C.POINTER(C.c_int64)
    L.lh_h5_open.argtypes = [C.c_char_p, C.c_int, i64p]
    L.lh_h5_close.argtypes = [C.c_int64]
    L.lh_h5_exists.argtypes = [C.c_int64, C.c_char_p]
    L.lh_h5_create_group.argtypes = [C.c_int64, C.c_char_p]
    L.lh_h5_delete.argtypes = [C.c_int64, C.c_char_p]
    L.lh_h5_write.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_int, i64p, C.c_void_p]
    L.lh_h5_create_extendable.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_int, i64p, i64p]
    L.lh_h5_write_slab.argtypes = [C.c_int64, C.c_char_p, C.c_int64, C.c_void_p]
    L.lh_h5_info.argtypes = [C.c_int64, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), i64p]
    L.lh_h5_read.argtypes = [C.c_int64, C.c_char_p, C.c_void_p, C.c_int64]
    L.lh_h5_list.argtypes = [C.c_int64, C.c_char_p, C.c_char_p, C.c_int64, C.POINTER(C.c_int)]
    _h5lib = L
    return L


def _check(st, what):
    if st < 0:
        raise OSError(f"{what}: {_H5_ERR.get(st, st)}")
    return st


def _kind(a):
    if a.dtype == np.float64:
        return H5_F64
    if a.dtype == np.complex128:
        return H5_C128
    if a.dtype == np.uint64:
        return H5_U64
    if a.dtype == np.bool_ or a.dtype == np.uint8:
        return H5_U8
    if np.issubdtype(a.dtype, np.integer):
        return H5_I64
    raise TypeError(f"no HDF5 mapping for {a.dtype}")


_NP = {H5_F64: np.float64, H5_C128: np.complex128, H5_I64: np.int64, H5_U64: np.uint64, H5_U8: np.uint8}


def _dims(shape):
    return (C.c_int64 * max(len(shape), 1))(*shape)


class H5File:
    """An open HDF5 file (``HDF5.h5open``); mode "r", "r+" or "w"."""

    def __init__(self, path, mode="r"):
        self.L = h5lib()
        self.path = path
        h = C.c_int64()
        _check(self.L.lh_h5_open(path.encode(), {"r": 0, "r+": 1, "w": 2}[mode], C.byref(h)), f"open {path}")
        self.h = h.value

    def close(self):
        if self.h is not None:
            self.L.lh_h5_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __contains__(self, path):
        return _check(self.L.lh_h5_exists(self.h, path.encode()), f"exists {path}") == 1

    def create_group(self, path):
        _check(self.L.lh_h5_create_group(self.h, path.encode()), f"create_group {path}")

    def delete(self, path):
        _check(self.L.lh_h5_delete(self.h, path.encode()), f"delete {path}")

    def keys(self, path="/"):
        buf = C.create_string_buffer(1 << 16)
        n = C.c_int()
        _check(self.L.lh_h5_list(self.h, path.encode(), buf, len(buf), C.byref(n)), f"keys {path}")
        return [s.decode() for s in buf.raw.split(b"\0")[:n.value]]

    def is_group(self, path):
        t, nd = C.c_int(), C.c_int()
        d = (C.c_int64 * 8)()
        st = self.L.lh_h5_info(self.h, path.encode(), C.byref(t), C.byref(nd), d)
        return st == -4 and path in self

    def info(self, path):
        t, nd = C.c_int(), C.c_int()
        d = (C.c_int64 * 8)()
        _check(self.L.lh_h5_info(self.h, path.encode(), C.byref(t), C.byref(nd), d), f"info {path}")
        return t.value, tuple(d[i] for i in range(nd.value))

    def write(self, path, value):
        """``parent[name] = value`` / ``write(parent[name], value)``; a dict
        becomes a group of datasets."""
        if isinstance(value, dict):
            for k, v in value.items():
                self.write(f"{path}/{k}", v)
            return
        if isinstance(value, str):
            _check(self.L.lh_h5_write(self.h, path.encode(), H5_STR, 0, None, value.encode()), f"write {path}")
            return
        if isinstance(value, (list, tuple)) and value and all(isinstance(v, str) for v in value):
            arr = (C.c_char_p * len(value))(*[v.encode() for v in value])
            _check(self.L.lh_h5_write(self.h, path.encode(), H5_STR, 1, _dims((len(value),)), arr),
                   f"write {path}")
            return
        a = np.asarray(value)
        if a.dtype == np.bool_:
            a = a.astype(np.uint8)
        elif a.dtype.kind == "f":
            a = a.astype(np.float64)
        elif a.dtype.kind == "c":
            a = a.astype(np.complex128)
        k = _kind(a)
        # (ascontiguousarray would turn a 0-d array into shape (1,))
        a = np.asfortranarray(a, dtype=_NP[k]) if a.ndim else np.array(a, dtype=_NP[k])
        _check(self.L.lh_h5_write(self.h, path.encode(), k, a.ndim, _dims(a.shape), a.ctypes.data), f"write {path}")

    def create_extendable(self, path, kind, dims, chunk):
        _check(self.L.lh_h5_create_extendable(self.h, path.encode(), kind, len(dims), _dims(dims), _dims(chunk)),
               f"create_dataset {path}")

    def write_slab(self, path, index, value):
        t, dims = self.info(path)
        a = np.asfortranarray(np.asarray(value, dtype=_NP[t]).reshape(dims[:-1], order="F"))
        _check(self.L.lh_h5_write_slab(self.h, path.encode(), index, a.ctypes.data), f"write_slab {path}")

    def read(self, path):
        """``read(file[path])``: an array in column-major shape, a scalar, a
        string, a list of strings, or a dict for a group."""
        if path == "/" or self.is_group(path):
            return {k: self.read(f"{path.rstrip('/')}/{k}") for k in self.keys(path)}
        t, dims = self.info(path)
        if t == H5_STR:
            n = int(np.prod(dims)) if dims else 1
            buf = C.create_string_buffer(1 << 22)
            _check(self.L.lh_h5_read(self.h, path.encode(), buf, len(buf)), f"read {path}")
            parts = [p.decode() for p in buf.raw.split(b"\0")[:n]]
            return parts[0] if not dims else parts
        out = np.empty(dims, dtype=_NP[t], order="F")
        _check(self.L.lh_h5_read(self.h, path.encode(), out.ctypes.data, out.nbytes), f"read {path}")
        return out[()] if not dims else out


def git_commit():
    """``Utils.git_commit``: the commit of the source tree, or "unavailable"."""
    try:
        r = subprocess.run(["git", "-C", _HERE, "rev-parse", "HEAD"], capture_output=True, text=True, timeout=10)
        return r.stdout.strip() if r.returncode == 0 and r.stdout.strip() else "unavailable"
    except (OSError, subprocess.SubprocessError):
        return "unavailable"


def sourcecode():
    """``Utils.sourcecode``: the text of the package's sources, file by file."""
    parts = []
    for sub in ("", "csrc"):
        d = os.path.join(_HERE, sub)
        for name in sorted(os.listdir(d)):
            if name.endswith((".py", ".hip", ".h", ".c")):
                with open(os.path.join(d, name), encoding="utf-8") as f:
                    parts.append(f"# ---- {os.path.join(sub, name)}\n{f.read()}")
    return "\n".join(parts)


def save_grid(tmin, tmax, saveN):
    """``GridCondition(tmin, tmax, saveN).grid``: ``range(tmin, tmax, length=saveN)``."""
    if saveN == 1:
        return np.array([float(tmin)])
    i = np.arange(saveN, dtype=np.longdouble)
    a, b = np.longdouble(tmin), np.longdouble(tmax)
    return ((a * (saveN - 1 - i) + b * i) / (saveN - 1)).astype(np.float64)


def cachehash(statsnames, ydims):
    """The guard against continuing a different propagation from a cache: the
    reference stores Julia's ``hash((sort(statsnames), size(y)))``
    (src/Output.jl:228-229, 286-288); this is FNV-1a over the same tuple, so
    equal tuples give equal hashes (the value itself differs from Julia's)."""
    h = 0xcbf29ce484222325
    for b in ("\0".join(sorted(statsnames)) + "|" + ",".join(str(int(d)) for d in ydims)).encode():
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return np.uint64(h)


class _OutputBase:
    yname = "Eω"
    tname = "z"

    def _stats_names(self):
        if self.statsfun is None:
            return []
        if hasattr(self.statsfun, "names"):
            return self.statsfun.names()
        return list(self.statsfun.to_dict(np.zeros((0, _nstat()))).keys())

    def check_cache(self, y, t, dt):
        """No cache outside HDF5Output: ``check_cache(o, y, t, dt) = y, t, dt``."""
        return y, t, dt


def _per_step_last(v):
    """Values of n steps, (n,) or (n, k...), as an array whose last dimension
    is the step (Julia's layout of a statistic over a propagation)."""
    v = np.asarray(v)
    if v.dtype.kind in "iub":
        v = v.astype(np.int64) if v.dtype.kind != "b" else v.astype(np.uint8)
    elif v.dtype.kind == "c":
        v = v.astype(np.complex128)
    else:
        v = v.astype(np.float64)
    return np.moveaxis(v, 0, -1) if v.ndim > 1 else v


def _nstat():
    from . import _capi
    return _capi.LH_NSTAT


class MemoryOutput(_OutputBase):
    """``Output.MemoryOutput(tmin, tmax, saveN, statsfun)`` with
    ``GridCondition`` saves (src/Output.jl:15-141)."""

    def __init__(self, tmin, tmax, saveN, statsfun=None, yname="Eω", tname="z", script=None):
        self.tmin, self.tmax, self.saveN = tmin, tmax, saveN
        self.save_cond = save_grid(tmin, tmax, saveN)
        self.statsfun = statsfun
        self.yname, self.tname = yname, tname
        self.saved = 0
        self.data = {"stats": {}, "meta": {"sourcecode": sourcecode(), "git_commit": git_commit()}}
        if script is not None:
            self.data["meta"]["script_code"] = script

    def __getitem__(self, k):
        k = k if isinstance(k, tuple) else (k,)
        v = self.data
        while k and isinstance(k[0], str):
            v = v[k[0]]
            k = k[1:]
        return v[k] if k else v

    def __contains__(self, k):
        return k in self.data

    def __call__(self, key, val=None, force=False, meta=False, group=None):
        """``output(dict; force, meta, group)`` / ``output(key, val; ...)``
        (src/Output.jl:104-130): a key already present raises unless force."""
        if isinstance(key, dict):
            for k, v in key.items():
                self(k, v, force=force, meta=meta, group=group)
            return
        parent = self.data["meta"] if meta else self.data
        target = parent.setdefault(group, {}) if group is not None else parent
        if key in target:
            if not force:
                raise KeyError(f"Key {key} already present in dataset.")
            warnings.warn(f"Key {key} already exists and will be overwritten.")
        target[key] = val

    # the streamed form of the per-step call (Output.jl:65-81)
    def on_step(self, first_save, saves, zs, z, dz, y, records):
        if self.yname not in self.data:
            self.data[self.yname] = np.zeros(y.shape + (self.saveN,), dtype=np.complex128, order="F")
            self.data[self.tname] = np.zeros(self.saveN)
        for j in range(saves.shape[-1]):
            self.data[self.yname][..., first_save + j] = saves[..., j]
            self.data[self.tname][first_save + j] = zs[j]
        self.saved = first_save + saves.shape[-1]
        self._append_stats(records)

    def _append_stats(self, records):
        """``append_stats!`` (src/Output.jl:83-102): a scalar statistic grows a
        vector, an array-valued one an array with one more (last) dimension."""
        if self.statsfun is None or len(records) == 0:
            return
        for k, v in self.statsfun.to_dict(records).items():
            v = _per_step_last(v)
            old = self.data["stats"].get(k)
            self.data["stats"][k] = v if old is None else np.concatenate([old, v], axis=-1)

    def tofile(self, fpath):
        """``Output.tofile``: write everything to an HDF5 file (``Utils.save_dict_h5``)."""
        with H5File(fpath, "w") as f:
            for k, v in self.data.items():
                f.write(k, v)


class HDF5Output(_OutputBase):
    """``Output.HDF5Output(fpath, tmin, tmax, saveN, statsfun; yname, tname,
    compression, script, cache, readonly)`` (src/Output.jl:146-419): the file
    grows by one z-slice per save; ``meta/cache/{t, dt, y, saved}`` holds the
    latest saving step so an interrupted run continues from it."""

    def __init__(self, fpath, tmin=0.0, tmax=0.0, saveN=1, statsfun=None, yname="Eω", tname="z",
                 compression=False, script=None, cache=True, readonly=False):
        h5lib()
        if compression:
            raise NotImplementedError("blosc compression (H5Zblosc) is not available")
        self.fpath = fpath
        self.tmin, self.tmax, self.saveN = tmin, tmax, saveN
        self.save_cond = save_grid(tmin, tmax, saveN)
        self.statsfun = statsfun
        self.yname, self.tname = yname, tname
        self.cache, self.readonly = cache, readonly
        self.compression = compression
        self.saved = 0
        self.cachehash = np.uint64(0)
        if os.path.isfile(fpath) and cache:
            with H5File(fpath, "r+") as f:
                if "meta/cache" not in f:
                    raise RuntimeError("cached HDF5Output created, file exists, but has no cache")
                if "meta/cache/saved" in f and yname in f:
                    self.saved = int(f.read("meta/cache/saved"))
                    self.cachehash = cachehash(f.keys("stats"), f.info(yname)[1][:-1])
        elif not readonly:
            if os.path.isfile(fpath):
                warnings.warn(f"output file {fpath} already exists and will be overwritten!")
                os.remove(fpath)
            d = os.path.dirname(fpath)
            if d:
                os.makedirs(d, exist_ok=True)
            with H5File(fpath, "w") as f:
                f.create_group("stats")
                f.create_group("meta")
                f.write("meta/sourcecode", sourcecode())
                f.write("meta/git_commit", git_commit())
                if script is not None:
                    f.write("meta/script_code", script)
                if cache:
                    f.create_group("meta/cache")

    @classmethod
    def open(cls, fpath):
        """``HDF5Output(fpath)``: read-only access to an existing file."""
        if not os.path.isfile(fpath):
            raise FileNotFoundError("Cannot open read-only HDF5Output: file not found")
        return cls(fpath, 0, 0, 1, readonly=True)

    def __getitem__(self, idx):
        """``o[name]`` reads a whole dataset or group; ``o[name, I...]`` a slice of it."""
        idx = idx if isinstance(idx, tuple) else (idx,)
        names = []
        while idx and isinstance(idx[0], str):  # group members by name (src/Output.jl:241-247)
            names.append(idx[0])
            idx = idx[1:]
        with H5File(self.fpath, "r") as f:
            v = f.read("/".join(names))
        return v[idx] if idx else v

    def __contains__(self, key):
        if not os.path.isfile(self.fpath):
            return False
        with H5File(self.fpath, "r") as f:
            return key in f

    def _initialise(self, f, y):
        dims = tuple(y.shape) + (self.saveN,)
        f.create_extendable(self.yname, H5_C128, dims, tuple(y.shape) + (1,))
        f.create_extendable(self.tname, H5_F64, (self.saveN,), (1,))
        self.cachehash = cachehash(self._stats_names(), y.shape)
        f.write("meta/cachehash", self.cachehash)
        if self.cache:
            f.write("meta/cache/t", -np.inf)
            f.write("meta/cache/dt", -np.inf)
            f.write("meta/cache/y", y)
            f.write("meta/cache/saved", np.int64(0))

    def on_step(self, first_save, saves, zs, z, dz, y, records):
        """The output call of an accepted step that saves (src/Output.jl:262-302)."""
        if self.readonly:
            raise RuntimeError("Cannot add data to read-only output!")
        with H5File(self.fpath, "r+") as f:
            if self.yname not in f:
                self._initialise(f, y)
            if cachehash(self._stats_names(), y.shape) != self.cachehash:
                raise RuntimeError("the hash for this propagation does not agree with cache in file")
            for j in range(saves.shape[-1]):
                f.write_slab(self.yname, first_save + j, saves[..., j])
                f.write_slab(self.tname, first_save + j, np.array([zs[j]]))
            self.saved = first_save + saves.shape[-1]
            if self.statsfun is not None and len(records):
                # append_stats! (src/Output.jl:305-337): one chunk per step
                for k, v in self.statsfun.to_dict(records).items():
                    v = _per_step_last(v)
                    path = f"stats/{k}"
                    shape = v.shape[:-1]
                    if path not in f:
                        f.create_extendable(path, _kind(v), shape + (1,), shape + (1,))
                        n0 = 0
                    else:
                        n0 = f.info(path)[1][-1]
                    for i in range(v.shape[-1]):
                        f.write_slab(path, n0 + i, v[..., i])
            if self.cache:
                f.write("meta/cache/t", float(z))
                f.write("meta/cache/dt", float(dz))
                f.write("meta/cache/y", y)
                f.write("meta/cache/saved", np.int64(self.saved))

    def __call__(self, key, val=None, force=False, meta=False, group=None):
        """``output(dict; force, meta, group)`` / ``output(key, val; ...)``
        (src/Output.jl:345-419): an existing dataset is replaced with a warning
        when force, else rewritten in place after a warning."""
        if self.readonly:
            raise RuntimeError("Cannot add data to read-only output!")
        items = key.items() if isinstance(key, dict) else [(key, val)]
        with H5File(self.fpath, "r+") as f:
            parent = "meta" if meta else ""
            for k, v in items:
                path = "/".join(p for p in (parent, group, k) if p)
                if path in f:
                    if force:
                        warnings.warn(f"Dataset {k} already present in file {self.fpath} and will be overwritten")
                        f.delete(path)
                    else:
                        warnings.warn(f"File {self.fpath} already has dataset {k}. Pass force=true to overwrite")
                f.write(path, v)

    def check_cache(self, y, t, dt):
        """``check_cache`` (src/Output.jl:473-484): the cached (y, t, dt) when
        the cache is ahead of t, else the arguments."""
        if not self.cache or not os.path.isfile(self.fpath):
            return y, t, dt
        with H5File(self.fpath, "r") as f:
            if "meta/cache/t" not in f:
                return y, t, dt
            tc = float(f.read("meta/cache/t"))
            if tc < t:
                return y, t, dt
            return f.read("meta/cache/y"), tc, float(f.read("meta/cache/dt"))


def makefilename(name, scanidx):
    """``Scans.makefilename``: ``"<name>_<scanidx:05d>.h5"``."""
    return "%s_%05d.h5" % (getattr(name, "name", name), scanidx)


def savescan(out, scan, scanidx):
    """``savescan`` (src/Output.jl:562-583): the scan index, the values of the
    scan variables at it, the scan arrays, their order and shape, under meta."""
    out("scanidx", scanidx, meta=True)
    values = scan.values(scanidx)
    out({str(v): x for v, x in zip(scan.variables, values)}, meta=True, group="scanvars")
    out({str(v): np.asarray(a) for v, a in zip(scan.variables, scan.arrays)}, meta=True, group="scanarrays")
    out("scanorder", [str(v) for v in scan.variables], meta=True)
    out("scanshape", np.array([len(a) for a in scan.arrays], dtype=np.int64), meta=True)
    return out


def ScanHDF5Output(scan, scanidx, *args, fname=None, fdir=None, **kw):
    """``ScanHDF5Output`` (src/Output.jl:513-527)."""
    fpath = makefilename(scan if fname is None else fname, scanidx)
    if fdir is not None:
        fpath = os.path.join(fdir, fpath)
        os.makedirs(fdir, exist_ok=True)
    return savescan(HDF5Output(fpath, *args, **kw), scan, scanidx)


def ScanMemoryOutput(scan, scanidx, *args, **kw):
    return savescan(MemoryOutput(*args, **kw), scan, scanidx)
