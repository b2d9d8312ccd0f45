"""Parameter scans: ``Scans.Scan``, ``runscan`` and the batched form.

Mirrors ``src/Scans.jl``: a ``Scan`` holds named variables and their arrays;
its points are ``Iterators.product`` of the arrays (the first variable varies
fastest) numbered from 1 (``:120-175, 245-249``); ``runscan`` runs a function
at every point of the execution mode's share and isolates each point's
failure (``:275-311``); ``LocalExec``, ``RangeExec`` and ``BatchExec`` with
``chunks`` (``:14-44, 313-383``).  The file queue and cluster back-ends are
not provided.

``runscan_batched`` is the GPU form of the same loop: every point of the
share is built on the host, all of them are solved in one batched launch per
grid class (``lh_batch_solve_opts``, each point with its own options, step
control and status), and each point's output is then filled; a point whose
set-up raises or whose solve fails is reported and skipped, as ``runscan``'s
try/catch does.
"""
import itertools
import warnings

import numpy as np

from . import _capi
from .output import makefilename  # noqa: F401 (Scans.makefilename)


class LocalExec:
    """The whole scan in this process."""


class RangeExec:
    """The points ``r`` (1-based, inclusive, as a Julia ``UnitRange``)."""

    def __init__(self, r):
        self.r = r


class BatchExec:
    """Batch ``batch`` (1-based) of ``nbatches`` interleaved chunks."""

    def __init__(self, nbatches, batch):
        self.nbatches, self.batch = nbatches, batch


class Scan:
    """``Scan(name; variables...)`` / ``Scan(name, exec; variables...)``."""

    def __init__(self, name, exec=None, **variables):
        self.name = name
        self.exec = exec if exec is not None else LocalExec()
        self.variables = list(variables.keys())
        self.arrays = [np.asarray(a) for a in variables.values()]

    def addvariable(self, variable=None, array=None, **kw):
        """``addvariable!(scan, variable, array)`` / ``addvariable!(scan; kwargs...)``."""
        if variable is not None:
            self.variables.append(variable)
            self.arrays.append(np.asarray(array))
        for v, a in kw.items():
            self.variables.append(v)
            self.arrays.append(np.asarray(a))

    def __len__(self):
        return 0 if not self.arrays else int(np.prod([len(a) for a in self.arrays]))

    def size(self):
        return (0,) if not self.arrays else tuple(len(a) for a in self.arrays)

    def points(self):
        """All points in ``vec(collect(Iterators.product(arrays...)))`` order."""
        rev = itertools.product(*[list(a) for a in reversed(self.arrays)])
        return [tuple(reversed(p)) for p in rev]

    def values(self, scanidx):
        """The variables' values at the 1-based ``scanidx``."""
        return self.points()[scanidx - 1]

    def getvalue(self, variable, scanidx):
        """``getvalue(scan, variable, scanidx)``."""
        return self.values(scanidx)[self.variables.index(variable)]


def chunks(a, n):
    """``Scans.chunks``: entries dealt round-robin into ``n`` lists."""
    out = [[] for _ in range(n)]
    for i, x in enumerate(a):
        out[i % n].append(x)
    return out


def _share(scan):
    """(scanidx, args) pairs of the execution mode's share, numbered as the
    reference's runscan numbers them."""
    pts = scan.points()
    ex = scan.exec
    if isinstance(ex, LocalExec):
        return list(enumerate(pts, start=1))
    if isinstance(ex, RangeExec):
        lo, hi = ex.r[0], ex.r[-1]
        # runscan(f, ::Scan{RangeExec}) enumerates the sub-range from 1
        return list(enumerate(pts[lo - 1:hi], start=1))
    if isinstance(ex, BatchExec):
        idcs = chunks(list(range(1, len(pts) + 1)), ex.nbatches)[ex.batch - 1]
        return [(i, pts[i - 1]) for i in idcs]
    raise TypeError(f"execution mode {type(ex).__name__} is not provided")


def _warn(scanidx, e):
    warnings.warn(f"Error at scanidx {scanidx}:\n{type(e).__name__}: {e}", RuntimeWarning)


def runscan(f, scan):
    """``runscan(f, scan)``: ``f(scanidx, args...)`` at every point of the
    share; a point that raises is reported and skipped.  Returns the results."""
    out = []
    for scanidx, args in _share(scan):
        try:
            out.append(f(scanidx, *args))
        except Exception as e:  # noqa: BLE001 -- the reference's try/catch around each point
            _warn(scanidx, e)
    return out


def runscan_batched(scan, make_point, rtol=1e-6, atol=1e-10, safety=0.9, init_dz=1e-4, min_dz=0.0,
                    max_dz=None):
    """Every point of the share in one batched solve.  ``make_point(scanidx,
    *args)`` returns ``(Eω, grid, linop, transform, FT, output)`` as
    ``prop_capillary_args`` does, optionally followed by a dict of that
    point's ``Luna.run`` solver options (rtol, atol, safety, init_dz, min_dz,
    max_dz) over the defaults given here; the output's save grid and
    statistics are honoured per point.  Returns ``{scanidx: output}`` of the
    points that ran to the end; failures are reported (RuntimeWarning) and
    left out."""
    defaults = dict(rtol=rtol, atol=atol, safety=safety, init_dz=init_dz, min_dz=min_dz, max_dz=max_dz)
    built = []
    for scanidx, args in _share(scan):
        try:
            res = make_point(scanidx, *args)
            Ew, grid, linop, transform, _fT, output = res[:6]
            kw = dict(defaults)
            kw.update(res[6] if len(res) > 6 else {})
            if kw["max_dz"] is None:
                kw["max_dz"] = grid.zmax / 2
            ctx = transform.context(linop)
            ctx.set_step_stats(output.statsfun)
            opts = _capi.rk_opts(grid.zmax, save_n=output.saveN, save_z0=output.tmin, **kw)
            built.append((scanidx, Ew, grid, linop, transform, output, ctx, opts))
        except Exception as e:  # noqa: BLE001
            _warn(scanidx, e)
    if not built:
        return {}
    from .runtime import batch_solve_opts
    ctxs = [b[6] for b in built]
    outs, saves, zss, stats = batch_solve_opts(ctxs, [b[1] for b in built], [b[7] for b in built])
    done = {}
    from .interface import simtype
    for (scanidx, Ew, grid, linop, transform, output, ctx, opts), Ew_out, sv, zs, st in \
            zip(built, outs, saves, zss, stats):
        try:
            _capi.check(st.status, f"scan point {scanidx}")
            shape = (grid.n_omega,) + ((transform.n_r,) if transform.n_r > 1 else ())
            n = opts.save_n
            sv = sv.reshape(n, -1, grid.n_omega).transpose(2, 1, 0).reshape(shape + (n,), order="F")
            rec = ctx.step_stats() if output.statsfun is not None else np.zeros((0, _capi.LH_NSTAT))
            output(grid.to_dict(), group="grid")
            output(simtype(grid, transform, linop), group="simulation_type")
            output.on_step(0, sv, zs[:n], st.z, st.dz, Ew_out.reshape(shape, order="F"), rec)
            output("solver", st.as_dict(), meta=True, force=True)
            done[scanidx] = output
        except Exception as e:  # noqa: BLE001
            _warn(scanidx, e)
    return done
