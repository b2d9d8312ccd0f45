"""Constant linear operators, restating ``src/LinearOps.jl``.

Mode-averaged (``make_const_linop`` ``LinearOps.jl:256-319`` with ``αlim!``
``:237-241``) and radial free-space (``:130-165`` with ``_fill_linop_r!``
``:173-184``).  The operator is built once on the host and uploaded with the
transform constants.
"""
import numpy as np

from . import physdata as pd


def make_const_linop(grid, mode, ref_lambda):
    """Returns (linop, β array used by the norm, β1)."""
    omega = grid.omega
    s = grid.sidx
    beta1 = mode.dispersion1(pd.wlfreq(ref_lambda))
    beta = np.zeros_like(omega)
    beta[s] = mode.beta(omega[s])
    beta[~s] = 1.0
    alpha = np.zeros_like(omega)
    alpha[s] = mode.alpha(omega[s])
    alpha = np.clip(alpha, 0.0, 3000.0)
    if grid.kind == "real":
        im = -(beta - beta1 * omega)
    else:
        beta0 = float(mode.beta(np.array([pd.wlfreq(ref_lambda)]))[0])
        im = -(beta - beta1 * (omega - grid.omega0) - beta0)
    linop = (-alpha / 2) + 1j * im
    linop[~s] = 0
    return linop, beta, beta1


def make_const_linop_radial(grid, q, nfun_lambda, beta1=None):
    """``make_const_linop(grid, q::QDHT, nfun)`` for the field grid
    (``LinearOps.jl:130-144, 173-184``) and the envelope grid (``:146-164,
    211-225``)."""
    omega = grid.omega
    n = np.zeros_like(omega)
    s = grid.sidx
    n[s] = np.real(nfun_lambda(2 * np.pi * pd.c / omega[s]))
    if beta1 is None:
        from .modes import derivative1
        beta1 = derivative1(lambda w: w / pd.c * float(np.real(nfun_lambda(pd.wlfreq(w)))),
                            pd.wlfreq(grid.ref_lambda))
    # envelope grid (LinearOps.jl:146-164, 211-225): the same fill over the
    # absolute ω, less -iβref with βref = ω0/c·n(λ0) (thg = false)
    bref = 0.0
    if grid.kind != "real":
        bref = grid.omega0 / pd.c * float(np.real(nfun_lambda(2 * np.pi * pd.c / grid.omega0)))
    k2 = (n * omega / pd.c) ** 2
    kr2 = q.k ** 2
    out = np.empty((len(omega), q.N), dtype=np.complex128, order="F")
    for ir in range(q.N):
        bsq = k2 - kr2[ir]
        neg = bsq < 0
        col = np.empty(len(omega), dtype=np.complex128)
        col[neg] = -1j * (-beta1 * omega[neg]) - np.minimum(np.sqrt(np.abs(bsq[neg])), 200)
        col[~neg] = -1j * (np.sqrt(bsq[~neg]) - beta1 * omega[~neg])
        if bref != 0.0:
            col -= -1j * bref
        out[:, ir] = col
    return out, beta1
