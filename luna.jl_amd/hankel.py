"""Quasi-discrete Hankel transform of order 0: the ``Hankel.QDHT`` the radial
transform is built on.

``Hankel`` is an external Julia package (Project.toml pins ``Hankel = "0.5.5"``;
its source is not part of the reference).  Its published algorithm is
Guizar-Sicairos & Gutiérrez-Vega, J. Opt. Soc. Am. A 21, 53 (2004):
with j_n the zeros of J0 and S = j_{N+1},

    r_n = j_n R/S,   k_n = j_n/R,   K = S/R,
    T_nm = (2/S) J0(j_n j_m/S) / J1(j_m)^2,
    forward  (r → k, ``mul!``):  Y = (R/K) T A,
    inverse  (k → r, ``ldiv!``): A = (K/R) T Y,
    integrateR(A) = Σ_n (2/K²)/J1(j_n)² A_n,   integrateK(Y) = Σ_n (2/R²)/J1(j_n)² Y_n.

Call sites in the reference: ``src/NonlinearRHS.jl:466-474`` (both products),
``src/Fields.jl:380, 536-549`` (input transform and energies),
``src/LinearOps.jl:130-144`` (``q.k``).  ``dim`` selects the radial axis as in
``QDHT(R, N, dim=2)``.
"""
import numpy as np
from scipy import special


def bessel_j0_zeros(n):
    """First n positive zeros of J0, Newton-polished (J0' = -J1)."""
    z = special.jn_zeros(0, n).astype(float)
    for _ in range(3):
        z = z + special.j0(z) / special.j1(z)
    return z


class QDHT:
    def __init__(self, R, N, dim=2):
        roots = bessel_j0_zeros(N + 1)
        j, S = roots[:N], roots[N]
        self.R, self.N, self.S, self.dim = float(R), int(N), float(S), dim
        self.r = j * R / S
        self.K = S / R
        self.k = j * self.K / S
        J1sq = special.j1(j) ** 2
        self.J1sq = J1sq
        self.T = (2.0 / S) * special.j0(np.outer(j, j) / S) / J1sq[None, :]
        self.scaleR = 2.0 / self.K ** 2 / J1sq
        self.scaleK = 2.0 / self.R ** 2 / J1sq

    def _dot(self, A):
        ax = self.dim - 1
        A = np.moveaxis(np.asarray(A), ax, 0)
        out = np.tensordot(self.T, A, axes=(1, 0))
        return np.moveaxis(out, 0, ax)

    def forward(self, A):
        """``q * A``: real space → reciprocal space."""
        return self._dot(A) * (self.R / self.K)

    def inverse(self, Y):
        """``q \\ Y``: reciprocal space → real space."""
        return self._dot(Y) * (self.K / self.R)

    def integrateR(self, A, axis=0):
        return np.tensordot(self.scaleR, np.moveaxis(np.asarray(A), axis, 0), axes=(0, 0))

    def integrateK(self, Y, axis=0):
        return np.tensordot(self.scaleK, np.moveaxis(np.asarray(Y), axis, 0), axes=(0, 0))
