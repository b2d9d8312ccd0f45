"""CPU restatement of the spectral surrogate (1-D Fourier neural operator):
forward pass, mean-squared loss and the full parameter gradient.

Test infrastructure only: the library never calls this.  The network has no
counterpart in the reference; its definition is the one documented in
include/luna_fno.h (Li et al., ICLR 2021, FNO-1d with a truncated real DFT).

Two modes:
  * bf16=False: plain FP64 arithmetic.  Pinned by central finite differences
    of the loss (tests/test_fno_oracle.py).
  * bf16=True: every operand of the lift and Fourier-layer products is first
    rounded to BF16 (round to nearest even, via FP32), as the GPU kernels do on
    the way to the BF16 matrix cores, and products accumulate in FP64; the
    projection head stays unrounded (FP32 matrix cores on the GPU).  Used to
    check the GPU kernels; the difference between the two modes is the BF16
    effect itself.
"""
import numpy as np
from scipy.special import erf


def bf16_round(a):
    """FP64 → FP32 → BF16 (round to nearest even) → FP64."""
    a32 = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + (((u >> 16) & 1) + 0x7FFF)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


class FNOOracle:
    def __init__(self, n_x, cin, channels, modes, layers, hidden):
        self.X, self.cin, self.C, self.Mo, self.L, self.H = n_x, cin, channels, modes, layers, hidden
        C, H, Mo = channels, hidden, modes
        off = 0
        self.lay = {}

        def take(name, shape):
            nonlocal off
            n = int(np.prod(shape))
            self.lay[name] = (off, shape)
            off += n

        take("P", (C, cin))
        take("bP", (C,))
        for l in range(layers):
            take(f"W{l}", (C, C))
            take(f"b{l}", (C,))
            take(f"Rre{l}", (Mo, C, C))
            take(f"Rim{l}", (Mo, C, C))
        take("Q1", (H, C))
        take("bq1", (H,))
        take("Q2", (H,))
        take("bq2", (1,))
        self.n_params = off
        # truncated DFT matrices: F[2m] = cos, F[2m+1] = -sin; Ginv = irfft weights
        x = np.arange(n_x)
        F = np.zeros((2 * Mo, n_x))
        Gi = np.zeros((n_x, 2 * Mo))
        for m in range(Mo):
            ang = 2 * np.pi * ((m * x) % n_x) / n_x
            cm = 1.0 if m == 0 else 2.0
            F[2 * m], F[2 * m + 1] = np.cos(ang), -np.sin(ang)
            Gi[:, 2 * m], Gi[:, 2 * m + 1] = cm * np.cos(ang) / n_x, -cm * np.sin(ang) / n_x
        self.F, self.Gi = F.astype(np.float32).astype(np.float64), Gi.astype(np.float32).astype(np.float64)

    def unpack(self, p):
        return {k: np.asarray(p[o:o + int(np.prod(s))], dtype=np.float64).reshape(s) for k, (o, s) in self.lay.items()}

    def pack(self, d):
        out = np.zeros(self.n_params)
        for k, (o, s) in self.lay.items():
            out[o:o + int(np.prod(s))] = np.asarray(d[k]).reshape(-1)
        return out

    @staticmethod
    def rblk(Rre, Rim):
        return np.block([[Rre, Rim], [-Rim, Rre]])

    def forward(self, p, x, bf16=False):
        r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
        w = self.unpack(np.asarray(p, dtype=np.float32).astype(np.float64) if bf16 else p)
        B, X, C, Mo = x.shape[0], self.X, self.C, self.Mo
        F, Gi = r(self.F), r(self.Gi)
        cache = {"x": x}
        v = r(x.reshape(B * X, -1)) @ r(w["P"]).T + w["bP"]
        V, Z, Vh, Rb = [], [], [], []
        for l in range(self.L):
            vb = v.reshape(B, X, C)
            vhat = np.einsum("rx,bxc->brc", F, r(vb))                       # [B][2Mo][C]
            rbl = np.stack([self.rblk(w[f"Rre{l}"][m], w[f"Rim{l}"][m]) for m in range(Mo)])  # [Mo][2C][2C]
            vblk = vhat.reshape(B, Mo, 2 * C)
            ohat = np.einsum("bmk,mkn->bmn", r(vblk), r(rbl)).reshape(B, 2 * Mo, C)
            z = np.einsum("xr,brc->bxc", Gi, r(ohat)).reshape(B * X, C)
            z = z + r(v) @ r(w[f"W{l}"]).T + w[f"b{l}"]
            V.append(v); Z.append(z); Vh.append(vhat); Rb.append(rbl)
            v = gelu(z) if l + 1 < self.L else z
        z1 = v @ w["Q1"].T + w["bq1"]  # projection head: not rounded
        h = gelu(z1)
        y = h @ w["Q2"] + w["bq2"][0]
        cache.update(V=V, Z=Z, Vh=Vh, Rb=Rb, vL=v, z1=z1, h=h, w=w, r=r, F=F, Gi=Gi)
        return y.reshape(B, X), cache

    def loss_grad(self, p, x, t, bf16=False, wt=None):
        """Loss Σ wt (y - t)² / (B·X) (wt = 1 if None), its gradient, and y."""
        y, c = self.forward(p, x, bf16)
        r, w, F, Gi = c["r"], c["w"], c["F"], c["Gi"]
        B, X, C, Mo = x.shape[0], self.X, self.C, self.Mo
        n = B * X
        d = (y - t).reshape(-1)
        wt = np.ones(n) if wt is None else np.asarray(wt, dtype=np.float64).reshape(-1)
        loss = float(np.sum(wt * d * d) / n)
        dy = 2 * wt * d / n
        g = {}
        g["Q2"] = dy @ c["h"]
        g["bq2"] = np.array([dy.sum()])
        dz1 = dy[:, None] * w["Q2"][None, :] * gelu_d(c["z1"])
        g["Q1"] = dz1.T @ c["vL"]
        g["bq1"] = dz1.sum(0)
        dv = dz1 @ w["Q1"]
        for l in range(self.L - 1, -1, -1):
            dz = dv * gelu_d(c["Z"][l]) if l + 1 < self.L else dv
            v = c["V"][l]
            g[f"W{l}"] = r(dz).T @ r(v)
            g[f"b{l}"] = dz.sum(0)
            dohat = np.einsum("xr,bxc->brc", Gi, r(dz.reshape(B, X, C)))    # Ginvᵀ dz_b
            dob = dohat.reshape(B, Mo, 2 * C)
            vblk = c["Vh"][l].reshape(B, Mo, 2 * C)
            drb = np.einsum("bmk,bmn->mkn", r(vblk), r(dob))
            g[f"Rre{l}"] = drb[:, :C, :C] + drb[:, C:, C:]
            g[f"Rim{l}"] = drb[:, :C, C:] - drb[:, C:, :C]
            dvhat = np.einsum("bmn,mkn->bmk", r(dob), r(c["Rb"][l])).reshape(B, 2 * Mo, C)
            dv = r(dz) @ r(w[f"W{l}"]) + np.einsum("rx,brc->bxc", F, r(dvhat)).reshape(B * X, C)
        g["P"] = r(dv).T @ r(x.reshape(B * X, -1))
        g["bP"] = dv.sum(0)
        return loss, self.pack(g), y
