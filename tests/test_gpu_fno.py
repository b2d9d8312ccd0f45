"""GPU kernels of the spectral surrogate against the CPU restatement.

Forward output and every parameter gradient are compared with the oracle in
its BF16-operand mode (same rounding of every matrix-product operand, FP64
accumulation): the remaining differences are FP32 accumulation order.  The
distance to the plain FP64 oracle (the BF16 effect itself) is printed.
"""
import numpy as np
import pytest

from oracle import fno_oracle  # noqa: E402

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - np.asarray(b, float)) / np.linalg.norm(np.asarray(b, float)))


# BF16 rounding of every product operand makes the network's output a
# discontinuous function of its FP32 intermediates: perturbing them by 1e-7
# (FP32 accumulation order) moves the BF16-mode oracle's output by ~5e-4
# (shallow net) and ~3e-3 (full net), and its gradients by ~2e-3 / ~5e-3.
# The tolerances below are about four times those floors; a wrong index,
# sign or operand changes results at O(1).
CASES = {
    "shallow": (dict(n_x=256, cin=4, channels=32, modes=8, layers=1, hidden=32), 4, 50.0, 2e-3, 8e-3),
    "full": (dict(n_x=1024, cin=4, channels=64, modes=32, layers=4, hidden=128), 3, 200.0, 1.2e-2, 2e-2),
}


@pytest.fixture(scope="module", params=list(CASES))
def setup(pkg, request):
    from luna_jl_amd import surrogate
    dims, B, rscale, tol_f, tol_g = CASES[request.param]
    net = surrogate.FNO(max_batch=4, **dims)
    p = net.init_params(seed=0)
    for name, (o, shape) in net.layout.items():  # make the spectral path carry weight
        if name.startswith("R"):
            n = int(np.prod(shape))
            p[o:o + n] *= rscale
    net.set_params(p)
    orc = fno_oracle.FNOOracle(dims["n_x"], dims["cin"], dims["channels"], dims["modes"], dims["layers"],
                               dims["hidden"])
    assert orc.n_params == net.n_params
    rng = np.random.default_rng(7)
    x = rng.standard_normal((B, dims["n_x"], dims["cin"])).astype(np.float32)
    t = (0.5 * rng.standard_normal((B, dims["n_x"]))).astype(np.float32)
    return request.param, net, orc, p, x, t, tol_f, tol_g


def test_fno_forward_parity(setup):
    name, net, orc, p, x, t, tol_f, tol_g = setup
    print(f"case: fno forward parity ({name})", flush=True)
    y = net.forward(x)
    yb, _ = orc.forward(p.astype(np.float64), x.astype(np.float64), bf16=True)
    y64, _ = orc.forward(p.astype(np.float64), x.astype(np.float64), bf16=False)
    e_b, e_64 = rel(y, yb), rel(y, y64)
    print("forward rel err vs bf16 oracle", e_b, "vs fp64 oracle", e_64, flush=True)
    assert e_b <= tol_f
    assert e_64 <= 3 * tol_f


def test_fno_gradient_parity(setup):
    name, net, orc, p, x, t, tol_f, tol_g = setup
    print(f"case: fno gradient parity ({name})", flush=True)
    wt = np.random.default_rng(9).uniform(0.1, 2.0, t.shape).astype(np.float32)  # weighted loss
    loss, g = net.loss_grad(x, t, wt)
    args = (p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    lb, gb, _ = orc.loss_grad(*args, bf16=True, wt=wt)
    l64, g64, _ = orc.loss_grad(*args, bf16=False, wt=wt)
    print("loss gpu / bf16 oracle / fp64 oracle", loss, lb, l64, flush=True)
    assert abs(loss - lb) <= 1e-3 * abs(lb)
    worst = 0.0
    for pname, (o, shape) in net.layout.items():
        n = int(np.prod(shape))
        e = rel(g[o:o + n], gb[o:o + n])
        e64 = rel(g[o:o + n], g64[o:o + n])
        print(f"grad {pname}: vs bf16 oracle {e:.2e}, vs fp64 oracle {e64:.2e}", flush=True)
        worst = max(worst, e)
    assert worst <= tol_g


def test_fno_training_reduces_loss(pkg):
    print("case: fno training", flush=True)
    from luna_jl_amd import surrogate
    net = surrogate.FNO(max_batch=16)
    net.set_params(net.init_params(seed=1))
    rng = np.random.default_rng(3)
    n = 128
    a = rng.uniform(-1, 1, n)
    xs = np.linspace(-1, 1, net.dims["n_x"])
    x = np.zeros((n, net.dims["n_x"], 4), np.float32)
    x[:, :, 0] = np.exp(-xs ** 2 * 4)[None, :]
    x[:, :, 1] = a[:, None]
    x[:, :, 3] = xs[None, :]
    t = (np.exp(-(xs[None, :] - 0.3 * a[:, None]) ** 2 * 8) * (1 + 0.5 * a[:, None])).astype(np.float32)
    net.set_data(x, t)
    import time
    t0 = time.perf_counter()
    losses = net.train(400, batch=16, lr=2e-3, seed=0)
    print("train 400 steps of batch 16:", time.perf_counter() - t0, "s", flush=True)
    print("loss first/last", losses[:5].mean(), losses[-20:].mean(), flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.05 * losses[:5].mean()
