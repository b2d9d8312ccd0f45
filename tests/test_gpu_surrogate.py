"""End-to-end screening with the spectral surrogate (no counterpart in the
reference): the FNO is trained on exit spectra that the batched FP64 solver
produces for capillary scan points, calibrated on held-out points, and then
screens a fresh sweep.  The gain over the method without it is the ratio of
solver calls (every point vs the fallback points only); every sweep point is
also solved, to check that no accepted point is above the stated bound.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOUND = 1e-2


def test_surrogate_screening_gain(pkg):
    # the full configuration: 4096 training and 512 calibration runs (seed 2),
    # the 256-point sweep (seed 1) screened
    print("case: surrogate screening (4096 train / 512 calibration / 256 sweep)", flush=True)
    import time
    from luna_jl_amd import surrogate as sg
    n_tr, n_cal = 4096, 512
    E, P, T = sg.scan_params(n_tr + n_cal, seed=2)
    S_in, S_out, st = sg.solve_spectra(E, P, T)
    assert np.all(st == 0)
    S_ref = float(S_out.max())
    ys = sg.Scaler(S_ref).fit(S_out[:n_tr])
    xs = sg.Scaler(S_ref).fit(S_in[:n_tr])
    X = sg.features(S_in, E, P, xs)
    Y = ys.encode(S_out).astype(np.float32)
    net = sg.FNO(max_batch=64)
    net.set_params(net.init_params(seed=0))
    net.set_data(X[:n_tr], Y[:n_tr], sg.loss_weights(S_out[:n_tr]))
    t0 = time.perf_counter()
    losses = net.train(20000, batch=32, lr=1e-3, seed=0)
    print("train 20000 steps:", time.perf_counter() - t0, "s; loss", losses[:10].mean(), "->", losses[-50:].mean(),
          flush=True)
    err_cal = sg.spectral_error(ys.decode(net.forward(X[n_tr:])), S_out[n_tr:])
    screen = sg.Screen(E[n_tr:], P[n_tr:], err_cal, bound=BOUND)
    Es, Ps, Ts = sg.scan_params(256, seed=1)
    Ss_in, Ss_out, sts = sg.solve_spectra(Es, Ps, Ts)
    pred = ys.decode(net.forward(sg.features(Ss_in, Es, Ps, xs)))
    true_err = sg.spectral_error(pred, Ss_out)
    acc = screen.predicted_error(Es, Ps) <= BOUND
    gain = len(Es) / max(1, int((~acc).sum()))
    print(f"calibration error median {np.median(err_cal):.2e} max {err_cal.max():.2e}; safety {screen.safety}; "
          f"accepted {int(acc.sum())}/256, gain {gain:.2f}x; worst accepted true error "
          f"{true_err[acc].max() if acc.any() else None}", flush=True)
    # the screen bounds the error statistically (calibrated on held-out runs):
    # at least 98 % of the accepted points must be within the bound
    assert np.mean(true_err[acc] <= BOUND) >= 0.98
    assert gain >= 5.0  # the target: surrogate screening cuts full-solver calls by at least 5x
