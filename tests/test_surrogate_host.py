"""Host logic of the spectral surrogate: encodings, loss weights, the
predicted-error screen and its leave-one-out calibration (CPU only)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def sg(pkg):
    from luna_jl_amd import surrogate
    return surrogate


def test_axis_and_resample(sg):
    ax = sg.spectral_axis()
    assert ax.shape == (sg.N_BINS,) and np.all(np.diff(ax) > 0)
    w = np.linspace(0, 2 * ax[-1], 5000)
    Ew = np.exp(-((w - ax[500]) / (50 * (ax[1] - ax[0]))) ** 2) + 0j
    S = sg.resample(w, Ew, ax)
    assert np.argmax(S) == 500


def test_scaler_round_trip(sg):
    rng = np.random.default_rng(0)
    S = rng.uniform(0, 1, (5, sg.N_BINS)) ** 6
    sc = sg.Scaler(S.max()).fit(S)
    back = sc.decode(sc.encode(S))
    np.testing.assert_allclose(back, S, rtol=1e-9, atol=1e-15 * S.max())
    y = sc.encode(S)
    assert abs(y.mean()) < 1e-12 and abs(y.std() - 1) < 1e-12


def test_loss_weights(sg):
    S = np.zeros((2, 8))
    S[0, 3] = 4.0
    S[1] = 1.0
    w = sg.loss_weights(S, floor=1e-3)
    np.testing.assert_allclose(w.mean(axis=1), 1.0, rtol=1e-6)
    assert np.argmax(w[0]) == 3 and np.allclose(w[1], 1.0)


def test_features(sg):
    S = np.ones((3, sg.N_BINS))
    E = np.array([50e-6, 125e-6, 200e-6])
    P = np.array([0.5, 1.75, 3.0])
    x = sg.features(S, E, P, sg.Scaler(1.0, 0.0, 1.0))
    assert x.shape == (3, sg.N_BINS, 4) and x.dtype == np.float32
    np.testing.assert_allclose(x[:, 0, 1], [-1, 0, 1], atol=1e-6)
    np.testing.assert_allclose(x[:, 0, 2], [-1, 0, 1], atol=1e-6)
    np.testing.assert_allclose(x[0, [0, -1], 3], [-1, 1])


def test_screen_calibration_is_conservative(sg):
    # errors grow with energy: the calibrated screen must reject the region
    # where held-out errors exceed the bound and accept the rest
    rng = np.random.default_rng(1)
    E = rng.uniform(50e-6, 200e-6, 400)
    P = rng.uniform(0.5, 3.0, 400)
    err = 2e-3 + 1.5e-2 * ((E - 50e-6) / 150e-6) ** 4
    scr = sg.Screen(E, P, err, k=8, bound=1e-2)
    # leave-one-out on the calibration set itself: no accepted point above the bound
    loc = err[scr._nearest(scr.pts, exclude_self=True)].max(axis=1)
    acc = scr.safety * loc <= 1e-2
    assert not np.any(err[acc] > 1e-2)
    Et = np.array([60e-6, 195e-6])
    pred = scr.predicted_error(Et, np.array([1.0, 1.0]))
    assert pred[0] <= 1e-2 < pred[1]


def test_scan_params_seeded(sg):
    E1, P1, T1 = sg.scan_params(16, seed=2)
    E2, P2, T2 = sg.scan_params(16, seed=2)
    assert np.array_equal(E1, E2) and np.array_equal(P1, P2)
    assert set(np.unique(T1)) == {0.5e-12, 1e-12}
    assert E1.min() >= 50e-6 and E1.max() <= 200e-6 and P1.min() >= 0.5 and P1.max() <= 3.0
