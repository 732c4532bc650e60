"""The learned eta_max readout (tsuner_amd/eta_readout.py) on the CPU: its
features against direct evaluations of their definitions, the cells it
corrects, its numpy MLP against torch's layers, the fit's recovery of a
known offset, and the shipped weights."""
import os

import numpy as np
import pytest

from tsuner_amd import eta_readout as E


def _case(n=16, seed=0):
    """A coarse n x n beach (4 x 4 fine cells per block) sloping from 20 m
    deep to 5 m of land, with a partly wetted land strip."""
    rng = np.random.default_rng(seed)
    H0f = np.linspace(20.0, -5.0, 4 * n)[None, :] * np.ones((4 * n, 1)) + rng.normal(0, 0.3, (4 * n, 4 * n))
    H0c = H0f.reshape(n, 4, n, 4).mean(axis=(1, 3))
    eta = 1.0 + 0.1 * rng.normal(size=(n, n))
    land = H0c < 0
    eta[land] = np.where(rng.random(land.sum()) < 0.5, np.nan, -H0c[land] + 0.2)
    hmax = np.where(np.isfinite(eta), eta + H0c, 0.0)
    return eta, hmax, H0c, H0f


def test_features_follow_their_definitions():
    eta, hmax, H0c, H0f = _case()
    n = eta.shape[0]
    idx, X = E.features(eta, hmax, H0c, H0f)
    assert X.shape == (idx.size, E.NFEAT) and idx.size > 0
    zone = H0c < E.NEARSHORE_DEPTH
    wet = np.isfinite(eta)
    ecf = np.where(wet, eta, -H0c)
    peak = ecf[zone & wet].max()
    expect = np.flatnonzero(zone & wet & (H0c < 0) & (ecf >= peak - E.BAND))
    np.testing.assert_array_equal(idx, expect)
    for r, k in enumerate(idx):
        i, j = divmod(int(k), n)
        blk = -H0f[4 * i:4 * i + 4, 4 * j:4 * j + 4]
        below = blk < eta[i, j]
        nb = ecf[max(i - 1, 0):i + 2, max(j - 1, 0):j + 2]
        # the 3 x 3 neighbourhood repeats edge cells: its max is the clipped window's max
        assert X[r, 0] == eta[i, j] and X[r, 1] == hmax[i, j] and X[r, 2] == -H0c[i, j]
        assert np.isclose(X[r, 3], blk.min()) and np.isclose(X[r, 4], blk.max()) and np.isclose(X[r, 5], blk.std())
        assert np.isclose(X[r, 6], below.mean())
        assert np.isclose(X[r, 7], blk[below].mean() if below.any() else -H0c[i, j])
        assert np.isclose(X[r, 8], nb.max())
        assert np.isclose(X[r, 9], eta[i, j] + H0c[i, j])
        sea = zone & (H0c > 0)
        assert np.isclose(X[r, 10], ecf[sea].max())


def test_sea_cells_are_left_alone():
    """Sea cells within the band of the zone peak are not read out: only the
    land (front) cells are."""
    eta, hmax, H0c, H0f = _case(seed=6)
    land = H0c < 0
    eta[~land] = np.nanmax(eta[land])          # the sea as high as the highest front cell
    idx, _ = E.features(eta, hmax, H0c, H0f)
    assert idx.size > 0 and np.all(H0c.reshape(-1)[idx] < 0)


def test_apply_changes_only_the_band_cells():
    eta, hmax, H0c, H0f = _case(seed=1)
    m = {"mu": np.zeros(E.NFEAT), "sd": np.ones(E.NFEAT), "W1": np.zeros((E.NFEAT, 64)), "b1": np.zeros(64),
         "W2": np.zeros((64, 64)), "b2": np.zeros(64), "W3": np.zeros((64, 1)), "b3": np.array([0.05])}
    out = E.apply(m, eta, hmax, H0c, H0f)
    idx, _ = E.features(eta, hmax, H0c, H0f)
    changed = np.flatnonzero(np.isfinite(eta).reshape(-1) & (out.reshape(-1) != eta.reshape(-1)))
    np.testing.assert_array_equal(changed, idx)
    np.testing.assert_allclose(out.reshape(-1)[idx] - eta.reshape(-1)[idx], 0.05, atol=1e-12)
    np.testing.assert_array_equal(np.isnan(out), np.isnan(eta))
    # nothing wet in the zone: nothing to correct
    dry = np.full_like(eta, np.nan)
    np.testing.assert_array_equal(np.isnan(E.apply(m, dry, hmax * 0, H0c, H0f)), True)


def test_cell_cap_keeps_the_highest():
    eta, hmax, H0c, H0f = _case(n=64, seed=2)
    old = E.MAX_CELLS
    try:
        E.MAX_CELLS = 7
        idx, X = E.features(eta, hmax, H0c, H0f)
    finally:
        E.MAX_CELLS = old
    full, Xf = E.features(eta, hmax, H0c, H0f)
    assert idx.size == 7 and set(idx) <= set(full)
    assert X[:, 0].min() >= np.sort(Xf[:, 0])[-7]


def test_predict_equals_torch_layers():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    X = rng.normal(size=(50, E.NFEAT))
    torch.manual_seed(0)
    net = torch.nn.Sequential(torch.nn.Linear(E.NFEAT, 64), torch.nn.SiLU(), torch.nn.Linear(64, 64),
                              torch.nn.SiLU(), torch.nn.Linear(64, 1)).double()
    L = [m for m in net if isinstance(m, torch.nn.Linear)]
    mu, sd = rng.normal(size=E.NFEAT), 1 + rng.random(E.NFEAT)
    m = {"mu": mu, "sd": sd}
    for k, layer in enumerate(L, 1):
        m["W%d" % k] = layer.weight.detach().numpy().T
        m["b%d" % k] = layer.bias.detach().numpy()
    ref = net(torch.tensor((X - mu) / sd)).detach().numpy()[:, 0]
    np.testing.assert_allclose(E.predict(m, X), ref, rtol=1e-12, atol=1e-12)


def test_fit_recovers_a_constant_offset():
    """Samples whose fine block mean lies a fixed 4 cm below the coarse
    value in every cell: the fitted readout's zone peak lands on the fine
    one (the fold bias absorbs what the MLP leaves)."""
    pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    Xs, ys, gs, peaks = [], [], [], []
    for k in range(10):
        X = rng.normal(size=(120, E.NFEAT))
        X[:, 0] = 1.0 + 0.3 * rng.random(120)
        Xs.append(X)
        ys.append(np.full(120, -0.04))
        gs.append(np.full(120, k))
        peaks.append(X[:, 0].max() - 0.04)
    X, y, g = np.concatenate(Xs), np.concatenate(ys), np.concatenate(gs)
    model = E.fit(X, y, g, np.array(peaks), epochs=20)
    err = np.abs([(X[g == k, 0] + E.predict(model, X[g == k])).max() - peaks[k] for k in range(10)])
    # 4 cm without the readout; the max over 120 cells picks up the MLP's
    # per-cell scatter on these random features
    assert err.mean() < 0.01 and err.max() < 0.025, err


@pytest.mark.parametrize("path", [E.WEIGHTS, E.WEIGHTS_6X64])
def test_shipped_weights(path):
    if not os.path.exists(path):
        pytest.skip("no shipped readout weights at %s" % path)
    m = E.load(path)
    assert m["W1"].shape == (E.NFEAT, 64) and m["W2"].shape == (64, 64) and m["W3"].shape == (64, 1)
    assert np.all(m["sd"] > 0)
    eta, hmax, H0c, H0f = _case(seed=5)
    out = E.apply(m, eta, hmax, H0c, H0f)
    assert np.all(np.isfinite(out[np.isfinite(eta)]))


def test_silu_saturates_without_warnings():
    with np.errstate(all="raise"):
        y = E._silu(np.array([-1000.0, -50.0, 0.0, 50.0]))
    assert y[0] == 0 and abs(y[1]) < 1e-18 and y[2] == 0 and y[3] == 50.0


def test_samples_targets_and_peaks():
    """samples() on synthetic runs: one row per corrected cell with target
    fine block mean - coarse eta_max (NaN where the fine run never wetted the
    block), the member index, and per member the fine run's zone peak."""
    from tsuner_amd import correction as K
    fines, coarses = [], []
    for seed in (7, 8):
        eta, hmax, H0c, H0f = _case(seed=seed)
        n = eta.shape[0]
        rng = np.random.default_rng(seed)
        ef = np.repeat(np.repeat(np.where(np.isfinite(eta), eta - 0.03, np.nan), 4, 0), 4, 1)
        ef = ef + rng.normal(0, 0.001, ef.shape)
        ef[:4, :4] = np.nan                       # one block the fine run never wetted
        fines.append((ef, H0f))
        coarses.append((eta, hmax, H0c))
    fine = K.RunResult(1.0, np.zeros(2), [f[0] for f in fines], [np.zeros_like(f[0]) for f in fines],
                       [f[1] for f in fines])
    coarse = K.RunResult(1.0, np.zeros(2), [c[0] for c in coarses], [c[1] for c in coarses],
                         [c[2] for c in coarses])
    X, y, g, peak = E.samples(fine, coarse)
    assert X.shape[1] == E.NFEAT and len(X) == len(y) == len(g) and set(np.unique(g)) <= {0, 1}
    for k in (0, 1):
        idx, Xk = E.features(*coarses[k], fines[k][1])
        np.testing.assert_array_equal(X[g == k], Xk)
        efb = K._block_mean_nan(fines[k][0])
        np.testing.assert_array_equal(y[g == k], efb.reshape(-1)[idx] - Xk[:, 0])
        zone = coarses[k][2] < E.NEARSHORE_DEPTH
        assert peak[k] == np.nanmax(efb[zone])
