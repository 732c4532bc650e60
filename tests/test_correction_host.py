"""Host-side pieces of the learned-correction pipeline (no GPU): the 4x4
restriction of scenarios, NaN-aware block means of extrema fields, runup and
the error metrics."""
import numpy as np

from tsuner_amd import correction as K
from tsuner_amd import scenarios as S


def test_coarse_scenario_is_block_mean():
    f = S.gaussian_source(32, 10.0, seed=4)
    c = K.coarse_scenario(f)
    assert (c.npx, c.npy) == (8, 8)
    assert c.params["dx"] == 40.0 and c.params["nvol"] == 64
    np.testing.assert_allclose(c.V[0, 2, 3], f.V[0, 8:12, 12:16].mean(), rtol=1e-15)
    np.testing.assert_allclose(c.V.sum(axis=(1, 2)) * 16, f.V.sum(axis=(1, 2)), rtol=1e-12)
    blk = f.fixH[8:12, 12:16]
    assert c.fixH[2, 3] == blk.min()


def test_block_mean_ignores_never_wet_cells():
    a = np.full((8, 8), np.nan)
    a[0, 0] = 1.0
    a[0, 1] = 3.0
    a[4:8, 4:8] = 2.0
    m = K._block_mean_nan(a)
    assert m[0, 0] == 2.0 and m[1, 1] == 2.0
    assert np.isnan(m[0, 1]) and np.isnan(m[1, 0])


def test_runup_is_highest_wetted_ground():
    H0 = np.array([[5.0, -1.0, -2.0, -3.0]])
    hmax = np.array([[6.0, 0.5, 0.002, 0.0]])
    assert K.runup(H0, hmax) == 2.0           # z = 2 m wetted by 2 mm, z = 3 m dry
    assert K.runup(H0, np.zeros_like(hmax)) == 0.0


def test_metrics_zero_for_identical_runs():
    n = 8
    H0f = np.tile(np.linspace(20.0, -4.0, 4 * n), (4 * n, 1))
    ef = np.where(H0f > -2.0, 0.5, np.nan)
    hf = np.where(H0f > -2.0, 1.0, 0.0)
    fine = K.RunResult(1.0, np.zeros(1), [ef], [hf], [H0f])
    H0c = H0f.reshape(n, 4, n, 4).mean(axis=(1, 3))
    ec = K._block_mean_nan(ef)
    hc = hf.reshape(n, 4, n, 4).max(axis=(1, 3))
    coarse = K.RunResult(1.0, np.zeros(1), [ec], [hc], [H0c])
    m = K.metrics(fine, coarse, 0)
    assert m["peak_eta_err"] == 0.0 and m["mean_abs_eta_err"] == 0.0
    assert m["runup_fine"] > 0
