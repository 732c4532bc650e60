"""Host-side pieces of the learned-correction pipeline (no GPU): the 4x4
restriction of scenarios, NaN-aware block means of extrema fields, runup and
the error metrics."""
import numpy as np
import pytest

from tsuner_amd import correction as K
from tsuner_amd import scenarios as S


def test_coarse_scenario_restricts_the_free_surface():
    f = S.gaussian_source(32, 10.0, seed=4)
    c = K.coarse_scenario(f)
    assert (c.npx, c.npy) == (8, 8)
    assert c.params["dx"] == 40.0 and c.params["nvol"] == 64
    # bed: block means; fully wet blocks: block means of h as well
    np.testing.assert_allclose(c.V[3], f.V[3].reshape(8, 4, 8, 4).mean(axis=(1, 3)), rtol=1e-14)
    wet = (f.V[0] > K.RESTRICT_WET_H).reshape(8, 4, 8, 4).sum(axis=(1, 3))
    full = wet == 16
    assert full.any()
    np.testing.assert_allclose(c.V[0][full], f.V[0].reshape(8, 4, 8, 4).mean(axis=(1, 3))[full], rtol=1e-12)
    blk = f.fixH[8:12, 12:16]
    assert c.fixH[2, 3] == blk.min()


def test_restriction_keeps_a_lake_at_rest_across_the_shoreline():
    """A 4x4 block half sea (1 m deep) and half land (1 m high): a block mean
    of h would give a 0.5 m surface above datum; the restriction keeps the
    surface at datum (dry coarse cell, bed at datum), and wet blocks keep
    their velocity and concentration."""
    V = np.zeros((5, 8, 8))
    V[3] = 1.0                      # depth below datum
    V[3][:, :2] = -1.0              # land in the two left columns
    V[0] = np.maximum(V[3], 0.0)    # still water at datum
    R = K.restrict_state(V)
    assert R[3, 0, 0] == 0.0 and R[0, 0, 0] == 0.0           # straddling block: dry, bed at datum
    assert np.all(R[0, :, 1] == 1.0) and np.all(R[3, :, 1] == 1.0)
    assert np.all(R[0] - R[3] <= 0.0 + 1e-15)                # surface never above datum
    V[1] = 0.3 * V[0]
    V[4] = 0.01 * V[0]
    R = K.restrict_state(V)
    np.testing.assert_allclose(R[1, :, 1], 0.3)
    np.testing.assert_allclose(R[4, :, 1], 0.01)
    # a wave crest of 0.2 m on the wet half raises the straddling block's
    # surface to 0.2 m over a bed at datum
    V[0][:, 2:4] += 0.2
    R = K.restrict_state(V)
    np.testing.assert_allclose(R[0, :, 0], 0.2)


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
    assert np.isnan(m["bed_change_mae"])                 # no bed changes given


def test_bed_change_error_over_moving_cells():
    """Bed change error: mean |coarse - fine| over the cells where either run
    moved the bed by more than 1 mm, beside the fine run's mean |change|
    there."""
    n = 4
    zf, zc = np.zeros((4 * n, 4 * n)), np.zeros((n, n))
    fine = lambda dH: K.RunResult(1.0, np.zeros(1), [zf], [zf], [zf], dH=[dH])
    coarse = lambda dH: K.RunResult(1.0, np.zeros(1), [zc], [zc], [zc], dH=[dH])
    dfine = np.zeros((n, n))
    dfine[1, 2] = 0.02
    dfine[3, 0] = -0.01
    dcoarse = dfine.copy()
    dcoarse[1, 2] = 0.01
    dcoarse[0, 0] = 0.0005                               # below 1 mm: not counted
    m = K.metrics(fine(dfine), coarse(dcoarse), 0)
    assert m["bed_change_mae"] == pytest.approx(0.005)   # (0.01 + 0) / 2 cells
    assert m["bed_change_fine_mean_abs"] == pytest.approx(0.015)


def test_subgrid_runup_reads_the_fine_ground():
    """A coarse cell reached by water floods the fine land cells of its block
    lying more than WET_H below its highest surface; the runup is the highest
    of them, so a coarse run whose surface matches the fine run's gives the
    fine run's runup, not the coarse cells' mean ground."""
    n = 4
    z = np.tile(np.linspace(-3.0, 3.0, 4 * n), (4 * n, 1))        # ground rising to the right
    H0 = -z
    eta = 1.3                                                      # still surface reached 1.3 m
    hmax_f = np.where(z < eta, eta - z, 0.0)
    rf = K.runup(H0, hmax_f)
    assert rf == z[0][z[0] < eta].max()
    Hc = H0.reshape(n, 4, n, 4).mean(axis=(1, 3))
    hc = np.maximum(eta + Hc, 0.0)                                 # coarse depth for the same surface
    ec = np.where(hc > K.WET_H, eta, np.nan)
    assert K.runup_subgrid(H0, ec, hc) == rf
    assert K.runup(Hc, hc) != rf                                   # block-mean ground differs
    # a coarse cell never reached floods nothing
    assert K.runup_subgrid(H0, np.full((n, n), np.nan), np.zeros((n, n))) == 0.0



def test_readout_metrics_move_only_the_front_cells():
    """readout_metrics with a readout that adds 5 cm to every cell it reads
    out: the zone peak, which sits on wetted land, rises by 5 cm; the sea
    cells' peak and every runup entry stay those of the run itself."""
    from tsuner_amd import eta_readout as E
    n = 16
    H0f = np.tile(np.linspace(20.0, -4.0, 4 * n), (4 * n, 1))
    wet_f = H0f > -2.0
    ef = np.where(wet_f, np.where(H0f > 0, 0.5, 0.5 + 0.01 * (-H0f)), np.nan)
    hf = np.where(wet_f, 1.0, 0.0)
    fine = K.RunResult(1.0, np.zeros(1), [ef], [hf], [H0f])
    H0c = H0f.reshape(n, 4, n, 4).mean(axis=(1, 3))
    ec = K._block_mean_nan(ef)
    hc = hf.reshape(n, 4, n, 4).max(axis=(1, 3))
    run = K.RunResult(1.0, np.zeros(1), [ec], [hc], [H0c])
    m_run = [K.metrics(fine, run, 0)]
    model = {"mu": np.zeros(E.NFEAT), "sd": np.ones(E.NFEAT), "W1": np.zeros((E.NFEAT, 64)), "b1": np.zeros(64),
             "W2": np.zeros((64, 64)), "b2": np.zeros(64), "W3": np.zeros((64, 1)), "b3": np.array([0.05])}
    mr, t = K.readout_metrics(fine, run, m_run, model)
    assert len(mr) == 1 and t >= 0
    assert np.isclose(mr[0]["peak_eta_coarse"], m_run[0]["peak_eta_coarse"] + 0.05)
    assert mr[0]["peak_eta_sea_coarse"] == m_run[0]["peak_eta_sea_coarse"]
    for k in K.RUNUP_KEYS:
        assert mr[0][k] == m_run[0][k]
    # the highest coarse cell is wetted land
    zone = H0c < E.NEARSHORE_DEPTH
    top = np.nanargmax(np.where(zone, ec, np.nan))
    assert H0c.reshape(-1)[top] < 0


def test_fallback_policy_calibrates_on_the_first_half():
    """Fallback thresholds: chosen on the first half of the members as the
    one with the most fallbacks that keeps >= 8x there, then applied to the
    second half, where a fallback member costs one fine run and has zero
    error.  With 20 members, a coarse wall of 1 s for all of them and 10 s
    of fine run per member, the first half affords one fallback
    (100 / (0.5 + 10) >= 8, two would give 4.9x)."""
    n = 20
    rng = np.random.default_rng(3)
    resid = rng.permutation(np.arange(n, dtype=np.float64))
    err = rng.uniform(0.01, 0.2, n)
    mc = [{"peak_eta_coarse": 1.0 + 0.1 * i, "peak_eta_err": err[i]} for i in range(n)]
    mb = [{"peak_eta_coarse": 1.0, "peak_eta_err": 2 * err[i]} for i in range(n)]
    corr = K.RunResult(1.0, np.zeros(n), [], [], [], resid=resid)
    base = K.RunResult(0.5, np.zeros(n), [], [], [])

    def summary(ms, use_fine=None):
        e = [0.0 if (use_fine is not None and use_fine[i]) else m["peak_eta_err"] for i, m in enumerate(ms)]
        return {"peak_eta_err_mean_m": float(np.mean(e))}

    out = K._fallback_policies(mb, mc, corr, base, 10.0, 8.0, summary)
    pol = out["policies"]["resid_max"]
    cal, test = resid[:10], resid[10:]
    tau = np.sort(cal)[-2]                       # exactly one calibration member above it
    assert pol["threshold"] == tau
    fb = test > tau
    assert pol["fallback_frac"] == fb.mean() and pol["scenarios"] == 10
    np.testing.assert_allclose(pol["peak_eta_err_mean_m"], np.mean(np.where(fb, 0.0, err[10:])))
    np.testing.assert_allclose(pol["speedup_vs_fine"], 10.0 * 10 / (1.0 * 10 / n + 10.0 * fb.sum()))
    # an indicator that is NaN never sends a member to the fine solver
    corr_nan = K.RunResult(1.0, np.zeros(n), [], [], [], resid=np.full(n, np.nan))
    pol = K._fallback_policies(mb, mc, corr_nan, base, 10.0, 8.0, summary)["policies"]["resid_max"]
    assert pol["fallback_frac"] == 0.0
    np.testing.assert_allclose(pol["peak_eta_err_mean_m"], np.mean(err[10:]))
