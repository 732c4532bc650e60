"""Solver-side hooks of the learned correction: network inputs from member
states, 4x4 free-surface restriction, training targets, applying a correction, per-member
extrema, and stop times with the crossing step clamped.  Checked against
numpy restatements of the documented formulas and against the single-grid
solver."""
import numpy as np
import pytest

import nn_oracle as O
import pyoracle
from tsuner_amd import correction as K
from tsuner_amd import scenarios as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SCALE_IN = [1 / 50.0, 1 / 10.0, 1 / 10.0, 1 / 50.0, 10.0, 10.0, 1.0, 1.0]
SCALE_OUT = [0.05, 0.5, 0.5, 0.01]
WET = 1e-3


def members(n, side, dx, seeds, steps=0, **kw):
    return [S.gaussian_source(side, dx, seed=s, steps=steps, **kw) for s in seeds]


def features_ref(V, dx, dy):
    h, qx, qy, H = V[0], V[1], V[2], V[3]
    z = -H
    gx = np.zeros_like(z)
    gy = np.zeros_like(z)
    gx[:, 1:-1] = (z[:, 2:] - z[:, :-2]) / (2 * dx)
    gx[:, 0] = (z[:, 1] - z[:, 0]) / dx
    gx[:, -1] = (z[:, -1] - z[:, -2]) / dx
    gy[1:-1, :] = (z[2:, :] - z[:-2, :]) / (2 * dy)
    gy[0, :] = (z[1, :] - z[0, :]) / dy
    gy[-1, :] = (z[-1, :] - z[-2, :]) / dy
    wet = h > WET
    f = [h, qx, qy, z, gx, gy, np.where(wet, h - H, 0.0), wet.astype(np.float64)]
    return np.stack([(fi.astype(np.float32) * np.float32(s)) for fi, s in zip(f, SCALE_IN)], axis=-1)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_features_restrict_target_apply(precision):
    from tsuner_amd.ensemble import Ensemble
    fine_sc = members(2, 64, 10.0, [3, 4])
    coarse_sc = members(2, 16, 40.0, [3, 4])
    fine = Ensemble(fine_sc, precision=precision)
    coarse = Ensemble(coarse_sc, precision=precision)
    fine.step(5)
    coarse.restrict_from(fine)
    for i in range(2):
        F = fine.state(i)
        Rf = K.restrict_state(F)
        # blocks across the shoreline exercise the free-surface rule
        wet = (F[0] > K.RESTRICT_WET_H).reshape(16, 4, 16, 4).sum(axis=(1, 3))
        assert ((wet > 0) & (wet < 16)).any()
        np.testing.assert_allclose(coarse.state(i), Rf, rtol=1e-6 if precision == "fp32" else 1e-13, atol=1e-12)
    coarse.step(3)
    # inputs
    # the hooks run on the ensemble's own stream: finish torch's fills first
    x = torch.zeros(2 * 16 * 16 * 8, dtype=torch.bfloat16, device="cuda")
    torch.cuda.synchronize()
    coarse.nn_features(SCALE_IN, WET, x.data_ptr())
    coarse.synchronize()
    got = x.float().cpu().numpy().reshape(2, 16, 16, 8)
    for i in range(2):
        ref = O.bf16(features_ref(coarse.state(i), 40.0, 40.0))
        err = np.abs(got[i] - ref)
        assert np.all(err <= 2.0 ** -7 * np.abs(ref) + 1e-30), np.argwhere(err > 2.0 ** -7 * np.abs(ref))[:3]
    # targets
    tgt = torch.zeros(2 * 16 * 16 * 4, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    coarse.nn_target(fine, SCALE_OUT, tgt.data_ptr())
    coarse.synchronize()
    t = tgt.cpu().numpy().reshape(2, 16, 16, 4)
    for i in range(2):
        F, Cs = fine.state(i), coarse.state(i)
        Rf = K.restrict_state(F)
        for j in range(4):
            np.testing.assert_allclose(t[i, ..., j], ((Rf[j] - Cs[j]) / SCALE_OUT[j]).astype(np.float32),
                                       rtol=1e-5, atol=1e-4)
    # apply to the members whose step count is 3 (both)
    assert list(coarse.steps_done()) == [3, 3]
    rng = np.random.default_rng(0)
    out = rng.normal(0, 1, size=(2, 16, 16, 4)).astype(np.float32)
    before = [coarse.state(i) for i in range(2)]
    tout = torch.from_numpy(out).cuda()
    resid = torch.zeros(2, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    coarse.nn_apply(SCALE_OUT, WET, tout.data_ptr(), 3, resid.data_ptr())
    coarse.synchronize()
    # residual indicator: largest |correction of h - H| over each member's wet cells
    got_r = resid.cpu().numpy()
    for i in range(2):
        wet = before[i][0] > WET
        d = np.abs(out[i, ..., 0] * np.float32(SCALE_OUT[0]) - out[i, ..., 3] * np.float32(SCALE_OUT[3]))
        np.testing.assert_allclose(got_r[i], d[wet].max(), rtol=1e-6)
    for i in range(2):
        B = before[i]
        was_wet = B[0] > WET
        assert was_wet.any() and (~was_wet).any()
        h2 = B[0] + out[i, ..., 0].astype(np.float32) * np.float32(SCALE_OUT[0])
        h2 = np.where(was_wet, np.maximum(h2, 0.0), B[0])
        A = coarse.state(i)
        rt = 1e-6 if precision == "fp32" else 1e-12
        np.testing.assert_allclose(A[0], h2, rtol=rt, atol=1e-9)
        H2 = np.where(was_wet, B[3] + out[i, ..., 3] * np.float32(SCALE_OUT[3]), B[3])
        np.testing.assert_allclose(A[3], H2, rtol=rt, atol=1e-9)
        wet = was_wet & (h2 > 0)
        np.testing.assert_allclose(A[1][wet], (B[1] + out[i, ..., 1] * np.float32(SCALE_OUT[1]))[wet], rtol=rt,
                                   atol=1e-6)
        assert np.all(A[1][was_wet & ~wet] == 0) and np.all(A[2][was_wet & ~wet] == 0)
        np.testing.assert_array_equal(A[:, ~was_wet], B[:, ~was_wet])   # dry cells are left alone
        # the next step's dt is the CFL dt of the corrected state
        # (GPU/main_TSUNER.cu:147-148 applied to it), not the one the last
        # stage computed before the correction
        dtc = min(pyoracle.cell_dt(coarse_sc[i].params, *A[:3, r, c]) for r in range(16) for c in range(16))
        got_dt = coarse.time(i)[1]
        if precision == "fp64":
            assert got_dt == dtc
        else:
            assert got_dt == pytest.approx(dtc, rel=1e-5)
    # shoreline mode: with the inputs given, a dry cell next to a wet one (by
    # the inputs' wet flags) takes a depth correction that leaves it deeper
    # than shore_h; other dry cells stay as they are
    snap = [coarse.state(i) for i in range(2)]
    shore_h = 0.02
    flags = got[..., 7] != 0
    coarse.nn_apply(SCALE_OUT, WET, tout.data_ptr(), 3, 0, x.data_ptr(), shore_h)
    coarse.synchronize()
    for i in range(2):
        B, A = snap[i], coarse.state(i)
        nb = np.zeros_like(flags[i])
        nb[:, 1:] |= flags[i][:, :-1]
        nb[:, :-1] |= flags[i][:, 1:]
        nb[1:, :] |= flags[i][:-1, :]
        nb[:-1, :] |= flags[i][1:, :]
        h2 = B[0] + out[i, ..., 0].astype(np.float32) * np.float32(SCALE_OUT[0])
        dry = ~(B[0] > WET)
        takes = dry & nb & (h2 > shore_h)
        np.testing.assert_allclose(A[0][takes], h2[takes], rtol=1e-6)
        np.testing.assert_array_equal(A[3][takes], B[3][takes])
        np.testing.assert_array_equal(A[:, dry & ~takes], B[:, dry & ~takes])
    # members whose step count differs are untouched
    snap = [coarse.state(i) for i in range(2)]
    dts = [coarse.time(i)[1] for i in range(2)]
    coarse.nn_apply(SCALE_OUT, WET, tout.data_ptr(), 4)
    coarse.synchronize()
    for i in range(2):
        np.testing.assert_array_equal(coarse.state(i), snap[i])
        assert coarse.time(i)[1] == dts[i]


def test_ensemble_extrema_match_single_solver():
    from tsuner_amd.ensemble import Ensemble
    from tsuner_amd.solver import Solver
    scs = [S.gaussian_source(48, 20.0, seed=s, extrema=1) for s in (7, 8)]
    e = Ensemble(scs, precision="fp64")
    e.step(12)
    for i, sc in enumerate(scs):
        g = Solver.from_scenario(sc, precision="fp64")
        g.step(12)
        ref = g.extrema()
        got = e.extrema(i)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(e.state(i), g.state())
        g.close()


def test_clamped_stop_time_lands_exactly():
    from tsuner_amd.ensemble import Ensemble
    from tsuner_amd.solver import Solver
    sc = S.gaussian_source(40, 25.0, seed=9)
    free = Solver.from_scenario(sc, precision="fp64")
    free.step(6)
    t6 = free.t
    free.step(1)
    t7 = free.t
    stop = 0.5 * (t6 + t7)
    g = Solver.from_scenario(sc, precision="fp64")
    g.set_stop_time(stop, clamp=True)
    g.step(6)
    ref6 = Solver.from_scenario(sc, precision="fp64")
    ref6.step(6)
    np.testing.assert_array_equal(g.state(), ref6.state())   # steps before the crossing are unchanged
    g.step(5)
    assert g.t == stop and g.steps == 7
    assert not np.array_equal(g.state(), free.state())
    # ensemble members with per-member clamped stop times agree with the solver
    g6 = Solver.from_scenario(sc, precision="fp64")
    g6.set_stop_time(t6, clamp=True)
    g6.step(10)
    e = Ensemble([sc, sc], precision="fp64")
    e.set_stop_time([stop, t6], clamp=True)
    e.step(10)
    assert e.time(0)[0] == stop and e.time(1)[0] == t6 == g6.t
    np.testing.assert_array_equal(e.state(0), g.state())
    np.testing.assert_array_equal(e.state(1), g6.state())
    for s in (free, g, ref6, g6):
        s.close()
