"""The learned-correction pipeline end to end at a small size (fine 64^2,
coarse 16^2): training pairs line up in time, a zero network leaves the
coarse run unchanged bit for bit, and training on the pairs lowers the loss;
and at full size, what the shipped network gains over the coarse solver
alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_pairs_zero_network_and_training():
    from tsuner_amd import correction as K
    from tsuner_amd import nn
    scale = [0.05, 0.5, 0.5, 0.01]
    ds = K.generate([11, 12], windows=4, scale_out=scale, fine_n=64)
    assert ds.count == 8
    x = ds.x.float().cpu().numpy()
    y = ds.y.cpu().numpy()
    assert x.shape == (8, 16, 16, 8) and y.shape == (8, 16, 16, 4)
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(y))
    assert np.abs(y).max() > 0      # the coarse solver does differ from the restricted fine one

    # a network whose output is identically zero must not change the run
    fine_sc = [K.fine_scenario(s, extrema=1, n=64) for s in (21, 22)]
    coarse_sc = [K.coarse_scenario(f) for f in fine_sc]
    t_end = np.array([30.0, 45.0])
    zero = nn.Net(np.zeros(nn.NPARAM, np.float32), max_batch=2, h=16, w=16)
    a = K.run_coarse(coarse_sc, t_end)
    b = K.run_coarse(coarse_sc, t_end, net=zero, scale_out=scale)
    assert np.all(a.t_end == t_end) and np.all(b.t_end == t_end)
    for i in range(2):
        np.testing.assert_array_equal(np.nan_to_num(a.eta_max[i], nan=-7.0), np.nan_to_num(b.eta_max[i], nan=-7.0))
        np.testing.assert_array_equal(a.hmax[i], b.hmax[i])
    assert np.all(b.resid == 0)

    # a few epochs on the pairs lower the loss over the whole set
    def full_loss(net):
        tot = 0.0
        for b in range(2):
            xa = ds.x.data_ptr() + b * 4 * 16 * 16 * 8 * 2
            ya = ds.y.data_ptr() + b * 4 * 16 * 16 * 4 * 4
            net.forward(xa, 4)
            net.backward(xa, ya, 4)
            tot += net.loss(4)
        return tot

    net = nn.Net(nn.init_params(3), max_batch=4, h=16, w=16)
    before = full_loss(net)
    net, hist = K.train(ds, epochs=30, batch=4, lr=3e-3, net=net, val_frac=0.0)
    assert full_loss(net) < 0.8 * before


def test_networks_meet_the_bars_and_beat_coarse_alone():
    """What the trained networks gain over the method without it, on the
    benchmark's 200 held-out scenarios at full size (fine 1024^2, coarse
    256^2, 3000 fine steps; a sample of 24 is too small for the absolute
    bars: runup errors of single scenarios reach 30 %): the configured
    network (four convolutions, 32 channels) and the
    wider shipped one (six, 64 channels) against the coarse solver alone and
    the fine solver, with the metrics the benchmark reports.  Bars asserted:
    coastal peak eta_max error <= 5 cm and runup error <= 10 % (means over
    the scenarios) at >= 8x the fine solver's throughput, each network
    improving on coarse alone in coastal peak, zone peak, nearshore mean and
    runup.  The zone peak (inundation front included) is reported, not held
    to 5 cm: on 200 scenarios it is about 7.4 cm for both networks.  The
    learned eta_max readout of the configured network's run must lower the
    zone peak error and keep the coastal peak and runup bars; with the
    residual fallback it holds the zone peak to 5 cm at >= 8x (4.7 cm at
    9.4x when written)."""
    from tsuner_amd import correction as K
    from tsuner_amd import eta_readout
    nets = {"conv4_w32": K.load_net(K.WEIGHTS_4X32), "conv6_w64": K.load_net(K.WEIGHTS)}
    net, so, _ = nets["conv4_w32"]
    ev = K.evaluate(net, so, list(range(300001, 300201)), fine_steps=3000,
                    others={"conv6_w64": nets["conv6_w64"][:2] + (eta_readout.load(eta_readout.WEIGHTS_6X64),)},
                    readout=eta_readout.load())
    base = ev["coarse_alone"]
    rows = {"conv4_w32": ev["coarse_nn"], "conv6_w64": ev["others"]["conv6_w64"]["coarse_nn"]}
    print("coarse_alone", base, flush=True)
    for k, corr in rows.items():
        print(k, corr, flush=True)
        assert corr["peak_eta_sea_err_mean_m"] <= 0.05, (k, corr["peak_eta_sea_err_mean_m"])
        assert corr["runup_rel_err_mean"] <= 0.10, (k, corr["runup_rel_err_mean"])
        assert corr["speedup_vs_fine"] >= 8.0, (k, corr["speedup_vs_fine"])
        assert corr["peak_eta_sea_err_mean_m"] < 0.8 * base["peak_eta_sea_err_mean_m"]
        assert corr["peak_eta_err_mean_m"] < base["peak_eta_err_mean_m"]
        assert corr["nearshore_mean_abs_eta_err_m"] < 0.8 * base["nearshore_mean_abs_eta_err_m"]
        assert corr["runup_rel_err_mean"] < base["runup_rel_err_mean"]
    ro = ev["coarse_nn_readout"]
    print("conv4_w32 + eta readout", ro, flush=True)
    for name, pol in ev["readout_fallback_policies"]["policies"].items():
        print("readout fallback", name, {k: pol[k] for k in ("peak_eta_err_mean_m", "fallback_frac",
                                                             "speedup_vs_fine")}, flush=True)
    assert ro["peak_eta_err_mean_m"] < 0.85 * ev["coarse_nn"]["peak_eta_err_mean_m"]
    assert ro["peak_eta_sea_err_mean_m"] <= 0.05 and ro["runup_rel_err_mean"] <= 0.10
    assert ro["speedup_vs_fine"] >= 8.0
    # with the residual fallback (threshold from scenarios 300001-300100,
    # reported on 300101-300200): the zone peak bar at >= 8x
    ro64 = ev["others"]["conv6_w64"]["coarse_nn_readout"]
    print("conv6_w64 + eta readout", ro64, flush=True)
    for name, pol in ev["others"]["conv6_w64"]["readout_fallback_policies"]["policies"].items():
        print("conv6_w64 readout fallback", name, {k: pol[k] for k in ("peak_eta_err_mean_m", "fallback_frac",
                                                                       "speedup_vs_fine")}, flush=True)
    assert ro64["peak_eta_err_mean_m"] < 0.85 * ev["others"]["conv6_w64"]["coarse_nn"]["peak_eta_err_mean_m"]
    assert ro64["peak_eta_sea_err_mean_m"] <= 0.05 and ro64["runup_rel_err_mean"] <= 0.10
    pol = ev["readout_fallback_policies"]["policies"]["resid_max"]
    print("readout + resid_max fallback", pol, flush=True)
    assert pol["peak_eta_err_mean_m"] <= 0.05 and pol["speedup_vs_fine"] >= 8.0
    assert pol["peak_eta_sea_err_mean_m"] <= 0.05 and pol["runup_rel_err_mean"] <= 0.10
