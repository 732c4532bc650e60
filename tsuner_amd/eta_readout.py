"""Learned readout of the coarse run's eta_max near the shore.

The coarse + network run's eta_max stands on 40 m cells whose ground is the
mean of 4 x 4 fine cells.  At the inundation front a partly flooded block's
water rests on that mean ground, while the fine run's block mean of eta_max
covers only its flooded (lower) fine cells: over the 200 held-out
scenarios the coarse value at the zone's highest cell lies 7 cm above the
fine one on average.  The readout is a small per-cell MLP that maps what the
coarse run and the fine bathymetry (an input of every scenario) say about a
cell to the fine run's block mean of eta_max there.  It is applied to the
land cells of the nearshore zone that the coarse run wetted whose eta_max
lies within BAND of the zone's highest value (the front cells that can
hold the zone peak; at most the MAX_CELLS highest of them); every other
cell, the sea cells among them, keeps the run's own value.

Features per cell (coarse grid, 4 x 4 fine cells per block):
  eta_max, h_max, ground z = -H0, min / max / std of the fine ground in the
  block, the fraction of fine cells below eta_max and their mean ground,
  the 3 x 3 neighbourhood's largest and mean eta_max (dry cells at their
  ground) and dry count, eta_max - z, the largest eta_max over the zone's
  sea cells, the local ground slope and the neighbourhood's largest h_max.

Training (`python -m tsuner_amd.eta_readout train`): fine and coarse +
network runs of the correction's training scenarios (seeds 1-96), the
cells above as samples, target fine block mean - coarse eta_max over the
samples the fine run wetted.  The MLP's bias on the zone peak, estimated by
five scenario-grouped folds (median of the held-out folds' signed peak
errors), is subtracted from its output layer.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

from . import correction as K

BAND = 1.0                 # m below the coarse zone peak
MAX_CELLS = 3000           # at most this many cells per member, the highest
NEARSHORE_DEPTH = 10.0     # zone: coarse cells initially shallower than this [m]
WEIGHTS = os.path.join(K.HERE, "eta_readout_4x32.npz")          # for the 4x32 network's runs
WEIGHTS_6X64 = os.path.join(K.HERE, "eta_readout_6x64.npz")     # for the 64-wide network's runs
NFEAT = 15


def features(eta_max: np.ndarray, hmax: np.ndarray, H0c: np.ndarray, H0f: np.ndarray):
    """(flat indices of the cells the readout corrects, their features
    (m, NFEAT)).  eta_max is NaN where the coarse run never wetted a cell;
    H0c and H0f are the initial depths below datum of the coarse and fine
    grids.  The cells: land cells (H0 < 0) of the nearshore zone that the
    coarse run wetted, with eta_max within BAND of the zone's highest value
    (at most the MAX_CELLS highest)."""
    n = eta_max.shape[0]
    if H0f.shape[0] != n * K.FACTOR:
        raise ValueError("fine bathymetry must be %d x the coarse grid" % K.FACTOR)
    zc = -H0c
    wet = np.isfinite(eta_max)
    zone = H0c < NEARSHORE_DEPTH
    cand = zone & wet
    if not cand.any():
        return np.zeros(0, np.int64), np.zeros((0, NFEAT))
    ecf = np.where(wet, eta_max, zc)
    peak = ecf[cand].max()
    idx = np.flatnonzero(cand & (H0c < 0) & (ecf >= peak - BAND))
    if idx.size > MAX_CELLS:
        idx = np.sort(idx[np.argsort(-ecf.reshape(-1)[idx], kind="stable")[:MAX_CELLS]])
    if idx.size == 0:
        return idx, np.zeros((0, NFEAT))
    i, j = np.divmod(idx, n)
    # the selected cells' fine blocks and 3 x 3 neighbourhoods (edges repeat)
    blk = -H0f.reshape(n, K.FACTOR, n, K.FACTOR)[i, :, j, :].reshape(idx.size, -1)
    e = ecf[i, j]
    below = blk < e[:, None]
    nb_below = below.sum(axis=1)
    zb = np.where(nb_below > 0, np.where(below, blk, 0.0).sum(axis=1) / np.maximum(nb_below, 1), zc[i, j])
    sea = zone & (H0c > 0)
    peak_sea = ecf[sea].max() if sea.any() else 0.0

    # neighbour indices of the selected cells (edges repeat)
    rows = [np.clip(i + u, 0, n - 1) for u in (-1, 0, 1)]
    cols = [np.clip(j + v, 0, n - 1) for v in (-1, 0, 1)]

    def nine(a):
        return np.stack([a[r, c] for r in rows for c in cols])

    e9, z9 = nine(ecf), nine(zc)
    slope = np.maximum(np.abs(z9[5] - z9[3]), np.abs(z9[7] - z9[1])) / 2
    X = np.stack([e, hmax[i, j], zc[i, j], blk.min(axis=1), blk.max(axis=1), blk.std(axis=1),
                  nb_below / float(blk.shape[1]), zb, e9.max(axis=0), e - zc[i, j], np.full(idx.size, peak_sea),
                  e9.mean(axis=0), (~nine(wet)).sum(axis=0).astype(np.float64), slope, nine(hmax).max(axis=0)],
                 axis=-1)
    return idx, X


def _silu(x):
    # exp(-x) overflows to inf for x below about -709; x / inf is the -0
    # the function tends to there
    with np.errstate(over="ignore"):
        return x / (1.0 + np.exp(-x))


def predict(model: dict, X: np.ndarray) -> np.ndarray:
    """The MLP's correction (fine block mean - coarse eta_max) per row."""
    h = (X - model["mu"]) / model["sd"]
    h = _silu(h @ model["W1"] + model["b1"])
    h = _silu(h @ model["W2"] + model["b2"])
    return (h @ model["W3"] + model["b3"])[:, 0]


def apply(model: dict, eta_max: np.ndarray, hmax: np.ndarray, H0c: np.ndarray, H0f: np.ndarray) -> np.ndarray:
    """eta_max with the readout's value on the cells it corrects."""
    out = eta_max.copy()
    idx, X = features(eta_max, hmax, H0c, H0f)
    if idx.size:
        out.reshape(-1)[idx] = X[:, 0] + predict(model, X)
    return out


def load(path: str = WEIGHTS) -> dict:
    z = np.load(path)
    return {k: np.asarray(z[k], np.float64) for k in ("mu", "sd", "W1", "b1", "W2", "b2", "W3", "b3")}


def samples(fine: "K.RunResult", coarse: "K.RunResult"):
    """(X, y, group, peak) over the members: features of the corrected
    cells, target fine block mean - coarse eta_max (NaN where the fine run
    never wetted the block), member index; per member the fine run's zone
    peak (highest block mean of eta_max over the nearshore zone, NaN if it
    never wetted the zone)."""
    import warnings
    Xs, ys, gs, peaks = [], [], [], []
    for i in range(len(coarse.eta_max)):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            ef = K._block_mean_nan(fine.eta_max[i])
        zone = coarse.H0[i] < NEARSHORE_DEPTH
        peaks.append(float(np.nanmax(ef[zone])) if np.isfinite(ef[zone]).any() else np.nan)
        idx, X = features(coarse.eta_max[i], coarse.hmax[i], coarse.H0[i], fine.H0[i])
        Xs.append(X)
        ys.append(ef.reshape(-1)[idx] - X[:, 0])
        gs.append(np.full(len(idx), i, np.int32))
    return np.concatenate(Xs), np.concatenate(ys), np.concatenate(gs), np.array(peaks)


def _fit_mlp(X, y, epochs, seed, mu, sd, width=64):
    import torch
    torch.manual_seed(seed)
    net = torch.nn.Sequential(torch.nn.Linear(X.shape[1], width), torch.nn.SiLU(), torch.nn.Linear(width, width),
                              torch.nn.SiLU(), torch.nn.Linear(width, 1)).double()
    opt = torch.optim.Adam(net.parameters(), 2e-3)
    Xt = torch.tensor((X - mu) / sd)
    yt = torch.tensor(y)[:, None]
    g = torch.Generator().manual_seed(seed)
    for _ in range(epochs):
        perm = torch.randperm(len(Xt), generator=g)
        for i in range(0, len(Xt), 1024):
            b = perm[i:i + 1024]
            loss = torch.nn.functional.smooth_l1_loss(net(Xt[b]), yt[b], beta=0.05)
            opt.zero_grad()
            loss.backward()
            opt.step()
    L = [m for m in net if isinstance(m, torch.nn.Linear)]
    return {"mu": mu, "sd": sd,
            "W1": L[0].weight.detach().numpy().T.copy(), "b1": L[0].bias.detach().numpy().copy(),
            "W2": L[1].weight.detach().numpy().T.copy(), "b2": L[1].bias.detach().numpy().copy(),
            "W3": L[2].weight.detach().numpy().T.copy(), "b3": L[2].bias.detach().numpy().copy()}


def _signed_peak_errors(model, X, g, peak):
    """Per member with samples and a fine zone peak: the highest readout
    value over its samples minus that peak."""
    p = X[:, 0] + predict(model, X)
    return np.array([p[g == k].max() - peak[k] for k in np.unique(g) if np.isfinite(peak[k])])


def fit(X, y, g, peak, epochs: int = 30, seed: int = 0, folds: int = 5) -> dict:
    """The readout from samples: MLP on the fine-wetted samples, then its
    zone-peak bias from scenario-grouped folds folded into the output bias
    (peak: per member the fine zone peak, as samples returns it)."""
    w = np.isfinite(y)
    mu = X[w].mean(axis=0)
    sd = X[w].std(axis=0) + 1e-6
    model = _fit_mlp(X[w], y[w], epochs, seed, mu, sd)
    members = np.unique(g)
    signed = []
    for f in range(folds):
        val = np.isin(g, members[f::folds])
        trn = ~val & w
        m = _fit_mlp(X[trn], y[trn], epochs, seed + 1 + f, mu, sd)
        signed.append(_signed_peak_errors(m, X[val], g[val], peak))
    bias = float(np.median(np.concatenate(signed)))
    model["b3"] = model["b3"] - bias
    model["fold_bias"] = np.array(bias)
    return model


def save(path: str, model: dict, info: dict):
    np.savez(path, **{k: v for k, v in model.items()}, info=json.dumps(info))


def run_pairs(seeds, net, scale_out, fine_steps: int = 3000):
    """Fine and coarse + network runs of the scenarios (GPU)."""
    fine_sc = [K.fine_scenario(s, extrema=1) for s in seeds]
    coarse_sc = [K.coarse_scenario(f) for f in fine_sc]
    fine = K.run_fine(fine_sc, fine_steps)
    del fine_sc
    coarse = K.run_coarse(coarse_sc, fine.t_end, net=net, scale_out=scale_out, shore_h=K.SHORE_H)
    return fine, coarse


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="learned eta_max readout of the coarse + network run")
    ap.add_argument("cmd", choices=["samples", "fit", "train"],
                    help="samples: runs (GPU) to a samples file; fit: samples file to weights (CPU); train: both")
    ap.add_argument("--seeds", type=int, default=96, help="training scenarios (seeds 1..N)")
    ap.add_argument("--network", default=K.WEIGHTS_4X32, help="the correction network the readout follows")
    ap.add_argument("--samples", default=None, help="samples file (.npz) written by `samples`, read by `fit`")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    if a.cmd in ("samples", "train"):
        import torch
        torch.cuda.init()
        net, so, _ = K.load_net(a.network)
        fine, coarse = run_pairs(list(range(1, a.seeds + 1)), net, so)
        X, y, g, peak = samples(fine, coarse)
        K.log("[readout] %d samples from %d scenarios (%.1f s)" % (len(X), a.seeds, time.perf_counter() - t0))
        if a.samples:
            np.savez_compressed(a.samples, X=X, y=y, g=g, peak=peak, seeds=a.seeds,
                                network=os.path.basename(a.network))
    network = os.path.basename(a.network)
    if a.cmd == "fit":
        z = np.load(a.samples)
        X, y, g, peak = z["X"], z["y"], z["g"], z["peak"]
        if "network" in z:
            network = str(z["network"])
    if a.cmd in ("fit", "train"):
        model = fit(X, y, g, peak, epochs=a.epochs)
        info = {"train_seeds": "1-%d" % int(g.max() + 1), "samples": int(len(X)),
                "wet_samples": int(np.isfinite(y).sum()), "epochs": a.epochs, "band_m": BAND,
                "fold_bias_m": float(model["fold_bias"]), "network": network}
        save(a.out, model, info)
        K.log("[readout] saved %s: %s" % (a.out, info))


if __name__ == "__main__":
    main()
