"""Learned coarse-grid correction for ensembles: the solver on a 4x coarser
grid plus a convolutional network that, every WINDOW coarse steps, adds the
difference between the restricted fine solution and the coarse one.

Pipeline (all arithmetic in libtsuner's kernels; torch allocates the
buffers, copies pairs into place and names the stream):
  1. data: fine members (1024^2, dx 10 m) and their 4x4 restrictions (256^2,
     dx 40 m) advance window by window.  At each window start the coarse
     state is reset to restrict(fine); after WINDOW coarse steps the fine
     members advance to exactly the coarse members' times (clamped stop
     times); the pair is (network inputs of the coarse state,
     (restrict(fine) - coarse) / scale_out).
  2. training: Adam on the library's forward / backward kernels.
  3. deployment: a coarse ensemble steps WINDOW steps, the network's output
     is added to every member that completed the window, repeat.  A
     run-time indicator above a threshold sends a member to the fine solver
     (fallback); evaluate() calibrates the threshold on half of the
     scenarios and reports on the other half.

Scenarios: tsuner_amd.scenarios.gaussian_source (fractal shelf bed, Gaussian
eta source, walls, sediment on).
"""
from __future__ import annotations

import json
import os
import sys
import time
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

from . import nn
from . import scenarios as S

FINE_N, FACTOR, FINE_DX = 1024, 4, 10.0
COARSE_N, COARSE_DX = FINE_N // FACTOR, FINE_DX * FACTOR
WINDOW = 10
WET_H = 1e-3
# a correction may wet a dry shoreline cell only beyond this depth [m]
SHORE_H = 1e-2
# inputs: h, qx, qy, z, dz/dx, dz/dy, eta (wet), wet flag
SCALE_IN = [1 / 50.0, 1 / 5.0, 1 / 5.0, 1 / 50.0, 20.0, 20.0, 2.0, 1.0]
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "correction_weights.npz")              # 8 -> 64 x5 -> 4 (six convolutions)
WEIGHTS_4X32 = os.path.join(HERE, "correction_weights_4x32.npz")    # 8 -> 32 x3 -> 4 (four convolutions)


def log(msg):
    # progress goes to stderr: stdout carries only results (JSON)
    print(msg, file=sys.stderr, flush=True)


RESTRICT_WET_H = 1e-3   # fine cells deeper than this are wet in the restriction [m]


def restrict_state(V: np.ndarray, factor: int = FACTOR) -> np.ndarray:
    """Free-surface restriction of a (5, N, N) state onto (5, N/f, N/f), the
    host twin of the library's restriction (tsu_ensemble_restrict): bed =
    block mean of H; surface = mean of h - H over the block's fine cells with
    h > RESTRICT_WET_H; depth = max(surface + bed, 0), 0 when none is wet;
    qx, qy, hc = block totals scaled by depth / block total of h.  A lake at
    rest restricts to a lake at rest, which block means of h do not give
    where a block straddles the shoreline."""
    n = V.shape[1] // factor
    B = V.reshape(5, n, factor, n, factor)
    h, H = B[0], B[3]
    sh = h.sum(axis=(1, 3))
    Hc = H.sum(axis=(1, 3)) / (factor * factor)
    wet = h > RESTRICT_WET_H
    nw = wet.sum(axis=(1, 3))
    seta = np.where(wet, h - H, 0.0).sum(axis=(1, 3))
    with np.errstate(invalid="ignore", divide="ignore"):
        hc = np.where(nw > 0, np.maximum(seta / np.maximum(nw, 1) + Hc, 0.0), 0.0)
        k = np.where((hc > 0) & (sh > 0), hc / np.where(sh > 0, sh, 1.0), 0.0)
    out = np.empty((5, n, n))
    out[0] = hc
    out[1] = B[1].sum(axis=(1, 3)) * k
    out[2] = B[2].sum(axis=(1, 3)) * k
    out[3] = Hc
    out[4] = B[4].sum(axis=(1, 3)) * k
    return out


def coarse_scenario(fine: S.Scenario) -> S.Scenario:
    """4x4 free-surface restriction of a fine scenario (restrict_state; a
    coarse cell is erodible only if all its fine cells are)."""
    n = fine.npx // FACTOR
    p = dict(fine.params)
    p.update(npx=n, npy=n, nvol=n * n, dx=fine.params["dx"] * FACTOR, dy=fine.params["dy"] * FACTOR)
    V = restrict_state(fine.V)
    fix = fine.fixH.reshape(n, FACTOR, n, FACTOR).min(axis=(1, 3)).astype(np.int32)
    return S.Scenario(fine.name + "_coarse", p, dict(fine.scheme), np.ascontiguousarray(V), fix, steps=fine.steps)


def fine_scenario(seed: int, extrema: int = 0, n: int = FINE_N) -> S.Scenario:
    return S.gaussian_source(n, FINE_DX, seed=seed, extrema=extrema)


class _Stream:
    """Makes torch's current stream the ensemble's stream, so library calls
    made through torch-allocated buffers run in order with the solver."""

    def __init__(self, ens):
        import torch
        self.ctx = torch.cuda.stream(torch.cuda.ExternalStream(ens.stream))

    def __enter__(self):
        self.ctx.__enter__()

    def __exit__(self, *a):
        self.ctx.__exit__(*a)


def advance_fine_to(fine, times: np.ndarray, chunk: int = 16, max_steps: int = 100000):
    """Steps the fine ensemble until every member's time equals its entry of
    `times` (clamped stop times)."""
    fine.set_stop_time(times, clamp=True)
    done = 0
    while done < max_steps:
        fine.step(chunk)
        done += chunk
        if np.all(fine.times() >= times):
            return done
    raise RuntimeError("fine members did not reach the window end")


@dataclass
class Dataset:
    x: object            # torch BF16 [N][n][n][8]
    y: object            # torch FP32 [N][n][n][4]
    scale_out: List[float]
    count: int
    seeds: List[int] = field(default_factory=list)


def generate(seeds: Sequence[int], windows: int, scale_out: Sequence[float], precision: str = "fp32",
             capacity: int = None, seed_perm: int = 0, fine_n: int = FINE_N, net: "nn.Net" = None,
             shore_h: float = SHORE_H) -> Dataset:
    """Training pairs from fine/coarse runs of `seeds`, `windows` windows each.
    Without `net` every window starts from restrict(fine).  With `net` the
    coarse members follow their own corrected trajectory instead (the
    network's correction is applied after each window, as in deployment), so
    the pairs show the network the states it meets when it runs.  Pairs are
    written in a random order (contiguous batches are then random)."""
    import torch
    from .ensemble import Ensemble
    fine_sc = [fine_scenario(s, n=fine_n) for s in seeds]
    coarse_sc = [coarse_scenario(f) for f in fine_sc]
    m = len(seeds)
    total = m * windows if capacity is None else capacity
    npx = fine_n // FACTOR
    X = torch.zeros((total, npx, npx, 8), dtype=torch.bfloat16, device="cuda")
    Y = torch.zeros((total, npx, npx, 4), dtype=torch.float32, device="cuda")
    perm = np.random.default_rng(seed_perm).permutation(total)
    fine = Ensemble(fine_sc, precision=precision)
    coarse = Ensemble(coarse_sc, precision=precision)
    del fine_sc, coarse_sc
    xb = torch.zeros((m, npx, npx, 8), dtype=torch.bfloat16, device="cuda")
    yb = torch.zeros((m, npx, npx, 4), dtype=torch.float32, device="cuda")
    ob = torch.zeros((m, npx, npx, 4), dtype=torch.float32, device="cuda") if net is not None else None
    torch.cuda.synchronize()
    k = 0
    t0 = time.perf_counter()
    for w in range(windows):
        if w > 0 and net is None:
            coarse.restrict_from(fine)
        coarse.step(WINDOW)
        tc = coarse.times()
        advance_fine_to(fine, tc)
        coarse.nn_features(SCALE_IN, WET_H, xb.data_ptr())
        coarse.nn_target(fine, scale_out, yb.data_ptr())
        coarse.synchronize()
        if net is not None:
            for b0 in range(0, m, net.max_batch):
                nb = min(net.max_batch, m - b0)
                net.forward(xb.data_ptr() + b0 * npx * npx * 16, nb, out=ob.data_ptr() + b0 * npx * npx * 16)
            torch.cuda.synchronize()
            coarse.nn_apply(scale_out, WET_H, ob.data_ptr(), WINDOW * (w + 1), 0, xb.data_ptr(), shore_h)
            coarse.synchronize()
        for i in range(m):
            if k < total:
                X[perm[k]].copy_(xb[i])
                Y[perm[k]].copy_(yb[i])
            k += 1
        torch.cuda.synchronize()
        if w % 10 == 0:
            log("[data] window %d/%d t=%.1f s (%.1f s)" % (w + 1, windows, float(tc.mean()), time.perf_counter() - t0))
    fine.close()
    coarse.close()
    return Dataset(X, Y, list(scale_out), min(k, total), list(seeds))


# smallest output scales: 1 mm of depth, 0.01 m^2/s of discharge, 0.1 mm of bed
SCALE_OUT_FLOOR = [1e-3, 1e-2, 1e-2, 1e-4]


def normalise(ds: Dataset, chunk: int = 64, weights: Sequence[float] = (1.0, 1.0, 1.0, 1.0)) -> List[float]:
    """Sets the output scales to the RMS of the raw targets over the whole
    set (not below SCALE_OUT_FLOOR) divided by sqrt(weight), and divides the
    targets by them, in chunks through host memory, so that channel c weighs
    weights[c] in the mean-squared-error loss.  Returns the scales."""
    import torch
    n = ds.count
    ss = np.zeros(4)
    cnt = 0
    for i in range(0, n, chunk):
        y = ds.y[i:min(n, i + chunk)].cpu().numpy().reshape(-1, 4).astype(np.float64)
        ss += (y * y).sum(axis=0)
        cnt += y.shape[0]
    scale = [float(max(np.sqrt(s / max(cnt, 1)), f) / np.sqrt(w)) for s, f, w in zip(ss, SCALE_OUT_FLOOR, weights)]
    inv = (1.0 / np.asarray(scale, np.float64)).astype(np.float32)
    for i in range(0, n, chunk):
        y = ds.y[i:min(n, i + chunk)].cpu().numpy()
        ds.y[i:min(n, i + chunk)].copy_(torch.from_numpy(np.ascontiguousarray(y * inv)))
    torch.cuda.synchronize()
    ds.scale_out = scale
    return scale


def pilot_scales(seeds: Sequence[int], windows: int, fine_n: int = FINE_N) -> List[float]:
    """RMS of the raw corrections (h, qx, qy, H) over a short pilot run, not
    below SCALE_OUT_FLOOR."""
    ds = generate(seeds, windows, [1.0, 1.0, 1.0, 1.0], fine_n=fine_n)
    y = ds.y[:ds.count].cpu().numpy().reshape(-1, 4).astype(np.float64)
    rms = np.sqrt(np.mean(y * y, axis=0))
    return [float(max(r, f)) for r, f in zip(rms, SCALE_OUT_FLOOR)]


def train(ds, epochs: int, batch: int = 8, lr: float = 2e-3, seed: int = 0, net: "nn.Net" = None,
          log_every: int = 1, val_frac: float = 0.1, arch: "nn.Arch" = None):
    """Adam over contiguous batches of the (pre-shuffled) dataset or list of
    datasets, the last `val_frac` of each one's batches held out for
    validation; the batch order is reshuffled every epoch.  Returns the
    network and the per-epoch validation losses (mean squared error of the
    normalised targets; 1.0 is the loss of predicting no correction when the
    scales are the RMS)."""
    sets = ds if isinstance(ds, (list, tuple)) else [ds]
    h = sets[0].x.shape[1]
    xstride = h * h * 8 * 2
    ystride = h * h * 4 * 4
    train_b, val_b = [], []          # (x address, y address) per batch
    for d in sets:
        nb = d.count // batch
        nval = min(int(round(nb * val_frac)), nb - 1) if nb > 1 else 0
        for b in range(nb):
            a = (d.x.data_ptr() + b * batch * xstride, d.y.data_ptr() + b * batch * ystride)
            (val_b if b >= nb - nval else train_b).append(a)
    if net is None:
        arch = arch or nn.DEFAULT
        net = nn.Net(nn.init_params(seed, arch), max_batch=batch, h=h, w=h, arch=arch)
    rng = np.random.default_rng(seed)
    hist = []

    def val_loss():
        if not val_b:
            return float("nan")
        tot = 0.0
        for xa, ya in val_b:
            net.forward(xa, batch)
            net.backward(xa, ya, batch)
            tot += net.loss(batch)
        return tot / len(val_b)

    t0 = time.perf_counter()
    hist.append(val_loss())
    log("[train] before training: validation loss %.4f" % hist[-1])
    for ep in range(epochs):
        lr_ep = lr * 0.5 * (1.0 + np.cos(np.pi * ep / max(1, epochs)))
        for j in rng.permutation(len(train_b)):
            xa, ya = train_b[j]
            net.forward(xa, batch)
            net.backward(xa, ya, batch)
            net.adam(lr_ep)
        hist.append(val_loss())
        if ep % log_every == 0 or ep == epochs - 1:
            log("[train] epoch %d/%d lr %.2e validation loss %.4f (%.1f s)" % (ep + 1, epochs, lr_ep, hist[-1],
                                                                              time.perf_counter() - t0))
    return net, hist


@dataclass
class RunResult:
    wall_s: float
    t_end: np.ndarray                 # per member
    eta_max: List[np.ndarray]         # per member, NaN where never wet
    hmax: List[np.ndarray]
    H0: List[np.ndarray]              # initial depth below datum
    resid: np.ndarray = None          # per member max |correction of eta| (coarse + network)
    steps: np.ndarray = None
    dH: List[np.ndarray] = None       # per member bed change H_end - H0 on the coarse grid


def _bed_change(ens, i, H0: np.ndarray, factor: int) -> np.ndarray:
    """H_end - H0 of member i, averaged over factor x factor blocks (the
    restriction's bed is the block mean)."""
    d = ens.state(i)[3] - H0
    if factor == 1:
        return d
    n = d.shape[0] // factor
    return d.reshape(n, factor, n, factor).mean(axis=(1, 3))


def _extrema(ens, i):
    hmax, _, emax = ens.extrema(i)
    emax = np.where(emax > -1e29, emax, np.nan)
    hmax = np.where(hmax > -1e29, hmax, 0.0)
    return emax, hmax


def run_fine(scs: Sequence[S.Scenario], steps: int, precision: str = "fp32", chunk: int = 50) -> RunResult:
    """Fine ensemble, `steps` steps per member, extrema on; wall time of the
    stepping only (states already resident)."""
    from .ensemble import Ensemble
    ens = Ensemble(scs, precision=precision)
    ens.set_stop_step(steps)
    ens.set_graph_chunk(chunk)
    ens.synchronize()
    t0 = time.perf_counter()
    ens.step(steps)
    wall = time.perf_counter() - t0
    res = RunResult(wall, ens.times(), [], [], [sc.V[3].copy() for sc in scs], steps=ens.steps_done(), dH=[])
    for i in range(len(scs)):
        e, h = _extrema(ens, i)
        res.eta_max.append(e)
        res.hmax.append(h)
        res.dH.append(_bed_change(ens, i, res.H0[i], FACTOR))
    ens.close()
    return res


def run_coarse(scs: Sequence[S.Scenario], t_end: np.ndarray, net: "nn.Net" = None, scale_out=None,
               precision: str = "fp32", max_windows: int = 100000, shore_h: float = None) -> RunResult:
    """Coarse ensemble to the per-member end times (clamped), with the
    network's correction after every WINDOW steps when `net` is given (wet
    cells only, or also dry shoreline cells deeper than shore_h after the
    correction when shore_h is given)."""
    import torch
    from .ensemble import Ensemble
    n = len(scs)
    ens = Ensemble(scs, precision=precision)
    ens.set_stop_time(np.asarray(t_end, np.float64), clamp=True)
    npx = scs[0].npx
    x = out = resid = None
    if net is not None:
        x = torch.zeros((n, npx, npx, 8), dtype=torch.bfloat16, device="cuda")
        out = torch.zeros((n, npx, npx, 4), dtype=torch.float32, device="cuda")
        resid = torch.zeros(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    w = 0
    # Members advance in lockstep (WINDOW steps per window) until their own
    # clamped end time; a member is corrected only if it completed the window.
    with _Stream(ens):
        while w < max_windows:
            w += 1
            ens.step(WINDOW, sync=False)
            if net is not None:
                ens.nn_features(SCALE_IN, WET_H, x.data_ptr())
                for b0 in range(0, n, net.max_batch):
                    nb = min(net.max_batch, n - b0)
                    net.forward(x.data_ptr() + b0 * npx * npx * 16, nb, out=out.data_ptr() + b0 * npx * npx * 16)
                ens.nn_apply(scale_out, WET_H, out.data_ptr(), WINDOW * w, resid.data_ptr(),
                             x.data_ptr() if shore_h is not None else 0, shore_h or 0.0)
            if np.all(ens.times() >= t_end):
                break
    wall = time.perf_counter() - t0
    res = RunResult(wall, ens.times(), [], [], [sc.V[3].copy() for sc in scs], steps=ens.steps_done(), dH=[])
    if resid is not None:
        res.resid = resid.cpu().numpy()
    for i in range(n):
        e, h = _extrema(ens, i)
        res.eta_max.append(e)
        res.hmax.append(h)
        res.dH.append(_bed_change(ens, i, res.H0[i], 1))
    ens.close()
    return res


def _block_mean_nan(a: np.ndarray) -> np.ndarray:
    """Means over FACTOR x FACTOR blocks ignoring NaN (never-wet) cells; NaN
    where a whole block is NaN."""
    n = a.shape[0] // FACTOR
    b = a.reshape(n, FACTOR, n, FACTOR)
    cnt = np.isfinite(b).sum(axis=(1, 3))
    s = np.where(np.isfinite(b), b, 0.0).sum(axis=(1, 3))
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(cnt > 0, s / np.maximum(cnt, 1), np.nan)


def runup(H0: np.ndarray, hmax: np.ndarray) -> float:
    """Largest ground elevation (z = -H > 0) reached by water (h > WET_H)."""
    land = H0 < 0
    wet = hmax > WET_H
    sel = land & wet
    return float((-H0[sel]).max()) if sel.any() else 0.0


def runup_subgrid(H0_fine: np.ndarray, eta_max: np.ndarray, hmax: np.ndarray) -> float:
    """Runup of a coarse run read on the fine ground: within each coarse cell
    the water reached (hmax > WET_H), the fine land cells lying more than
    WET_H below that cell's highest surface (eta_max) count as flooded; the
    runup is the highest of them.  Puts the coarse run on the same elevation
    set as the fine run's runup instead of the coarse cells' mean ground."""
    n = eta_max.shape[0]
    z = -H0_fine.reshape(n, FACTOR, n, FACTOR)
    reached = (hmax > WET_H) & np.isfinite(eta_max)
    surf = np.where(reached, eta_max, -np.inf)[:, None, :, None]
    flooded = (z > 0) & (z < surf - WET_H)
    return float(z[flooded].max()) if flooded.any() else 0.0


def metrics(fine: RunResult, coarse: RunResult, i: int, nearshore_depth: float = 10.0) -> dict:
    """Errors of coarse member i against fine member i: peak eta_max over the
    nearshore zone (coarse cells with initial depth < nearshore_depth,
    including land), mean |eta_max error| there, and runup."""
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        ef = _block_mean_nan(fine.eta_max[i])
    ec = coarse.eta_max[i]
    zone = coarse.H0[i] < nearshore_depth
    both = zone & np.isfinite(ef) & np.isfinite(ec)

    def peak(a, z):
        return float(np.nanmax(np.where(z, a, np.nan))) if np.isfinite(a[z]).any() else float("nan")

    peak_f, peak_c = peak(ef, zone), peak(ec, zone)
    # the same over the sea part of the zone only (initially wet, depth
    # below nearshore_depth): the coastal wave height without the
    # inundation front, whose coarse surface sits on block-mean ground
    sea = zone & (coarse.H0[i] > 0)
    peak_fs, peak_cs = peak(ef, sea), peak(ec, sea)
    # runup: the fine run's highest wetted fine ground against the coarse
    # run's sub-grid reading of the same fine ground; the coarse run's own
    # highest wetted (block-mean) ground is reported beside it
    rf = runup(fine.H0[i], fine.hmax[i])
    rc = runup_subgrid(fine.H0[i], ec, coarse.hmax[i])
    rcg = runup(coarse.H0[i], coarse.hmax[i])

    def rel(a, b):
        return abs(a - b) / b if b > 0 else (0.0 if a == 0 else float("inf"))

    # bed change (H_end - H0, block means of the fine run): mean absolute
    # error over the cells where either run moved the bed by more than 1 mm,
    # and the fine run's mean |change| there
    bed_mae = bed_fine = float("nan")
    if fine.dH is not None and coarse.dH is not None:
        df, dc = fine.dH[i], coarse.dH[i]
        act = (np.abs(df) > 1e-3) | (np.abs(dc) > 1e-3)
        if act.any():
            bed_mae = float(np.mean(np.abs(dc[act] - df[act])))
            bed_fine = float(np.mean(np.abs(df[act])))
    return {"peak_eta_fine": peak_f, "peak_eta_coarse": peak_c, "peak_eta_err": abs(peak_c - peak_f),
            "bed_change_mae": bed_mae, "bed_change_fine_mean_abs": bed_fine,
            "peak_eta_sea_fine": peak_fs, "peak_eta_sea_coarse": peak_cs, "peak_eta_sea_err": abs(peak_cs - peak_fs),
            "mean_abs_eta_err": float(np.mean(np.abs(ec[both] - ef[both]))) if both.any() else float("nan"),
            "runup_fine": rf, "runup_coarse": rc, "runup_rel_err": rel(rc, rf),
            "runup_coarse_ground": rcg, "runup_rel_err_coarse_ground": rel(rcg, rf)}


RUNUP_KEYS = ("runup_fine", "runup_coarse", "runup_rel_err", "runup_coarse_ground", "runup_rel_err_coarse_ground")


def readout_metrics(fine: RunResult, run: RunResult, m_run: List[dict], model: dict):
    """metrics() of every member with the run's eta_max read out by the
    eta_readout `model` for the eta_max metrics (zone peak, coastal peak,
    nearshore mean); the runup entries stay those of m_run (the run's own
    reading).  fine.H0 is the fine scenarios' input bathymetry (copied
    before the fine run), which a deployment has without running the fine
    grid.  Returns (per-member metrics, seconds spent in the readout)."""
    from . import eta_readout
    import dataclasses
    t0 = time.perf_counter()
    ro = dataclasses.replace(run, eta_max=[eta_readout.apply(model, run.eta_max[i], run.hmax[i], run.H0[i],
                                                             fine.H0[i]) for i in range(len(run.eta_max))])
    t_ro = time.perf_counter() - t0
    out = []
    for i in range(len(run.eta_max)):
        m = metrics(fine, ro, i)
        for k in RUNUP_KEYS:
            m[k] = m_run[i][k]
        out.append(m)
    return out, t_ro


def save(path: str, net: "nn.Net", scale_out, info: dict):
    info = dict(info, arch=net.arch.as_dict())
    np.savez(path, params=net.params(), scale_out=np.asarray(scale_out, np.float64),
             scale_in=np.asarray(SCALE_IN, np.float64), info=json.dumps(info))


def load(path: str = WEIGHTS):
    """(params, scale_out, info) of a saved network; info["arch"] holds its
    width and number of middle layers (arch_of)."""
    with np.load(path, allow_pickle=False) as f:
        return f["params"].astype(np.float32), [float(v) for v in f["scale_out"]], json.loads(str(f["info"]))


def arch_of(info: dict) -> "nn.Arch":
    a = info.get("arch", {"width": 32, "mids": 2})
    return nn.Arch(int(a["width"]), int(a["mids"]))


def load_net(path: str = WEIGHTS, max_batch: int = 32, h: int = None, w: int = None):
    """The saved network ready for inference on h x w images (default: the
    coarse grid); returns (net, scale_out, info)."""
    params, scale_out, info = load(path)
    h = h or COARSE_N
    net = nn.Net(params, max_batch=max_batch, h=h, w=w or h, arch=arch_of(info))
    return net, scale_out, info


def evaluate(net, scale_out, seeds: Sequence[int], fine_steps: int = 3000, resid_tau: float = 0.5, fine_n: int = FINE_N,
             max_batch: int = 32, min_speedup: float = 8.0, others: dict = None, readout: dict = None) -> dict:
    """Held-out evaluation: fine members (fine_steps steps) against the coarse
    solver alone and coarse + network, with the residual fallback.  others
    ({name: (net, scale_out)}) are further networks evaluated against the
    same fine and coarse-alone runs (their summaries and fallback policies
    under out["others"][name]; a third tuple element is that network's
    eta_max readout).  readout (an eta_readout model for `net`): the coarse
    + network run's eta_max read out by it, reported as
    out["coarse_nn_readout"] with its own fallback policies."""
    t0 = time.perf_counter()
    fine_sc = [fine_scenario(s, extrema=1, n=fine_n) for s in seeds]
    coarse_sc = [coarse_scenario(f) for f in fine_sc]
    log("[eval] %d scenarios generated (%.1f s)" % (len(seeds), time.perf_counter() - t0))
    fine = run_fine(fine_sc, fine_steps)
    del fine_sc
    log("[eval] fine: %.2f s wall, %.0f sims/hr" % (fine.wall_s, len(seeds) * 3600 / fine.wall_s))
    base = run_coarse(coarse_sc, fine.t_end)
    log("[eval] coarse alone: %.2f s wall" % base.wall_s)
    corr = run_coarse(coarse_sc, fine.t_end, net=net, scale_out=scale_out, shore_h=SHORE_H)
    log("[eval] coarse + network: %.2f s wall" % corr.wall_s)
    wet_only = run_coarse(coarse_sc, fine.t_end, net=net, scale_out=scale_out)
    log("[eval] coarse + network, wet cells only: %.2f s wall" % wet_only.wall_s)
    # the network's bed channel left out (scale 0): the bed evolves by the
    # coarse solver alone, the surface and momenta are still corrected
    fixed_bed = run_coarse(coarse_sc, fine.t_end, net=net, scale_out=list(scale_out[:3]) + [0.0], shore_h=SHORE_H)
    log("[eval] coarse + network, bed uncorrected: %.2f s wall" % fixed_bed.wall_s)
    mb = [metrics(fine, base, i) for i in range(len(seeds))]
    mc = [metrics(fine, corr, i) for i in range(len(seeds))]
    mw = [metrics(fine, wet_only, i) for i in range(len(seeds))]
    mfb = [metrics(fine, fixed_bed, i) for i in range(len(seeds))]
    flagged = corr.resid > resid_tau
    n = len(seeds)
    fine_per = fine.wall_s / n
    wall_fallback = corr.wall_s + fine_per * int(flagged.sum())

    def summary(ms, use_fine=None):
        pe = np.array([0.0 if (use_fine is not None and use_fine[i]) else m["peak_eta_err"] for i, m in enumerate(ms)])
        ps = np.array([0.0 if (use_fine is not None and use_fine[i]) else m["peak_eta_sea_err"]
                       for i, m in enumerate(ms)])
        oks = np.isfinite(ps)
        bm = np.array([0.0 if (use_fine is not None and use_fine[i]) else m["bed_change_mae"]
                       for i, m in enumerate(ms)])
        bf = np.array([m["bed_change_fine_mean_abs"] for m in ms])
        me = np.array([0.0 if (use_fine is not None and use_fine[i]) else m["mean_abs_eta_err"] for i, m in enumerate(ms)])
        ru = np.array([0.0 if (use_fine is not None and use_fine[i]) else m["runup_rel_err"] for i, m in enumerate(ms)])
        rf = np.array([0.0 if (use_fine is not None and use_fine[i]) else m["runup_rel_err_coarse_ground"]
                       for i, m in enumerate(ms)])
        ok = np.isfinite(pe)
        okr = np.isfinite(ru)

        def stat(f, a):
            return float(f(a)) if a.size else None

        return {"peak_eta_err_mean_m": stat(np.mean, pe[ok]),
                "peak_eta_err_p95_m": stat(lambda a: np.percentile(a, 95), pe[ok]),
                "peak_eta_err_max_m": stat(np.max, pe[ok]),
                "nearshore_mean_abs_eta_err_m": stat(np.mean, me[np.isfinite(me)]),
                "runup_rel_err_mean": stat(np.mean, ru[okr]),
                "runup_rel_err_p95": stat(lambda a: np.percentile(a, 95), ru[okr]),
                "runup_rel_err_coarse_ground_mean": stat(np.mean, rf[np.isfinite(rf)]),
                "frac_peak_eta_within_5cm": stat(np.mean, (pe[ok] <= 0.05).astype(float)),
                "peak_eta_sea_err_mean_m": stat(np.mean, ps[oks]),
                "bed_change_mae_mean_m": stat(np.mean, bm[np.isfinite(bm)]),
                "bed_change_fine_mean_abs_m": stat(np.mean, bf[np.isfinite(bf)]),
                "frac_peak_eta_sea_within_5cm": stat(np.mean, (ps[oks] <= 0.05).astype(float)),
                "frac_runup_within_10pct": stat(np.mean, (ru[okr] <= 0.10).astype(float)),
                "scenarios_with_runup": int(sum(m["runup_fine"] > 0 for m in ms)),
                "scenarios_with_peak": int(ok.sum())}

    out = {
        "scenarios": n, "fine_steps": fine_steps,
        "fine": {"wall_s": fine.wall_s, "sims_per_hr": n * 3600 / fine.wall_s},
        "coarse_alone": dict(summary(mb), wall_s=base.wall_s, sims_per_hr=n * 3600 / base.wall_s),
        "coarse_nn": dict(summary(mc), wall_s=corr.wall_s, sims_per_hr=n * 3600 / corr.wall_s,
                          speedup_vs_fine=fine.wall_s / corr.wall_s, shore_h_m=SHORE_H),
        "coarse_nn_wet_cells_only": dict(summary(mw), wall_s=wet_only.wall_s,
                                         sims_per_hr=n * 3600 / wet_only.wall_s,
                                         speedup_vs_fine=fine.wall_s / wet_only.wall_s),
        "coarse_nn_bed_uncorrected": dict(summary(mfb), wall_s=fixed_bed.wall_s,
                                          sims_per_hr=n * 3600 / fixed_bed.wall_s,
                                          speedup_vs_fine=fine.wall_s / fixed_bed.wall_s),
        "coarse_nn_with_fallback": dict(summary(mc, flagged), resid_tau_m=resid_tau,
                                        fallback_frac=float(flagged.mean()), wall_s=wall_fallback,
                                        sims_per_hr=n * 3600 / wall_fallback,
                                        speedup_vs_fine=fine.wall_s / wall_fallback),
        "resid_max_m": {"p50": float(np.median(corr.resid)), "p95": float(np.percentile(corr.resid, 95)),
                        "max": float(corr.resid.max())},
        # the same run post-processed at other thresholds (cost of a fallback:
        # the fine ensemble's wall time per member)
        "fallback_sweep": [dict(summary(mc, corr.resid > tau), resid_tau_m=tau,
                                fallback_frac=float((corr.resid > tau).mean()),
                                speedup_vs_fine=fine.wall_s / (corr.wall_s + fine_per * int((corr.resid > tau).sum())))
                           for tau in (0.05, 0.1, 0.2, 0.3, 0.5)],
    }
    out["fallback_policies"] = _fallback_policies(mb, mc, corr, base, fine_per, min_speedup, summary)

    def read_out(run, m_run, model):
        """readout_metrics of the run, summarised; returns (summary,
        fallback policies), the readout's own time in the coarse wall."""
        import dataclasses
        mr, t_ro = readout_metrics(fine, run, m_run, model)
        wall_ro = run.wall_s + t_ro
        summ = dict(summary(mr), wall_s=wall_ro, readout_s=t_ro, sims_per_hr=n * 3600 / wall_ro,
                    speedup_vs_fine=fine.wall_s / wall_ro)
        pol = _fallback_policies(mb, mr, dataclasses.replace(run, wall_s=wall_ro), base, fine_per, min_speedup,
                                 summary)
        return summ, pol

    if readout is not None:
        out["coarse_nn_readout"], out["readout_fallback_policies"] = read_out(corr, mc, readout)
        log("[eval] readout: zone peak %.4f m (network alone %.4f m)"
            % (out["coarse_nn_readout"]["peak_eta_err_mean_m"], out["coarse_nn"]["peak_eta_err_mean_m"]))
    disagree = np.array([abs(mc[i]["peak_eta_coarse"] - mb[i]["peak_eta_coarse"]) for i in range(n)])
    if others:
        out["others"] = {}
        for name, entry in others.items():
            net2, so2 = entry[0], entry[1]
            c2 = run_coarse(coarse_sc, fine.t_end, net=net2, scale_out=so2, shore_h=SHORE_H)
            log("[eval] coarse + network %s: %.2f s wall" % (name, c2.wall_s))
            m2 = [metrics(fine, c2, i) for i in range(n)]
            out["others"][name] = {
                "coarse_nn": dict(summary(m2), wall_s=c2.wall_s, sims_per_hr=n * 3600 / c2.wall_s,
                                  speedup_vs_fine=fine.wall_s / c2.wall_s, shore_h_m=SHORE_H),
                "fallback_policies": _fallback_policies(mb, m2, c2, base, fine_per, min_speedup, summary),
                "per_scenario": {"peak_eta_err_nn": [m["peak_eta_err"] for m in m2],
                                 "peak_eta_nn": [m["peak_eta_coarse"] for m in m2],
                                 "runup_rel_err_nn": [m["runup_rel_err"] for m in m2]}}
            if len(entry) > 2 and entry[2] is not None:
                o = out["others"][name]
                o["coarse_nn_readout"], o["readout_fallback_policies"] = read_out(c2, m2, entry[2])
                log("[eval] readout %s: zone peak %.4f m (network alone %.4f m)"
                    % (name, o["coarse_nn_readout"]["peak_eta_err_mean_m"], o["coarse_nn"]["peak_eta_err_mean_m"]))
    out["per_scenario"] = {"seed": [int(s) for s in seeds],
                           "peak_eta_fine": [m["peak_eta_fine"] for m in mc],
                           "peak_eta_err_coarse": [m["peak_eta_err"] for m in mb],
                           "peak_eta_err_nn": [m["peak_eta_err"] for m in mc],
                           "peak_eta_coarse": [m["peak_eta_coarse"] for m in mb],
                           "peak_eta_nn": [m["peak_eta_coarse"] for m in mc],
                           "peak_eta_sea_fine": [m["peak_eta_sea_fine"] for m in mc],
                           "peak_eta_sea_nn": [m["peak_eta_sea_coarse"] for m in mc],
                           "runup_fine": [m["runup_fine"] for m in mc],
                           "runup_rel_err_coarse": [m["runup_rel_err"] for m in mb],
                           "runup_rel_err_nn": [m["runup_rel_err"] for m in mc],
                           "resid_max": [float(v) for v in corr.resid],
                           "nn_vs_coarse_peak_disagreement": [float(v) for v in disagree]}
    return out


def _fallback_policies(mb, mc, corr, base, fine_per, min_speedup, summary) -> dict:
    """Fallback policies from run-time indicators, each threshold chosen on
    the first half of the scenarios (the largest fallback share that keeps
    >= min_speedup there) and reported on the second half; a fallback costs
    the fine ensemble's per-member wall time and has zero error."""
    n = len(mc)
    disagree = np.array([abs(mc[i]["peak_eta_coarse"] - mb[i]["peak_eta_coarse"]) for i in range(n)])
    peak_c = np.array([mc[i]["peak_eta_coarse"] for i in range(n)])
    indicators = {"resid_max": (corr.resid, corr.wall_s),
                  "coarse_peak_eta": (peak_c, corr.wall_s),
                  "nn_vs_coarse_peak_disagreement": (disagree, corr.wall_s + base.wall_s)}
    half = n // 2
    cal, test = np.arange(half), np.arange(half, n)

    def cost(ind, tau, idx, base_wall):
        fb = ind[idx] > tau
        return base_wall * len(idx) / n + fine_per * int(fb.sum()), fb

    policies = {}
    for name, (ind, bw) in indicators.items():
        ind = np.nan_to_num(np.asarray(ind, np.float64), nan=-np.inf)
        best = None
        for tau in np.unique(np.concatenate([ind[cal], [np.inf]])):
            wall, fb = cost(ind, tau, cal, bw)
            sp = fine_per * len(cal) / wall
            if sp >= min_speedup and (best is None or fb.mean() > best[1]):
                best = (float(tau), float(fb.mean()))
        tau = best[0] if best else float("inf")
        wall, fb = cost(ind, tau, test, bw)
        sub = [mc[i] for i in test]
        policies[name] = dict(summary(sub, fb), threshold=tau, fallback_frac=float(fb.mean()),
                              speedup_vs_fine=fine_per * len(test) / wall, scenarios=int(len(test)))
    return {"min_speedup_on_calibration_half": min_speedup, "policies": policies}


def time_network(net: "nn.Net", n: int, reps: int = 10) -> dict:
    """hipEvent-timed forward passes of the network over n images (inputs
    zero, so the timing is data-independent); FLOPs counted as the MFMA
    work of the convolutions on the real channel counts."""
    import torch
    h, w = net.h, net.w
    x = torch.zeros((n, h, w, 8), dtype=torch.bfloat16, device="cuda")
    torch.cuda.synchronize()
    net.forward(x, n)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        net.forward(x, n)
    b.record()
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    flop_px = net.arch.flop_per_pixel
    flops = flop_px * n * h * w
    return {"images": n, "ms": ms, "tflops": flops / (ms * 1e-3) / 1e12, "flop_per_pixel": flop_px}


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="learned coarse-grid correction: train / evaluate")
    ap.add_argument("cmd", choices=["train", "evaluate", "all", "network"],
                    help="network: time the saved network's forward pass on 32 coarse images")
    ap.add_argument("--seeds", type=int, default=48, help="training scenarios")
    ap.add_argument("--windows", type=int, default=75, help="windows per training scenario")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--rounds", type=int, default=0, help="on-trajectory data rounds after the first training")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--width", type=int, default=32, choices=[32, 64], help="hidden channels")
    ap.add_argument("--loss-weights", default="1,1,1,1", help="loss weights of the h, qx, qy, H corrections")
    ap.add_argument("--mids", type=int, default=2, help="hidden 3x3 layers between the input and output ones")
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--heldout", type=int, default=200)
    ap.add_argument("--fine-steps", type=int, default=3000)
    ap.add_argument("--resid-tau", type=float, default=0.2)
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    import torch
    torch.cuda.init()
    result = {}
    if a.cmd in ("train", "all"):
        t0 = time.perf_counter()
        ds = generate(list(range(1, a.seeds + 1)), a.windows, [1.0, 1.0, 1.0, 1.0])
        scale_out = normalise(ds, weights=[float(v) for v in a.loss_weights.split(",")])
        log("[train] output scales %s" % scale_out)
        t_data = time.perf_counter() - t0
        log("[train] %d pairs in %.1f s" % (ds.count, t_data))
        t1 = time.perf_counter()
        net, hist = train(ds, a.epochs, batch=a.batch, lr=a.lr, arch=nn.Arch(a.width, a.mids))
        sets, pairs, hists = [ds], ds.count, [hist]
        # on-trajectory rounds: new scenarios run with the current network
        # applied, so training also sees the states the network itself leads to
        for rnd in range(a.rounds):
            t2 = time.perf_counter()
            seeds = list(range(1000 * (rnd + 1) + 1, 1000 * (rnd + 1) + 1 + a.seeds))
            ds2 = generate(seeds, a.windows, scale_out, net=net, seed_perm=rnd + 1)
            log("[train] round %d: %d on-trajectory pairs in %.1f s" % (rnd + 1, ds2.count, time.perf_counter() - t2))
            sets.append(ds2)
            pairs += ds2.count
            net, hist = train(sets, max(1, a.epochs // 2), batch=a.batch, lr=a.lr / 2, net=net)
            hists.append(hist)
        t_train = time.perf_counter() - t1
        info = {"pairs": pairs, "train_seeds": ds.seeds, "windows": a.windows, "epochs": a.epochs,
                "loss_weights": a.loss_weights,
                "rounds": a.rounds, "batch": a.batch, "lr": a.lr, "loss": hists, "data_s": t_data,
                "train_s": t_train}
        save(a.weights, net, scale_out, info)
        result["train"] = {k: v for k, v in info.items() if k != "train_seeds"}
        del ds, sets
    if a.cmd == "network":
        net, _, _ = load_net(a.weights)
        result["network_forward"] = time_network(net, 32, reps=20)
    if a.cmd in ("evaluate", "all"):
        net, scale_out, _ = load_net(a.weights)
        seeds = list(range(100001, 100001 + a.heldout))
        result["network_forward"] = time_network(net, 32)
        log("[eval] network forward: %s" % result["network_forward"])
        result["evaluate"] = evaluate(net, scale_out, seeds, fine_steps=a.fine_steps, resid_tau=a.resid_tau)
    js = json.dumps(result, indent=1)
    print(js, flush=True)
    if a.json:
        with open(a.json, "w") as f:
            f.write(js)


if __name__ == "__main__":
    main()
