#!/usr/bin/env python3
"""Solver benchmark: cell-updates/s of the fused RK2 step on one MI355X.

Workload (the roofline domain): 8192 x 8192 cells, dx = 5 m, seeded fractal
shelf bed, sediment on (Ashida-Michiue bedload and erosion/deposition),
N-wave inflow at x = 0, second-order MUSCL (van Leer) + HLL + two-stage RK.
A cell-update is one cell advanced one full time step (both stages).

Prints one JSON line.  value = FP64 cell-updates/s with the state resident in
HBM; the FP32-plane variant is reported beside it.  roofline.achieved uses
algorithmic bytes: 200 B per cell-update in FP64 (stage 1 reads and writes the
5 planes, stage 2 reads both states and writes U^{n+1}), 100 B in FP32, i.e.
100 B (FP64) per cell per stage-kernel launch, divided by the stage kernel's
mean launch duration from HIP events recorded on the solver's stream during
the timed steps.  cpu_baseline times the CPU oracle (serial FP64
restatement of the reference arithmetic, 1 thread) on a window of the same
initial state.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per stage-kernel launch on the 8192^2 workload, from rocprofv3
# FETCH_SIZE and WRITE_SIZE runs scaled by what the same counters read for a
# plain copy kernel of known size; the summary with the commands, the build
# (PROFILE_BUILD, a commit of this repository) and the scale factors is
# TRAFFIC_PROFILE.  None at other sizes.  The stage kernels' work depends on
# the state (waves at rest skip quotients and Riemann solves), their HBM
# traffic does not.
TRAFFIC_8192 = {"fp64": 7.203e9, "fp32": 5.810e9}
TRAFFIC_PROFILE = "profiles/stage_kernel_8192_counters.md"
PROFILE_BUILD = "593c74d"
METRIC = "cell-updates/s at 8192² FP64 vs FP32; ensemble sims/hr coarse+NN vs fine"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class _StdoutToStderr:
    """Sends file descriptor 1 to stderr for the duration (the library's
    TSUNER() prints the reference's progress lines on stdout, and this
    program's stdout carries exactly one JSON line)."""

    def __enter__(self):
        import ctypes
        self._libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self._saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        self._libc.fflush(None)
        os.dup2(self._saved, 1)
        os.close(self._saved)
        return False


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def stage_kernel_name(sc, precision):
    """The stage kernel the library launches for this workload: with
    sediment, the variant that evaluates bedload among the HLL arithmetic
    when the run uses the Ashida-Michiue law (the benchmark's constants keep
    its FP64 square roots in the unscaled range)."""
    from tsuner_amd import scenarios as S
    sed = bool(sc.scheme["bedload"] or sc.scheme["suspension"])
    fast = sed and bool(sc.scheme["bedload"]) and sc.scheme["bedload_law"] == S.BEDLOAD_AM
    return "stage_kernel<%s, %s, %s>" % ("double, double" if precision == "fp64" else "float, float",
                                         "true" if sed else "false", "true" if fast else "false")


def run_variant(sc, precision, steps, warmup, tile_rows):
    from tsuner_amd.solver import Solver
    # algorithmic bytes per cell-update of RK2: 5 evolving planes with
    # sediment (2 reads + 1 write in stage 2, 1 + 1 in stage 1), 3 evolving
    # planes plus the read-only bed without it
    sed = bool(sc.scheme["bedload"] or sc.scheme["suspension"])
    bytes_per_update = (200.0 if sed else 136.0) * (1.0 if precision == "fp64" else 0.5)
    log("[bench] %s: upload %dx%d" % (precision, sc.npx, sc.npy))
    tu = time.perf_counter()
    g = Solver.from_scenario(sc, precision=precision, tile_rows=tile_rows)
    g.synchronize()
    upload_s = time.perf_counter() - tu
    log("[bench] %s: warmup %d steps (tile rows %d)" % (precision, warmup, g.tile_rows))
    g.step(warmup)
    g.set_timing(True)
    g.timing(reset=True)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(steps, sync=False)
    g.synchronize()
    t1 = time.perf_counter()
    ms, launches = g.timing(reset=True)
    g.set_timing(False)
    t, dt, n = g.time()
    cells = sc.npx * sc.npy
    wall = t1 - t0
    nlaunch = launches[0] + launches[1]
    mean_ms = (ms[0] + ms[1]) / max(nlaunch, 1)
    per_launch_bytes = bytes_per_update / 2.0 * cells
    achieved = per_launch_bytes / (mean_ms * 1e-3) / 1e9
    res = {
        "value": cells * steps / wall,
        "bytes_per_cell_update": bytes_per_update,
        "ms_per_step": wall * 1e3 / steps,
        "sim_time": t, "dt": dt, "steps_done": n,
        "stage_ms_mean": [ms[0] / max(launches[0], 1), ms[1] / max(launches[1], 1)],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_8192[precision] if sc.npx == 8192 and sc.npy == 8192 else None,
                     "traffic_source": ("%s (build %s)" % (TRAFFIC_PROFILE, PROFILE_BUILD)
                                        if sc.npx == 8192 and sc.npy == 8192 else None),
                     "kernel": stage_kernel_name(sc, precision),
                     "algorithmic_bytes_per_launch": per_launch_bytes},
        "mass": g.mass(),
    }
    td = time.perf_counter()
    g.state()
    download_s = time.perf_counter() - td
    # host buffers in and out around the timed steps (create + upload + first
    # dt, then one download): the rate a caller handing over host arrays sees
    res["pcie_inclusive"] = {"upload_s": upload_s, "download_s": download_s,
                             "cell_updates_per_s": cells * steps / (upload_s + wall + download_s)}
    log("[bench] %s: %.3f G cell-updates/s, %.3f ms/step, stage kernels %.3f/%.3f ms, %.0f GB/s"
        % (precision, res["value"] / 1e9, res["ms_per_step"], res["stage_ms_mean"][0], res["stage_ms_mean"][1],
           achieved))
    g.close()
    return res


def run_small_grids(warmup):
    """The two smaller single-grid workloads at their configured step counts:
    512^2 plain shallow water (dx 20 m, Gaussian hump on a plane beach, no
    sediment, walls, 2000 steps) and 2048^2 sediment shelf (dx 10 m, fractal
    bed, non-erodible patches, 3 m / 600 s N-wave inflow, 5000 steps), FP64
    and FP32 each."""
    from tsuner_amd import scenarios as S
    out = {}
    for sc in (S.plain_swe(n=512, dx=20.0, steps=2000), S.sediment_shelf(n=2048, dx=10.0, steps=5000)):
        row = {"cells": sc.npx * sc.npy, "steps": sc.steps}
        for prec in ("fp64", "fp32"):
            r = run_variant(sc, prec, sc.steps, warmup, 0)
            row[prec] = {"value": r["value"], "ms_per_step": r["ms_per_step"],
                         "bytes_per_cell_update": r["bytes_per_cell_update"],
                         "roofline_achieved_gbs": r["roofline"]["achieved"],
                         "roofline_frac": r["roofline"]["frac"], "kernel": r["roofline"]["kernel"]}
        out[sc.name] = row
    return out


def precision_error(sc, steps):
    """FP32 (compensated h/H/hc) against FP64 at the same simulation time:
    the FP64 run takes `steps` steps to time T, the FP32 run steps to T with
    its crossing step clamped onto T.  Per field (h, qx, qy, H, hc): max and
    RMS of |FP32 - FP64| over all cells, and both relative to the field's
    largest FP64 magnitude (momenta: the larger of max |qx|, max |qy|);
    water and sediment totals of the two runs."""
    from tsuner_amd.solver import Solver
    g = Solver.from_scenario(sc, precision="fp64")
    g.step(steps)
    T = g.t
    A = g.state()
    mA = g.mass()
    g.close()
    g = Solver.from_scenario(sc, precision="fp32")
    g.set_stop_time(T, clamp=True)
    while g.t < T:
        g.step(500)
    steps32 = g.steps
    B = g.state()
    mB = g.mass()
    g.close()
    if not np.all(np.isfinite(B)):
        raise RuntimeError("the FP32 run produced non-finite values")
    qscale = max(np.abs(A[1]).max(), np.abs(A[2]).max())
    fields = {}
    for f, name in enumerate(("h", "qx", "qy", "H", "hc")):
        d = np.abs(B[f] - A[f])
        scale = qscale if f in (1, 2) else max(np.abs(A[f]).max(), 1e-30)
        mx, rms = float(d.max()), float(np.sqrt(np.mean(d * d)))
        fields[name] = {"max_abs": mx, "rms_abs": rms, "scale": float(scale), "max_rel": mx / scale,
                        "rms_rel": rms / scale}
    return {"time_s": T, "steps_fp64": steps, "steps_fp32": int(steps32), "fields": fields,
            "water_total_rel_diff": (mB[0] - mA[0]) / mA[0], "sediment_total_rel_diff": (mB[1] - mA[1]) / mA[1],
            "units": "h, H, hc in m; qx, qy in m^2/s; relative to the scale given"}


def run_fp32_vs_fp64():
    """The 2048^2 sediment shelf (dx 10 m, fractal bed, non-erodible patches,
    3 m / 600 s N-wave inflow) for its configured 5000 steps in both
    precisions, compared field by field at the same simulation time."""
    from tsuner_amd import scenarios as S
    sc = S.sediment_shelf(n=2048, dx=10.0, steps=5000)
    log("[bench] fp32 vs fp64: %s, %d steps" % (sc.name, sc.steps))
    r = precision_error(sc, sc.steps)
    r["workload"] = ("sediment shelf 2048x2048, dx=10 m, N-wave inflow (3 m, 600 s), sediment on, %d FP64 steps"
                     % sc.steps)
    log("[bench] fp32 vs fp64 at t=%.1f s: " % r["time_s"] + ", ".join(
        "%s max %.2e rms %.2e" % (k, v["max_rel"], v["rms_rel"]) for k, v in r["fields"].items()))
    return r


ENSEMBLE_SIDES = (256, 384, 512, 768, 1024)


def run_ensemble(members, steps, seed=2024, chunk=50):
    """Fine-grid ensemble: `members` Gaussian-source scenarios with sides drawn
    uniformly from ENSEMBLE_SIDES (dx = 10 m), FP32 state, tile heights
    autotuned per grid-size class, every member advanced `steps` steps by one
    batched stage launch per RK stage replayed from HIP graphs.  sims/hr =
    members * 3600 / wall seconds of the whole ensemble run."""
    from tsuner_amd import autotune
    from tsuner_amd import scenarios as S
    from tsuner_amd.ensemble import Ensemble
    rng = np.random.default_rng(seed)
    sides = rng.choice(ENSEMBLE_SIDES, size=members)
    log("[bench] ensemble: generating %d members" % members)
    t0 = time.perf_counter()
    scs = [S.gaussian_source(int(n), 10.0, seed=seed + k, steps=steps) for k, n in enumerate(sides)]
    cells = int(sum(int(n) * int(n) for n in sides))
    log("[bench] ensemble: generated in %.1f s, %.1f M cells" % (time.perf_counter() - t0, cells / 1e6))
    ens = Ensemble(scs, precision="fp32")
    del scs
    rows = autotune.tune_ensemble(ens, "fp32", use_cache=False, log=log)
    ens.set_stop_step(steps)
    ens.set_graph_chunk(chunk)
    ens.synchronize()
    t0 = time.perf_counter()
    ens.step(steps)
    wall = time.perf_counter() - t0
    steps_done = [ens.time(i)[2] for i in range(members)]
    ens.close()
    res = {"sims_per_hr": members * 3600.0 / wall, "wall_s": wall, "members": members, "steps": steps,
           "cells": cells, "cell_updates_per_s": cells * steps / wall,
           "hbm_gbs_algorithmic": cells * steps * 100.0 / wall / 1e9,
           "tile_rows": {str(k): v for k, v in rows.items()},
           "all_members_finished": all(s == steps for s in steps_done),
           "workload": "%d Gaussian-source members, sides uniform in %s, dx=10 m, fractal shelf bed, sediment on, "
                       "walls, FP32 state (compensated h/H/hc), %d steps each" % (members, list(ENSEMBLE_SIDES), steps)}
    log("[bench] ensemble: %.1f s wall, %.0f sims/hr, %.3f G cell-updates/s"
        % (wall, res["sims_per_hr"], res["cell_updates_per_s"] / 1e9))
    return res


BF16_MFMA_PEAK_TFLOPS = 2500.0


def run_reference_example():
    """The reference's own example (its data.dat: westport bathymetry and
    NCTR series, 100 x 100 cells, 7200 s, output every 20 s), from the loaded
    inputs committed as tests/golden/westport_100x100.npz: the GPU-resident
    run to 7200 s (cell-updates/s), then the whole TSUNER() call with its
    NetCDF output (wall seconds as TSUNER reports them)."""
    import tempfile
    from tsuner_amd import scenarios as S
    from tsuner_amd.solver import Solver, run_tsuner
    z = np.load(os.path.join(ROOT, "tests", "golden", "westport_100x100.npz"))
    dd = dict(zip([str(k) for k in z["data_dat_keys"]], z["data_dat"]))
    npx, npy = int(dd["npx"]), int(dd["npy"])
    xmin, xmax, ymin, ymax, dx, dy = (float(v) for v in z["extent"])
    p = S.base_params(npx, npy, dx, dy, friction=dd["FR"], R=dd["R"], DS=dd["DS"], porosity=dd["porosity"],
                      OC=dd["OC"], Dsg=dd["Dsg"], CFL=dd["CFL"], epsilon_h=dd["epsilon_h"],
                      total_time=dd["total_time"], save_time=dd["save_time"])
    p.update(xmin=xmin, xmax=xmax, ymin=ymin, ymax=ymax, dx=dx, dy=dy)
    V = S.pack_state(z["h"], 0 * z["h"], 0 * z["h"], z["H"])
    scheme = S.default_scheme(bc=[S.BC_INFLOW] * 4)
    total = float(dd["total_time"])
    log("[bench] reference example: %dx%d, %.0f s" % (npx, npy, total))
    g = Solver(p, scheme, V, z["fixH"], z["CC"], z["tCC"])
    g.set_stop_time(total)
    g.synchronize()
    t0 = time.perf_counter()
    while g.t < total:
        g.step(500)
    wall = time.perf_counter() - t0
    steps = g.steps
    g.close()
    with tempfile.TemporaryDirectory() as d, _StdoutToStderr():
        code, secs = run_tsuner(p, V, z["fixH"], os.path.join(d, "sim"), z["CC"], z["tCC"])
    return {"workload": "data.dat example: westport, %dx%d cells (dx %.1f m, dy %.1f m), %.0f s, inflow series "
                        "on all sides, sediment on, FP64" % (npx, npy, dx, dy, total),
            "steps": steps, "gpu_resident_s": wall, "cell_updates_per_s": npx * npy * steps / wall,
            "tsuner_call_s": secs if code == 0 else None,
            "tsuner_call": "TSUNER() with NetCDF output every %.0f s" % dd["save_time"]}


def run_correction(heldout, fine_steps):
    """Coarse + learned correction against the fine solver on held-out
    Gaussian-source scenarios (1024^2 fine, 256^2 coarse): the configured
    network (four 3x3 convolutions, 32 channels) as the headline row, the
    wider shipped one (six convolutions, 64 channels) beside it, both against
    one fine run; plus each network's forward throughput on the MFMA."""
    from tsuner_amd import correction as K
    nets = {}
    for name, path in (("conv4_w32", K.WEIGHTS_4X32), ("conv6_w64", K.WEIGHTS)):
        if os.path.exists(path):
            nets[name] = K.load_net(path)
        else:
            log("[bench] correction: no trained weights at %s" % path)
    if not nets:
        return None
    keys = ("peak_eta_err_mean_m", "frac_peak_eta_within_5cm", "peak_eta_sea_err_mean_m",
            "frac_peak_eta_sea_within_5cm", "nearshore_mean_abs_eta_err_m", "runup_rel_err_mean",
            "frac_runup_within_10pct", "bed_change_mae_mean_m", "bed_change_fine_mean_abs_m", "sims_per_hr",
            "speedup_vs_fine", "fallback_frac", "threshold")

    def pick(d):
        return {k: d[k] for k in keys if k in d}

    def network(net, info):
        tn = K.time_network(net, 32)
        log("[bench] correction network %s: %.3f ms per 32 images, %.1f TFLOP/s"
            % (net.arch.as_dict(), tn["ms"], tn["tflops"]))
        return {"roofline": {"bound": "mfma", "achieved": tn["tflops"], "peak": BF16_MFMA_PEAK_TFLOPS,
                             "unit": "TFLOP/s", "frac": tn["tflops"] / BF16_MFMA_PEAK_TFLOPS, "traffic": None},
                "forward_ms_per_32_images": tn["ms"], "flop_per_pixel": tn["flop_per_pixel"],
                "trained_on_pairs": info.get("pairs"), "arch": net.arch.as_dict()}

    names = list(nets)
    head = names[0]
    net, scale_out, info = nets[head]
    # each network's eta_max readout, trained on that network's runs
    from tsuner_amd import eta_readout
    ro_paths = {"conv4_w32": eta_readout.WEIGHTS, "conv6_w64": eta_readout.WEIGHTS_6X64}
    readouts = {k: (eta_readout.load(ro_paths[k]) if k in ro_paths and os.path.exists(ro_paths[k]) else None)
                for k in names}
    others = {k: (nets[k][0], nets[k][1], readouts[k]) for k in names[1:]}
    readout = readouts[head]
    ev = K.evaluate(net, scale_out, list(range(300001, 300001 + heldout)), fine_steps=fine_steps, others=others,
                    readout=readout)
    rows = {head: {"coarse_nn": pick(ev["coarse_nn"]),
                   "fallback_policies": {k: pick(v) for k, v in ev["fallback_policies"]["policies"].items()},
                   "network": network(net, info)}}
    if readout is not None:
        rows[head]["coarse_nn_eta_readout"] = pick(ev["coarse_nn_readout"])
        rows[head]["eta_readout_fallback_policies"] = {
            k: pick(v) for k, v in ev["readout_fallback_policies"]["policies"].items()}
    for k in names[1:]:
        o = ev["others"][k]
        rows[k] = {"coarse_nn": pick(o["coarse_nn"]),
                   "fallback_policies": {p: pick(v) for p, v in o["fallback_policies"]["policies"].items()},
                   "network": network(nets[k][0], nets[k][2])}
        if "coarse_nn_readout" in o:
            rows[k]["coarse_nn_eta_readout"] = pick(o["coarse_nn_readout"])
            rows[k]["eta_readout_fallback_policies"] = {
                p: pick(v) for p, v in o["readout_fallback_policies"]["policies"].items()}
    return {"workload": "%d held-out Gaussian-source scenarios, fine 1024^2 dx=10 m (%d steps) vs coarse 256^2 "
                        "dx=40 m to the same end times, FP32; network correction every %d coarse steps"
                        % (heldout, fine_steps, K.WINDOW),
            "headline": head,
            "fine_sims_per_hr": ev["fine"]["sims_per_hr"], "coarse_alone": pick(ev["coarse_alone"]),
            "networks": rows,
            "metrics": "peak_eta_err: highest eta_max over the nearshore zone (coarse cells initially < 10 m deep, "
                       "land included) vs the fine run's block means; peak_eta_sea_err: the same over the sea part "
                       "of the zone; runup: highest fine ground reached; fallback policies: threshold chosen on the "
                       "first half of the scenarios (most fallbacks keeping >= 8x there), reported on the second; "
                       "coarse_nn_eta_readout: the same run's eta_max read out near the zone peak by the learned "
                       "per-cell readout (tsuner_amd/eta_readout.py) for the eta_max metrics, runup unchanged"}


def calibrate(nbytes, reps=3):
    """The library's plain copy kernel on nbytes (read and written per
    launch): GB/s counting both directions, against the HBM peak."""
    import ctypes as C
    from tsuner_amd import _lib
    out = C.c_double()
    _lib.check(_lib.lib().tsu_bandwidth_probe(0, int(nbytes), int(reps), C.byref(out)), "tsu_bandwidth_probe")
    log("[bench] copy kernel: %d bytes x %d: %.1f GB/s" % (nbytes, reps, out.value))
    return {"copy_kernel_gbs": out.value, "bytes_each_way_per_launch": int(nbytes), "launches": reps,
            "frac_of_peak": out.value / HBM_PEAK_GBS}


def cpu_baseline(sc, budget_s=12.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    rows, cols = min(256, sc.npy), min(1024, sc.npx)
    p = dict(sc.params, npx=cols, npy=rows, nvol=cols * rows, xmax=cols * sc.params["dx"],
             ymax=rows * sc.params["dy"])
    s = dict(sc.scheme, bc=[1, 1, 1, 1])
    V = np.ascontiguousarray(sc.V[:, :rows, :cols])
    fix = np.ascontiguousarray(sc.fixH[:rows, :cols])
    o = O.OracleSim(p, s, V, fix)
    t0 = time.perf_counter()
    o.step(1)
    one = time.perf_counter() - t0
    n = int(max(1, min(200, budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    o.step(n)
    el = time.perf_counter() - t0
    o.close()
    return {"value": rows * cols * n / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "CPU oracle (serial FP64 C restatement, gcc -O2 -ffp-contract=off) on the first %d rows x %d "
                      "columns of the same initial state (walls), %d steps after 1 warm-up step; host CPU: %s"
                      % (rows, cols, n, cpu_model())}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-active", action="store_true",
                    help="skip the active-flow variant (every wet cell moving, sediment suspended)")
    ap.add_argument("--no-small", action="store_true", help="skip the 512^2 and 2048^2 single-grid workloads")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 256-member fine ensemble")
    ap.add_argument("--ensemble-members", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=3000)
    ap.add_argument("--no-correction", action="store_true", help="skip the coarse + learned-correction leg")
    ap.add_argument("--no-example", action="store_true", help="skip the reference's data.dat example case")
    ap.add_argument("--correction-scenarios", type=int, default=200)
    ap.add_argument("--calibrate", nargs="?", const=1e9, type=float, default=None, metavar="BYTES",
                    help="only run the plain copy kernel (BYTES read and written per launch, 3 launches after a "
                         "warm-up) and print its bandwidth: the achievable-bandwidth reference and the fixed "
                         "traffic for calibrating byte counters")
    args = ap.parse_args()
    if args.calibrate is not None:
        print(json.dumps(calibrate(int(args.calibrate))), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this solver path runs on a single MI355X and no multi-GPU "
              "figure is produced (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from tsuner_amd import scenarios as S
    log("[bench] generating roofline domain %d^2" % args.size)
    t0 = time.perf_counter()
    sc = S.roofline_domain(n=args.size, dx=5.0 * 8192 / args.size)
    log("[bench] generated in %.1f s" % (time.perf_counter() - t0))

    r64 = run_variant(sc, "fp64", args.steps, args.warmup, args.tile_rows)
    r32 = None if args.no_fp32 else run_variant(sc, "fp32", args.steps, args.warmup, args.tile_rows)
    active = None
    if not args.no_active:
        # Same domain with flow everywhere: the configured run is mostly still
        # water during its first 1000 steps, this one exercises every branch.
        sa = S.Scenario(sc.name + "_active", sc.params, sc.scheme, sc.V.copy(), sc.fixH, sc.CC, sc.tCC)
        sa.V[1] = sa.V[0] * 1.0
        sa.V[2] = sa.V[0] * 0.25
        sa.V[4] = sa.V[0] * 1e-3
        a64 = run_variant(sa, "fp64", args.steps, args.warmup, args.tile_rows)
        a32 = None if args.no_fp32 else run_variant(sa, "fp32", args.steps, args.warmup, args.tile_rows)
        active = {"fp64": a64["value"], "fp64_ms_per_step": a64["ms_per_step"], "fp64_roofline": a64["roofline"],
                  "fp32": None if a32 is None else a32["value"],
                  "fp32_ms_per_step": None if a32 is None else a32["ms_per_step"],
                  "fp32_roofline": None if a32 is None else a32["roofline"],
                  "state": "initial state plus u = 1 m/s, v = 0.25 m/s and c = 1e-3 in every wet cell"}
        del sa
    small = None if args.no_small else run_small_grids(args.warmup)
    example = None
    if not args.no_example:
        try:
            example = run_reference_example()
        except Exception as e:   # informative leg: reported, never fatal to the headline line
            log("[bench] reference example failed: %r" % (e,))
    prec_err = None if (args.no_small or args.no_fp32) else run_fp32_vs_fp64()
    config_laws = None
    if not args.no_small:
        # the configurations' own laws where they differ from the reference's:
        # Grass bedload (A_g = 1e-3) on the roofline domain, minmod limiter and
        # Manning friction (n = 0.025) on the 512^2 runup beach
        sg = S.Scenario(sc.name + "_grass", sc.params, dict(sc.scheme, bedload_law=S.BEDLOAD_GRASS, grass_ag=1e-3),
                        sc.V, sc.fixH, sc.CC, sc.tCC)
        rg = run_variant(sg, "fp64", args.steps, args.warmup, args.tile_rows)
        sm = S.plain_swe(n=512, dx=20.0, steps=2000, limiter=S.LIM_MINMOD, friction_law=S.FRIC_MANNING,
                         manning_n=0.025)
        rm = run_variant(sm, "fp64", sm.steps, args.warmup, 0)
        config_laws = {
            "roofline_grass_bedload_fp64": {"value": rg["value"], "ms_per_step": rg["ms_per_step"],
                                            "roofline_frac": rg["roofline"]["frac"]},
            "runup_beach_512_minmod_manning_fp64": {"value": rm["value"], "ms_per_step": rm["ms_per_step"]},
            "note": "the headline and small_grids use the reference's laws (van Leer, Ashida-Michiue, quadratic "
                    "friction); these rows use the configurations' stated laws"}
        del sg
    ensemble = None
    if not args.no_ensemble:
        ensemble = run_ensemble(args.ensemble_members, args.ensemble_steps)
    correction = None
    if not args.no_correction:
        correction = run_correction(args.correction_scenarios, args.ensemble_steps)
    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(sc)
        log("[bench] cpu %.3e cell-updates/s" % cpu["value"])

    out = {
        "metric": METRIC,
        "value": r64["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded k^-3 fractal shelf bed, 5% non-erodible mask, 3 m / 600 s N-wave inflow at x=0",
        "config": {"workload": "roofline domain %dx%d, dx=%g m, sediment on (Ashida-Michiue bedload + "
                               "erosion/deposition), MUSCL(van Leer) + HLL + RK2, FP64, %d steps"
                               % (args.size, args.size, sc.params["dx"], args.steps),
                   "cells": sc.npx * sc.npy, "bytes_per_cell_update": 200},
        "roofline": r64["roofline"],
        "cpu_baseline": cpu,
        "fp32": None if r32 is None else {"value": r32["value"], "ms_per_step": r32["ms_per_step"],
                                          "dtype": "f32 state and fluxes, compensated h/H/hc (+60 B/cell-update lo planes)",
                                          "bytes_per_cell_update": 100,
                                          "roofline": r32["roofline"],
                                          "stage_ms_mean": r32["stage_ms_mean"],
                                          "pcie_inclusive": r32["pcie_inclusive"]},
        "stage_ms_mean": r64["stage_ms_mean"],
        "pcie_inclusive": r64["pcie_inclusive"],
        "active_flow": active,
        "small_grids": small,
        "reference_example": example,
        "fp32_vs_fp64": prec_err,
        "config_laws": config_laws,
        "ensemble_fine": ensemble,
        "ensemble_coarse_nn": correction,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
