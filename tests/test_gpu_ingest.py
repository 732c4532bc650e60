"""Real-data input end to end on the GPU: longitude/latitude input files
(written here as NetCDF classic) -> tsu_load_input_files -> the solver with
the loaded boundary series on all sides (VolumeSolver.cu:172-215), against
the oracle running the oracle-loaded case; FP64 results value-equal.  Then
the `tsuner data.dat` program on a data.dat naming such files."""
import os
import subprocess
import sys

import numpy as np
import pytest

import pyingest as OI
from tsuner_amd import ingest
from tsuner_amd import scenarios as S
from tsuner_amd.solver import Solver

from ingest_files import coastal_case, westport_scenario

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(HERE), "tsuner_amd", "tsuner")
EPS = np.finfo(float).eps


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_case(b, e, npx, npy, slr, scheme):
    d = OI.load_problem_files(b, e, "0", slr, npx, npy)
    p = dict(S.base_params(npx, npy, 1.0, 1.0))
    for k in ("xmin", "xmax", "ymin", "ymax", "dx", "dy"):
        p[k] = d[k]
    V = S.pack_state(d["h"], 0 * d["h"], 0 * d["h"], d["H"])
    return p, V, d["fixH"], d["CC"], d["tCC"]


@pytest.mark.parametrize("npx,npy", [(96, 80), (131, 67)])
def test_loaded_case_matches_oracle(tmp_path, oracle, npx, npy):
    b, e = coastal_case(tmp_path, T=8, seed=npx)
    sc = ingest.scenario_from_files(b, e, S.base_params(npx, npy, 1.0, 1.0), sea_level_correction=0.1)
    p, V, fix, CC, tCC = oracle_case(b, e, npx, npy, 0.1, sc.scheme)
    np.testing.assert_array_equal(sc.V, V)
    np.testing.assert_array_equal(sc.CC, CC)
    say("ingest case %dx%d, dx %.3f dy %.3f, %d time levels" % (npx, npy, p["dx"], p["dy"], len(tCC)))
    o = oracle.OracleSim(p, sc.scheme, V, fix, CC, tCC)
    g = Solver.from_scenario(sc)
    assert g.dt == o.dt
    for c in range(4):   # t passes tCC[T-2] = 210 s: inflow, then walls
        o.step(60)
        g.step(60)
        a, w = o.state(), g.state()
        assert np.isfinite(a).all()
        np.testing.assert_array_equal(a, w, err_msg="after %d steps" % (60 * (c + 1)))
        t, dt, _ = g.time()
        assert t == o.t and dt == o.dt
    assert o.t > tCC[-2]


def test_westport_example_matches_oracle(oracle):
    """The reference's own example (data.dat: westport bathymetry and NCTR
    series on 100 x 100 cells, committed as the loaded fixture): 400 steps,
    about 1580 s, inflow until 270 s, wet/dry coast and moving bed."""
    sc = westport_scenario()
    say("westport 100x100 from fixture")
    o = oracle.from_scenario(sc)
    g = Solver.from_scenario(sc)
    assert g.dt == o.dt
    for c in range(4):
        o.step(100)
        g.step(100)
        a, w = o.state(), g.state()
        assert np.isfinite(a).all()
        np.testing.assert_array_equal(a, w, err_msg="westport after %d steps" % (100 * (c + 1)))
        t, dt, _ = g.time()
        assert t == o.t and dt == o.dt
    assert o.t > sc.tCC[-2] and np.abs(a[4]).max() > 0   # past the inflow window, sediment in suspension


def test_westport_example_fp32_accuracy(oracle):
    """The FP32 solver (compensated h, H, hc) on the reference's example
    against the FP64 oracle after 400 steps: time, free surface on wet cells
    and momentum relative to their scales."""
    sc = westport_scenario()
    o = oracle.from_scenario(sc)
    g = Solver.from_scenario(sc, precision="fp32")
    o.step(400)
    g.step(400)
    a, b = o.state(), g.state()
    t, _, _ = g.time()
    wet = (a[0] > 1e-3) & (b[0] > 1e-3)
    eta_err = np.abs((a[0] - a[3]) - (b[0] - b[3]))[wet].max()
    qscale = max(np.abs(a[1]).max(), np.abs(a[2]).max())
    q_err = max(np.abs(a[1] - b[1]).max(), np.abs(a[2] - b[2]).max()) / qscale
    H_err = np.abs(a[3] - b[3]).max()
    say("westport fp32: t %.6f vs %.6f, eta err %.3e m, q rel err %.3e, H err %.3e m, wet cells %d vs %d"
        % (t, o.t, eta_err, q_err, H_err, int((a[0] > 0).sum()), int((b[0] > 0).sum())))
    # measured on MI355X: t rel 2.5e-9, eta 0.94 mm, q 1.3e-5, H 0.99 mm
    assert abs(t - o.t) < 1e-6 * o.t
    assert eta_err < 5e-3 and q_err < 1e-4 and H_err < 5e-3


def test_cli_reads_input_files(tmp_path, oracle):
    b, e = coastal_case(tmp_path, T=6, seed=4)
    npx, npy, total, save = 48, 40, 120.0, 40.0
    lines = ["# data.dat with input files", "# xmin", "0", "# xmax", "1", "# ymin", "0", "# ymax", "1",
             "# total time", str(total), "# save time", str(save), "# npx", str(npx), "# npy", str(npy),
             "# CFL", "0.9", "# epsilon_h", "1e-4", "# R", "2.65", "# DS", "2e-4", "# friction", "2.5e-2",
             "# porosity", "0.4", "# critical theta", "0.0047", "# Dsg", "100e-6", "# output", "coast_run",
             "# bathymetry", os.path.basename(b), "# eta", os.path.basename(e), "# fixed bed", "0",
             "# sea level", "0.2", "# tides", "0"]
    (tmp_path / "data.dat").write_text("\n".join(lines) + "\n")
    say("cli with input files %dx%d" % (npx, npy))
    r = subprocess.run([CLI, "data.dat", "0"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0 and "Normal End." in r.stdout
    scheme = S.default_scheme(extrema=1, bc=[S.BC_INFLOW] * 4)
    p, V, fix, CC, tCC = oracle_case(b, e, npx, npy, 0.2, scheme)
    p["total_time"], p["save_time"] = total, save
    o = oracle.OracleSim(p, scheme, V, fix, CC, tCC)
    recs, times = [], []
    nxt = 0.0
    while o.t < total:   # GPU/main_TSUNER.cu:157-209 save cadence
        if o.t >= nxt:
            recs.append(o.state())
            times.append(o.t)
            nxt += save
        o.step(1)
    recs.append(o.state())
    times.append(o.t)
    from scipy.io import netcdf_file
    with netcdf_file(str(tmp_path / "coast_run.nc"), "r", mmap=False) as f:
        np.testing.assert_array_equal(f.variables["time"][:], times)
        for k, W in enumerate(recs):
            np.testing.assert_array_equal(f.variables["eta"][k], W[0] - W[3])
            np.testing.assert_array_equal(f.variables["ux"][k], W[1] / (W[0] + EPS))
            np.testing.assert_array_equal(f.variables["b"][k], -W[3])


def test_cli_rejects_tides_input(tmp_path):
    """A tides file in data.dat (the reference's unfinished tides input) is
    refused with a message and a failing exit status, before any run."""
    b, e = coastal_case(tmp_path, T=4, seed=6)
    lines = ["# data.dat with a tides file", "# xmin", "0", "# xmax", "1", "# ymin", "0", "# ymax", "1",
             "# total time", "60", "# save time", "20", "# npx", "32", "# npy", "24",
             "# CFL", "0.9", "# epsilon_h", "1e-4", "# R", "2.65", "# DS", "2e-4", "# friction", "2.5e-2",
             "# porosity", "0.4", "# critical theta", "0.0047", "# Dsg", "100e-6", "# output", "tide_run",
             "# bathymetry", os.path.basename(b), "# eta", os.path.basename(e), "# fixed bed", "0",
             "# sea level", "0", "# tides", "tides.nc"]
    (tmp_path / "data.dat").write_text("\n".join(lines) + "\n")
    r = subprocess.run([CLI, "data.dat", "0"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "tides" in r.stderr
    assert not (tmp_path / "tide_run.nc").exists()
