"""The `tsuner data.dat [device]` program end to end: data.dat parsing, the
TSUNER() time loop with its save cadence, and the NetCDF output, checked
against the CPU oracle running the same loop (GPU/main_TSUNER.cu:157-209)."""
import os
import subprocess

import numpy as np
import pytest

from tsuner_amd import scenarios as S
from tsuner_amd import _lib

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FIXTURE = os.path.join(HERE, "golden", "data_small.dat")
CLI = os.path.join(ROOT, "tsuner_amd", "tsuner")
EPS = np.finfo(float).eps


def reference_loop(oracle, p, V, fix, total, save):
    """Snapshots of the reference time loop: save whenever t >= next save
    time before stepping, and once more at the end."""
    s = S.default_scheme(extrema=1)
    o = oracle.OracleSim(p, s, V, fix)
    recs, times = [], []
    nxt = 0.0
    while o.t < total:
        if save >= 0 and o.t >= nxt:
            recs.append(o.state())
            times.append(o.t)
            nxt += save
        o.step(1)
    recs.append(o.state())
    times.append(o.t)
    return recs, times, o.extrema()


def test_cli_state_file_matches_oracle(tmp_path, oracle):
    import ctypes as C
    from scipy.io import netcdf_file
    p = _lib.SimParams()
    pre = C.create_string_buffer(256)
    files = C.create_string_buffer(1024)
    assert _lib.lib().tsu_parse_data_dat(FIXTURE.encode(), C.byref(p), pre, 256, files, 1024, None) == 0
    params = {name: getattr(p, name) for name, _ in _lib.SimParams._fields_}
    npx, npy = p.npx, p.npy
    x = p.xmin + 0.5 * p.dx + np.arange(npx) * p.dx
    y = p.ymin + 0.5 * p.dy + np.arange(npy) * p.dy
    X, Y = np.meshgrid(x, y)
    H = 1.5 + 0.3 * np.sin(X) - 0.4 * X
    h = np.maximum(0.0, H + 0.6 * np.exp(-((X + 1.5) ** 2 + Y ** 2)))
    V = S.pack_state(h, 0.2 * h, 0.0 * h, H)
    fix = np.ones((npy, npx), np.int32)
    fix[:, :10] = 0
    state = tmp_path / "state.bin"
    with open(state, "wb") as f:
        f.write(np.ascontiguousarray(V).tobytes())
        f.write(fix.tobytes())
    r = subprocess.run([CLI, FIXTURE, "0", "--state", str(state)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "Normal End." in r.stdout
    recs, times, ext = reference_loop(oracle, params, V, fix, p.total_time, p.save_time)
    with netcdf_file(str(tmp_path / "small_run.nc"), "r", mmap=False) as f:
        np.testing.assert_array_equal(f.variables["time"][:], times)
        for k, W in enumerate(recs):
            np.testing.assert_array_equal(f.variables["eta"][k], W[0] - W[3])
            np.testing.assert_array_equal(f.variables["ux"][k], W[1] / (W[0] + EPS))
            np.testing.assert_array_equal(f.variables["uy"][k], W[2] / (W[0] + EPS))
            np.testing.assert_array_equal(f.variables["b"][k], -W[3])
            np.testing.assert_array_equal(f.variables["c"][k], W[4] / (W[0] + EPS))
        np.testing.assert_array_equal(f.variables["hmax"][:], ext[0])
        np.testing.assert_array_equal(f.variables["bmin"][:], -ext[1])
        np.testing.assert_array_equal(f.variables["eta_max"][:], ext[2])


def test_cli_analytic_example_runs(tmp_path):
    r = subprocess.run([CLI, FIXTURE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    from scipy.io import netcdf_file
    with netcdf_file(str(tmp_path / "small_run.nc"), "r", mmap=False) as f:
        eta = f.variables["eta"][:]
        assert eta.shape[1:] == (64, 80) and np.all(np.isfinite(eta))


def test_cli_missing_input():
    r = subprocess.run([CLI, "/nonexistent/data.dat"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "not found" in r.stderr


@pytest.mark.parametrize("save_frac", [0.0, 0.3, 1.7])
def test_tsuner_save_interval_around_dt(tmp_path, oracle, save_frac):
    """Save intervals of zero, shorter than dt and longer than dt: the
    reference loop saves at most once per pass and then takes one step
    (GPU/main_TSUNER.cu:157-197), so with an interval shorter than dt every
    step writes a record of a new state; records and times must equal the
    oracle loop's."""
    from scipy.io import netcdf_file
    from tsuner_amd.solver import run_tsuner
    n, m, dx = 40, 32, 10.0
    x, y = S.cell_centres(n, m, dx, dx)
    X, Y = np.meshgrid(x, y)
    H = 5.0 - 0.02 * X + 0.3 * np.cos(Y / 40.0)
    h = np.maximum(0.0, H + 0.8 * np.exp(-((X - 150) ** 2 + (Y - 160) ** 2) / 2000.0))
    V = S.pack_state(h, 0.3 * h, 0 * h, H)
    fix = np.ones((m, n), np.int32)
    dt0 = oracle.OracleSim(S.base_params(n, m, dx, dx), S.default_scheme(), V, fix).dt
    p = S.base_params(n, m, dx, dx, total_time=12 * dt0, save_time=save_frac * dt0)
    code, secs = run_tsuner(p, V, fix, str(tmp_path / "run"))
    assert code == 0
    recs, times, ext = reference_loop(oracle, p, V, fix, p["total_time"], p["save_time"])
    with netcdf_file(str(tmp_path / "run.nc"), "r", mmap=False) as f:
        np.testing.assert_array_equal(f.variables["time"][:], times)
        assert len(times) == len(recs) and (save_frac > 1 or len(times) >= 8)
        for k, W in enumerate(recs):
            np.testing.assert_array_equal(f.variables["eta"][k], W[0] - W[3])
            np.testing.assert_array_equal(f.variables["ux"][k], W[1] / (W[0] + EPS))
        np.testing.assert_array_equal(f.variables["eta_max"][:], ext[2])
