"""The single large domain at its configured size: 8192 x 8192 cells (dx 5 m,
fractal shelf bed, sediment on), FP64 and FP32.

The reference cannot hold this grid at all: its byte sizes are `int`
products (GPU/main_TSUNER.cu:46,60-62), which overflow above 53.7 M cells.
The library's offsets are 64-bit; here planes 3 and 4 of every state buffer
lie beyond byte 2^31.  Checked:
- FP64, one full-domain step: windows of the result equal the oracle run on
  the same window with the GPU's dt, bit for bit, on cells at least four cells
  (two stages of a two-cell stencil) inside the window's cut edges; one
  window holds the last rows;
- lake at rest stays at rest (FP64 bit for bit, FP32 in its hi/lo pairs);
- total water and total sediment (H + hc) over 10 steps with walls.
"""
import sys

import numpy as np
import pytest

from tsuner_amd import scenarios as S
from tsuner_amd.solver import Solver

pytestmark = pytest.mark.gpu

N = 8192
DX = 5.0
HALO = 4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def active_domain():
    """The roofline domain with flow in every wet cell (u = 0.8 m/s,
    v = -0.3 m/s, c = 1e-3), so that one step changes every wet cell and
    exercises the Riemann, bedload and exchange paths everywhere."""
    say("generating %d^2 domain" % N)
    sc = S.roofline_domain(n=N, dx=DX)
    sc.V[1] = 0.8 * sc.V[0]
    sc.V[2] = -0.3 * sc.V[0]
    sc.V[4] = 1e-3 * sc.V[0]
    return sc


def _window(sc, r0, r1, c0, c1, top_wall):
    p = dict(sc.params, npx=c1 - c0, npy=r1 - r0, nvol=(c1 - c0) * (r1 - r0),
             xmin=c0 * DX, xmax=c1 * DX, ymin=r0 * DX, ymax=r1 * DX)
    s = dict(sc.scheme, bc=[S.BC_WALL] * 4)
    V = np.ascontiguousarray(sc.V[:, r0:r1, c0:c1])
    fix = np.ascontiguousarray(sc.fixH[r0:r1, c0:c1])
    return p, s, V, fix


def test_fp64_step_windows_equal_the_oracle(active_domain, oracle):
    sc = active_domain
    say("fp64 %d^2: upload and one step" % N)
    g = Solver.from_scenario(sc)
    dt0 = g.dt
    g.step(1)
    t1 = g.t
    assert t1 == dt0
    W = g.state()
    g.close()
    assert np.all(np.isfinite(W))
    windows = [(4000, 4096, 3000, 3096, False),       # deep shelf
               (2000, 2096, 6800, 6896, False),       # near the shoreline
               (N - 96, N, 5000, 5096, True)]         # last rows: the top wall of the domain
    for r0, r1, c0, c1, top in windows:
        say("window rows %d-%d cols %d-%d" % (r0, r1, c0, c1))
        p, s, V, fix = _window(sc, r0, r1, c0, c1, top)
        o = oracle.OracleSim(p, s, V, fix)
        o.set_time(0.0, dt0)
        o.step(1)
        A = o.state()
        rs = slice(HALO, (r1 - r0) if top else (r1 - r0 - HALO))
        cs = slice(HALO, c1 - c0 - HALO)
        got = W[:, r0:r1, c0:c1][:, rs, cs]
        want = A[:, rs, cs]
        assert np.abs(want - V[:, rs, cs]).max() > 0, "the step changed nothing in this window"
        if not np.array_equal(got, want):
            d = np.abs(got - want)
            f, j, i = np.unravel_index(np.argmax(d), d.shape)
            raise AssertionError("window %r: max |diff| %.3e at plane %d (%d, %d); %d cells differ"
                                 % ((r0, c0), d[f, j, i], f, j, i, int((got != want).sum())))


def _conservation(sc, precision, steps=10):
    g = Solver.from_scenario(sc, precision=precision)
    m0 = g.mass()
    g.step(steps)
    m1 = g.mass()
    W = g.state()
    g.close()
    return m0, m1, W


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_water_and_sediment_conserved(precision):
    """Walls on every side, a 2 m Gaussian hump in deep water on the shelf,
    flow and suspended sediment over the deep half (x < 20 km, 16 m of water
    or more): Sum h and Sum (H + hc) stay within round-off over 10 steps (no
    cell wets, dries or runs out of suspended load in that time)."""
    say("conservation %s %d^2" % (precision, N))
    sc = S.roofline_domain(n=N, dx=DX, inflow=False)
    x = (np.arange(N) + 0.5) * DX
    hump = 2.0 * np.exp(-((x - 8000.0) ** 2) / (2 * 600.0 ** 2))
    deep = (x < 20000.0)[None, :] & (sc.V[0] > 0)
    sc.V[0] = np.where(deep, sc.V[0] + hump[None, :], sc.V[0])
    sc.V[1] = np.where(deep, 0.3 * sc.V[0], 0.0)
    sc.V[4] = np.where(deep, 1e-3 * sc.V[0], 0.0)
    m0, m1, W = _conservation(sc, precision)
    dh, ds = (m1[0] - m0[0]) / m0[0], (m1[1] - m0[1]) / m0[1]
    say("%s relative drift: water %.3e, sediment %.3e" % (precision, dh, ds))
    assert abs(dh) <= 1e-12 and abs(ds) <= 1e-12, (dh, ds)
    assert np.abs(W[3] - sc.V[3]).max() > 0      # the bed moved


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lake_at_rest(precision):
    say("lake at rest %s %d^2" % (precision, N))
    sc = S.lake_at_rest(n=N, dx=DX, seed=3)
    g = Solver.from_scenario(sc, precision=precision)
    g.step(10)
    W = g.state()
    g.close()
    if precision == "fp64":
        assert np.array_equal(W, sc.V)
    else:
        def pair(v):
            hi = v.astype(np.float32)
            return hi.astype(np.float64) + (v - hi.astype(np.float64)).astype(np.float32).astype(np.float64)
        assert np.all(W[1] == 0) and np.all(W[2] == 0)
        assert np.array_equal(W[0], pair(sc.V[0]))
        assert np.array_equal(W[3], pair(sc.V[3]))
