"""Parity of the HIP path (through the libtsuner C ABI) with the CPU oracle.

FP64 results must equal the oracle's as values (+0 == -0): the kernels keep
the reference's operation order and are built without FMA contraction.
"""
import sys

import numpy as np
import pytest

from tsuner_amd import scenarios as S
from tsuner_amd import _lib
from tsuner_amd.solver import Solver

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cmp_state(a, b, what=""):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        f, j, i = np.unravel_index(np.nanargmax(d), d.shape)
        raise AssertionError("%s: max |diff| %.3e at plane %d cell (%d,%d): %r vs %r; %d cells differ"
                             % (what, d[f, j, i], f, j, i, a[f, j, i], b[f, j, i], int((a != b).sum())))


def run_pair(sc, steps, chunks=1, tile_rows=0, tol=None):
    import pyoracle as O
    say("case %s steps=%d tile_rows=%d" % (sc.name, steps, tile_rows))
    o = O.from_scenario(sc)
    g = Solver.from_scenario(sc, tile_rows=tile_rows)
    assert g.dt == o.dt, "initial dt %r vs %r" % (g.dt, o.dt)
    per = steps // chunks
    for c in range(chunks):
        o.step(per)
        g.step(per)
        t, dt, n = g.time()
        if tol is None:
            cmp_state(o.state(), g.state(), "%s after %d steps" % (sc.name, (c + 1) * per))
            assert t == o.t and dt == o.dt, (t, o.t, dt, o.dt)
        else:
            a, b = o.state(), g.state()
            scale = np.maximum(np.abs(a).max(axis=(1, 2), keepdims=True), 1e-30)
            err = (np.abs(a - b) / scale).max()
            assert err <= tol, "%s relative error %.3e > %.1e" % (sc.name, err, tol)
        assert n == (c + 1) * per
    return o, g


# ---------------------------------------------------------------- KATs ----

def random_faces(rng, n):
    h = rng.uniform(0, 10, n)
    r = rng.uniform(0, 1, n)
    h[r < 0.25] = 0.0
    sel = (r >= 0.25) & (r < 0.35)
    h[sel] = rng.uniform(0, 2e-6, sel.sum())
    h[(r >= 0.35) & (r < 0.37)] = 1e-6
    H = rng.uniform(-5, 20, n)
    qx = rng.normal(0, 15, n)
    qy = rng.normal(0, 15, n)
    c = rng.uniform(-0.01, 0.05, n)
    hc = h * c
    return np.stack([h, qx, qy, H, hc], axis=1)


@pytest.mark.parametrize("law, tiny_shields", [(S.BEDLOAD_AM, False), (S.BEDLOAD_GRASS, False),
                                               (S.BEDLOAD_AM, True)])
def test_kat_edge_flux(oracle, law, tiny_shields):
    """Edge fluxes on random wet/dry faces, bitwise against the oracle.  The
    Ashida-Michiue law takes the stage kernels' branch-free bedload pair for
    the configured constants; with a critical Shields number of 1e-220 the
    square roots may need operand scaling and the general bedload path runs."""
    rng = np.random.default_rng(11 + law + 5 * tiny_shields)
    n = 3000
    L, R = random_faces(rng, n), random_faces(rng, n)
    # shallow subset so bedload thresholds (h > 5 DS) and theta > theta_c are crossed
    L[:500, 0] = rng.uniform(0, 3e-3, 500)
    R[:500, 0] = rng.uniform(0, 3e-3, 500)
    normals = np.array([(1, 0), (-1, 0), (0, 1), (0, -1)], dtype=np.int32)[rng.integers(0, 4, n)]
    fixed = rng.integers(0, 2, n).astype(np.int32)
    sc = S.sediment_shelf(n=8, dx=10.0, inflow=False, bedload_law=law)
    p, s = sc.params, sc.scheme
    if tiny_shields:
        p = dict(p, OC=1e-220)
    Dm = np.zeros((n, 5))
    Dp = np.zeros((n, 5))
    L_ = np.ascontiguousarray(L)
    R_ = np.ascontiguousarray(R)
    lib = _lib.lib()
    _lib.check(lib.tsu_kat_edge_flux(_lib.make_params(p), _lib.make_scheme(s), n, _lib.ptr(L_), _lib.ptr(R_),
                                     _lib.ptr(np.ascontiguousarray(normals), _lib.C.c_int),
                                     _lib.ptr(fixed, _lib.C.c_int), _lib.ptr(Dm), _lib.ptr(Dp), 0), "kat")
    for i in range(n):
        dm, dp = oracle.edge_flux(p, s, L[i], R[i], normals[i], fixed[i])
        if not (np.array_equal(dm, Dm[i]) and np.array_equal(dp, Dp[i])):
            raise AssertionError("edge %d: L=%r R=%r n=%r fixed=%d\n oracle Dm=%r Dp=%r\n gpu Dm=%r Dp=%r"
                                 % (i, L[i], R[i], normals[i], fixed[i], dm, dp, Dm[i], Dp[i]))


def test_kat_edge_flux_waves_at_rest(oracle):
    """The solve is skipped for a wave (64 consecutive edges) in which no
    edge moves: equal hydrostatic depths and no velocity or concentration
    on either side.  Blocks of 64 edges at rest with one lane that moves by
    a depth jump, a tiny momentum, a tiny concentration or a step of the bed
    under a level surface (which is still at rest), against the oracle."""
    rng = np.random.default_rng(77)
    blocks = []
    for case in range(7):
        h = rng.uniform(0.5, 20.0, 64)
        H = rng.uniform(-5.0, 15.0, 64)
        eta = h - H
        L = np.stack([h, 0 * h, 0 * h, H, 0 * h], axis=1)
        R = L.copy()
        k = 17 + case
        if case == 1:
            R[k, 0] += 1e-3                     # depth jump
        elif case == 2:
            L[k, 1] = 1e-200                    # tiny normal momentum
        elif case == 3:
            R[k, 4] = 1e-300                    # tiny suspended load
        elif case == 4:
            R[:, 3] = H + rng.uniform(-2, 2, 64)   # bed steps under the level surface
            R[:, 0] = eta + R[:, 3]
            R[:, 0] = np.maximum(R[:, 0], 0.0)
        elif case == 5:
            L[::3, 0] = 0.0                     # dry-dry lanes among lanes at rest
            R[::3, 0] = 0.0
        elif case == 6:
            L[k, 2] = -0.5                      # tangential momentum only
        blocks.append((L, R))
    L = np.ascontiguousarray(np.concatenate([b[0] for b in blocks]))
    R = np.ascontiguousarray(np.concatenate([b[1] for b in blocks]))
    n = len(L)
    normals = np.tile(np.array([(1, 0), (0, 1), (-1, 0), (0, -1)], dtype=np.int32), (n // 4, 1))
    fixed = np.ones(n, dtype=np.int32)
    sc = S.sediment_shelf(n=8, dx=10.0, inflow=False)
    p, s = sc.params, sc.scheme
    Dm = np.zeros((n, 5))
    Dp = np.zeros((n, 5))
    lib = _lib.lib()
    _lib.check(lib.tsu_kat_edge_flux(_lib.make_params(p), _lib.make_scheme(s), n, _lib.ptr(L), _lib.ptr(R),
                                     _lib.ptr(np.ascontiguousarray(normals), _lib.C.c_int),
                                     _lib.ptr(fixed, _lib.C.c_int), _lib.ptr(Dm), _lib.ptr(Dp), 0), "kat")
    moving = 0
    for i in range(n):
        dm, dp = oracle.edge_flux(p, s, L[i], R[i], normals[i], fixed[i])
        moving += bool(np.any(dm != 0) or np.any(dp != 0))
        if not (np.array_equal(dm, Dm[i]) and np.array_equal(dp, Dp[i])):
            raise AssertionError("edge %d (block %d): L=%r R=%r n=%r\n oracle Dm=%r Dp=%r\n gpu Dm=%r Dp=%r"
                                 % (i, i // 64, L[i], R[i], normals[i], dm, dp, Dm[i], Dp[i]))
    assert moving >= 4     # the single moving lanes do produce fluxes


def _kat_pairs(p, s, L, R, normals, fixed):
    n = len(L)
    Dm = np.zeros((n, 5))
    Dp = np.zeros((n, 5))
    lib = _lib.lib()
    _lib.check(lib.tsu_kat_edge_flux_pairs(_lib.make_params(p), _lib.make_scheme(s), n,
                                           _lib.ptr(np.ascontiguousarray(L)), _lib.ptr(np.ascontiguousarray(R)),
                                           _lib.ptr(np.ascontiguousarray(normals), _lib.C.c_int),
                                           _lib.ptr(np.ascontiguousarray(fixed), _lib.C.c_int), _lib.ptr(Dm),
                                           _lib.ptr(Dp), 0), "kat pairs")
    return Dm, Dp


def _check_edges(oracle, p, s, L, R, normals, fixed, Dm, Dp, what):
    for i in range(len(L)):
        dm, dp = oracle.edge_flux(p, s, L[i], R[i], normals[i], fixed[i])
        if not (np.array_equal(dm, Dm[i]) and np.array_equal(dp, Dp[i])):
            raise AssertionError("%s, edge %d (pair %d): L=%r R=%r n=%r fixed=%d\n oracle Dm=%r Dp=%r\n"
                                 " gpu Dm=%r Dp=%r" % (what, i, i // 2, L[i], R[i], normals[i], fixed[i], dm, dp,
                                                       Dm[i], Dp[i]))


@pytest.mark.parametrize("law, tiny_shields", [(S.BEDLOAD_AM, False), (S.BEDLOAD_GRASS, False),
                                               (S.BEDLOAD_AM, True)])
def test_kat_edge_flux_pairs(oracle, law, tiny_shields):
    """Random wet/dry edges solved two at a time by one lane (the stage
    kernels' paired solve: one pass through the solver's phases with the
    wave-wide tests taken over both edges), each edge bitwise against the
    oracle."""
    rng = np.random.default_rng(31 + law + 5 * tiny_shields)
    n = 3000
    L, R = random_faces(rng, n), random_faces(rng, n)
    L[:500, 0] = rng.uniform(0, 3e-3, 500)
    R[:500, 0] = rng.uniform(0, 3e-3, 500)
    normals = np.array([(1, 0), (-1, 0), (0, 1), (0, -1)], dtype=np.int32)[rng.integers(0, 4, n)]
    fixed = rng.integers(0, 2, n).astype(np.int32)
    sc = S.sediment_shelf(n=8, dx=10.0, inflow=False, bedload_law=law)
    p, s = sc.params, sc.scheme
    if tiny_shields:
        p = dict(p, OC=1e-220)
    Dm, Dp = _kat_pairs(p, s, L, R, normals, fixed)
    _check_edges(oracle, p, s, L, R, normals, fixed, Dm, Dp, "random pairs")


def test_kat_edge_pairs_that_split_the_votes(oracle):
    """Pairs whose two edges take different sides of a wave-wide test, in
    blocks of 64 pairs (one wave): an edge dry on both sides next to a
    moving one; an edge at rest next to a moving one; an edge whose depth
    needs the scaled square root (1e-240 m) next to ordinary ones; and whole
    waves at rest.  Each edge bitwise against the oracle."""
    rng = np.random.default_rng(5)
    blocks = []
    for case in range(4):
        m = 128                                   # 64 pairs: edges 2k (a) and 2k+1 (b)
        h = rng.uniform(0.5, 20.0, m)
        H = rng.uniform(-5.0, 15.0, m)
        L = np.stack([h, 0 * h, 0 * h, H, 0 * h], axis=1)
        R = L.copy()
        moving = random_faces(rng, m)
        moving[:, 0] = rng.uniform(0.1, 10.0, m)
        if case == 0:                             # a dry on both sides, b moving
            L[0::2, 0] = 0.0
            R[0::2, 0] = 0.0
            L[1::2], R[1::2] = moving[1::2], random_faces(rng, m)[1::2]
        elif case == 1:                           # a at rest, b moving
            L[1::2], R[1::2] = moving[1::2], random_faces(rng, m)[1::2]
        elif case == 2:                           # a needs the scaled square root in some lanes
            L[:], R[:] = moving, random_faces(rng, m)
            L[0:64:8, 0] = 1e-240
            R[0:64:8, 0] = 3e-241
        # case 3: every edge at rest
        blocks.append((L, R))
    L = np.concatenate([b[0] for b in blocks])
    R = np.concatenate([b[1] for b in blocks])
    n = len(L)
    normals = np.tile(np.array([(1, 0), (0, 1), (-1, 0), (0, 1)], dtype=np.int32), (n // 4, 1))
    fixed = np.ones(n, dtype=np.int32)
    sc = S.sediment_shelf(n=8, dx=10.0, inflow=False)
    p, s = sc.params, sc.scheme
    Dm, Dp = _kat_pairs(p, s, L, R, normals, fixed)
    _check_edges(oracle, p, s, L, R, normals, fixed, Dm, Dp, "split votes")
    assert np.any(Dm[1:256:2] != 0)              # the moving edges do produce fluxes
    assert not np.any(Dm[384:]) and not np.any(Dp[384:])


def kat_point(p, s, what, a, b=None, c=None, d=None):
    n = len(a)
    arr = [np.ascontiguousarray(x, dtype=np.float64) if x is not None else None for x in (a, b, c, d)]
    outs = [np.zeros(n) for _ in range(4)]
    lib = _lib.lib()
    _lib.check(lib.tsu_kat_point(_lib.make_params(p), _lib.make_scheme(s), what, n,
                                 *[(_lib.ptr(x) if x is not None else None) for x in arr],
                                 *[_lib.ptr(o) for o in outs], 0), "kat_point")
    return outs


def test_kat_points(oracle):
    rng = np.random.default_rng(5)
    sc = S.sediment_shelf(n=8, dx=10.0, inflow=False)
    p, s = sc.params, sc.scheme
    n = 4000
    F = random_faces(rng, n)
    h, qx, qy, H, hc = F.T
    for law in (S.LIM_VANLEER, S.LIM_MINMOD):
        a = rng.normal(0, 1, n) * 10.0 ** rng.integers(-18, 2, n)
        b = rng.normal(0, 1, n) * 10.0 ** rng.integers(-18, 2, n)
        a[:50] = 0.0
        b[25:75] = -a[25:75]
        s2 = dict(s, limiter=law)
        o1 = kat_point(p, s2, 0, a, b)[0]
        ref = np.array([oracle.limiter(law, x, y) for x, y in zip(a, b)])
        np.testing.assert_array_equal(o1, ref)
    ux, uy, m, _ = kat_point(p, s, 1, h, qx, qy)
    ref = np.array([oracle.velocity(*v) for v in zip(h, qx, qy)])
    np.testing.assert_array_equal(ux, ref[:, 0])
    np.testing.assert_array_equal(uy, ref[:, 1])
    np.testing.assert_array_equal(m, np.sqrt(ref[:, 0] ** 2 + ref[:, 1] ** 2))
    o = kat_point(p, s, 2, h - 0.5, qx, qy, hc)
    ref = np.array([oracle.state_filter(*v) for v in zip(h - 0.5, qx, qy, hc)])
    for k in range(4):
        np.testing.assert_array_equal(o[k], ref[:, k])
    o1 = kat_point(p, s, 3, h, qx, qy)[0]
    ref = np.array([oracle.cell_dt(p, *v) for v in zip(h, qx, qy)])
    np.testing.assert_array_equal(o1, ref)
    o1 = kat_point(p, s, 4, h, qx, qy, np.abs(hc))[0]
    ref = np.array([oracle.erosion_rate(p, s, *v) for v in zip(h, qx, qy, np.abs(hc))])
    np.testing.assert_array_equal(o1, ref)
    un = rng.normal(0, 2, n)
    ut = rng.normal(0, 2, n)
    hh = rng.uniform(0, 3e-3, n)
    for law in (S.BEDLOAD_AM, S.BEDLOAD_GRASS):
        s2 = dict(s, bedload_law=law)
        o1 = kat_point(p, s2, 5, hh, un, ut)[0]
        ref = np.array([oracle.bedload_flux(p, s2, *v) for v in zip(hh, un, ut)])
        np.testing.assert_array_equal(o1, ref)


# ------------------------------------------------------- grid parity -------

def test_plain_swe_exact_dt_path():
    run_pair(S.plain_swe(n=64, dx=160.0), 40, chunks=2)


def test_plain_swe_narrowed_dt_multi_strip():
    # 130 x 70 cells: three 60-column strips, 9100 cells (float-narrowed dt)
    sc = S.plain_swe(n=130, ny=70, dx=80.0)
    run_pair(sc, 30, chunks=3)


def test_sediment_inflow_odd_size():
    sc = S.sediment_shelf(n=67, ny=37, dx=40.0)
    run_pair(sc, 40, chunks=2, tile_rows=5)


def test_sediment_inflow_all_sides():
    sc = S.sediment_shelf(n=61, ny=33, dx=40.0)
    sc.scheme["bc"] = [S.BC_INFLOW] * 4
    run_pair(sc, 25)


def test_periodic_both_axes():
    sc = S.sediment_shelf(n=70, ny=45, dx=30.0, inflow=False, seed=9)
    sc.scheme["bc"] = [S.BC_PERIODIC] * 4
    sc.V[1] += 1.0   # some flow across the periodic seams
    run_pair(sc, 30, tile_rows=7)


def test_free_boundaries_and_extrema():
    sc = S.gaussian_source(n=64, dx=30.0, seed=3)
    sc.scheme["bc"] = [S.BC_FREE, S.BC_WALL, S.BC_FREE, S.BC_WALL]
    sc.scheme["extrema"] = 1
    o, g = run_pair(sc, 30)
    for a, b in zip(o.extrema(), g.extrema()):
        np.testing.assert_array_equal(a, b)


def test_first_order():
    sc = S.sediment_shelf(n=50, ny=41, dx=40.0)
    sc.scheme["order"] = 1
    run_pair(sc, 31, chunks=1)


def test_minmod_grass():
    sc = S.sediment_shelf(n=65, ny=30, dx=40.0, limiter=S.LIM_MINMOD, bedload_law=S.BEDLOAD_GRASS)
    run_pair(sc, 30)


def test_manning_friction():
    # cbrt is not correctly rounded on either side: tolerance, not equality
    sc = S.plain_swe(n=64, dx=160.0, friction_law=S.FRIC_MANNING, manning_n=0.025)
    run_pair(sc, 30, tol=1e-11)


@pytest.mark.parametrize("rows", [1, 3, 16, 64])
def test_tile_height_independent(rows):
    sc = S.sediment_shelf(n=75, ny=50, dx=40.0, seed=4)
    run_pair(sc, 12, tile_rows=rows)


def test_graph_replay_matches_eager():
    sc = S.sediment_shelf(n=90, ny=60, dx=40.0, seed=6)
    a = Solver.from_scenario(sc)
    b = Solver.from_scenario(sc)
    b.set_graph_chunk(10)
    for k in range(3):
        a.step(10)
        b.step(10)
        cmp_state(a.state(), b.state(), "graph replay %d" % k)
        assert a.time() == b.time()
        b.mass()   # eager work between replays


def test_stop_time_makes_steps_noops():
    sc = S.plain_swe(n=64, dx=160.0)
    g = Solver.from_scenario(sc)
    g.step(5)
    t5 = g.t
    g.set_stop_time(t5)
    g.step(7)
    assert g.steps == 5 and g.t == t5
    g.set_stop_time(float("inf"))
    g.step(1)
    assert g.steps == 6


# ---------------------------------------------- full-size invariants --------

def test_lake_at_rest_2048():
    sc = S.lake_at_rest(n=2048)
    say("lake at rest 2048^2")
    g = Solver.from_scenario(sc)
    g.step(5)
    cmp_state(sc.V, g.state(), "lake at rest")


def test_mass_conservation_1024():
    sc = S.sediment_shelf(n=1024, dx=10.0, inflow=False, seed=8)
    sc.V[0] = np.maximum(0.0, sc.V[3] + 1.5 * np.exp(-((np.arange(1024) - 300.0) ** 2) / 2e3)[None, :])
    g = Solver.from_scenario(sc)
    m0 = g.mass()
    g.step(50)
    m1 = g.mass()
    assert abs(m1[0] - m0[0]) <= 1e-11 * m0[0], (m0, m1)
    assert abs(m1[1] - m0[1]) <= 1e-11 * abs(m0[1]), (m0, m1)


def test_transpose_symmetry_512():
    sc = S.gaussian_source(n=256, dx=40.0, seed=12)
    sct = S.gaussian_source(n=256, dx=40.0, seed=12)
    sct.V = np.ascontiguousarray(np.stack([sc.V[0].T, sc.V[2].T, sc.V[1].T, sc.V[3].T, sc.V[4].T]))
    sct.fixH = np.ascontiguousarray(sc.fixH.T)
    a = Solver.from_scenario(sc)
    b = Solver.from_scenario(sct)
    a.step(40)
    b.step(40)
    A, B = a.state(), b.state()
    Bt = np.stack([B[0].T, B[2].T, B[1].T, B[3].T, B[4].T])
    cmp_state(A, Bt, "transpose symmetry")


def test_fp32_solver_error_per_field(oracle):
    """FP32 state and fluxes against the FP64 oracle after 50 steps: the error
    of each field relative to that field's scale (momentum: max |q|)."""
    sc = S.sediment_shelf(n=96, ny=64, dx=40.0, seed=10)
    o = oracle.from_scenario(sc)
    g = Solver.from_scenario(sc, precision="fp32")
    o.step(50)
    g.step(50)
    a, b = o.state(), g.state()
    qscale = max(np.abs(a[1]).max(), np.abs(a[2]).max())
    # about ten times the measured errors (h 9.4e-8, qx 3.2e-6, qy 2.4e-6,
    # H 2.6e-11, hc 7.7e-6)
    tol = {"h": 1e-6, "qx": 3e-5, "qy": 3e-5, "H": 3e-10, "hc": 8e-5}
    for f, name in enumerate(["h", "qx", "qy", "H", "hc"]):
        scale = qscale if f in (1, 2) else max(np.abs(a[f]).max(), 1e-12)
        err = np.abs(a[f] - b[f]).max() / scale
        say("fp32 rel err %s %.3e" % (name, err))
        assert err < tol[name], (name, err)


def test_fp32_lake_at_rest_exact():
    sc = S.lake_at_rest(n=256)
    g = Solver.from_scenario(sc, precision="fp32")
    g.step(20)
    W = g.state()
    np.testing.assert_array_equal(W[1], 0.0)
    np.testing.assert_array_equal(W[2], 0.0)
    # the FP32 state holds hi + lo float pairs (48 significant bits)
    def pair(v):
        hi = v.astype(np.float32)
        return hi.astype(np.float64) + (v - hi.astype(np.float64)).astype(np.float32).astype(np.float64)
    np.testing.assert_array_equal(W[0], pair(sc.V[0]))
    np.testing.assert_array_equal(W[3], pair(sc.V[3]))


def test_fp32_compensated_conservation():
    """Deep water with suspension and bedload (no clipping): the compensated
    FP32 planes keep Sum h and Sum(H + hc) to FP64-level round-off."""
    n = 256
    p = S.base_params(n, n, 20.0, 20.0)
    s = S.default_scheme()
    x, y = S.cell_centres(n, n, 20.0, 20.0)
    X, Y = np.meshgrid(x, y)
    H = 6.0 + 0.5 * np.sin(X / 90.0) * np.cos(Y / 70.0)
    h = H + 0.8 * np.exp(-((X - 2000) ** 2 + (Y - 3000) ** 2) / 2e5)
    V = S.pack_state(h, 1.2 * h, 0.3 * h, H, 1e-3 * h)
    g = Solver(p, s, V, S.nonerodible_mask(n, n, 7), precision="fp32")
    m0 = g.mass()
    g.step(100)
    m1 = g.mass()
    say("fp32 mass drift %.3e %.3e" % ((m1[0] - m0[0]) / m0[0], (m1[1] - m0[1]) / m0[1]))
    assert abs(m1[0] - m0[0]) <= 1e-12 * m0[0]
    assert abs(m1[1] - m0[1]) <= 1e-12 * m0[1]


def test_carrier_greenspan_fp64_parity_and_fp32_accuracy(oracle):
    """Moving shoreline on a plane beach driven through the inflow series:
    the FP64 path equals the oracle bit for bit over half a period, and the
    FP32 path stays as close to the exact solution as FP64 after 1.25
    periods (mean free-surface error and runup)."""
    import math
    import analytic as AN
    A = 0.5
    sc, B, x = AN.carrier_greenspan_scenario(4.0, A=A)
    run_pair(sc, 1000, chunks=2)
    t_end = 1.25 * math.pi * B.T0
    amp = A / 4 * B.Z0
    ru_an = AN.cg_runup(A) * B.Z0
    res = {}
    for prec in ("fp64", "fp32"):
        g = Solver.from_scenario(sc, precision=prec)
        g.set_stop_time(t_end)
        while g.t < t_end:
            g.step(500)
        W = g.state()[:, 0]
        hmax = g.extrema()[0][0]
        e_an, _, wet_an = B.field(x, g.t, A)
        wet = (W[0] > 1e-3) & wet_an
        err = np.abs((W[0] - W[3]) - e_an)[wet].mean() / amp
        ru = (-sc.V[3, 0][hmax > 1e-3]).max() / ru_an - 1.0
        res[prec] = (err, ru)
        say("carrier-greenspan %s: mean eta error %.2e of amplitude, runup %+.3f" % (prec, err, ru))
        g.close()
    assert res["fp64"][0] < 2e-3 and abs(res["fp64"][1]) < 0.03
    assert res["fp32"][0] < 1.5 * res["fp64"][0] + 1e-4 and abs(res["fp32"][1]) < 0.03


def test_runup_beach_512_at_its_configured_size():
    """The 512^2 runup-beach configuration (dx 20 m, Gaussian hump, walls,
    quadratic friction, no sediment) at full size: 100 steps in four chunks
    (9 strips of 2-row tiles, float-narrowed dt), bit-identical to the oracle;
    and the sediment shelf at 256^2 with its N-wave inflow, 120 steps."""
    run_pair(S.plain_swe(n=512, dx=20.0), 100, chunks=4)
    run_pair(S.sediment_shelf(n=256, dx=10.0, seed=2), 120, chunks=2)


def test_shared_denominator_division_on_the_device():
    """The device's reciprocal (hardware estimate, two Newton refinements, one
    residual correction, no operand scaling) and shared-denominator quotient
    equal numpy's correctly rounded 1/a and b/a over [2^-60, 2^60]: random
    denominators, denominators with an all-ones significand (where a
    reciprocal refinement can land on a rounding tie), h + DBL_EPSILON forms
    and 1 + c forms; numerators down to subnormals."""
    rng = np.random.default_rng(11)
    sc = S.sediment_shelf(n=8, dx=10.0, inflow=False)
    n = 200000
    e = rng.integers(-60, 60, n)
    a = np.ldexp(1.0 + rng.random(n), e)
    ones = np.arange(n) % 4 == 0
    a[ones] = np.ldexp(2.0 - 2.0 ** -52, e[ones] - 1)            # significand all ones
    k = np.arange(n) % 4 == 1
    a[k] = np.abs(rng.normal(size=k.sum())) * 10.0 ** rng.integers(-12, 4, k.sum()) + np.finfo(float).eps
    k = np.arange(n) % 4 == 2
    a[k] = 1.0 + np.abs(rng.normal(size=k.sum())) * 10.0 ** rng.integers(-16, 12, k.sum())
    b = rng.normal(size=n) * 10.0 ** rng.integers(-30, 30, n)
    # the last quarter (whole waves) takes numerators from the smallest
    # subnormal to 2^-900, where the residual step can underflow and the
    # device falls back to the IEEE division
    t = np.arange(n) >= 3 * n // 4
    bits = (rng.integers(0, 124, t.sum()).astype(np.uint64) << np.uint64(52)) | \
        rng.integers(0, 2 ** 52, t.sum(), dtype=np.uint64) | (rng.integers(0, 2, t.sum()).astype(np.uint64) << np.uint64(63))
    b[t] = bits.view(np.float64)
    y, q, yi, qi = kat_point(sc.params, sc.scheme, 6, a, b)
    np.testing.assert_array_equal(yi, 1.0 / a)                  # the device's IEEE division
    np.testing.assert_array_equal(qi, b / a)
    bad = np.flatnonzero(y != 1.0 / a)
    assert bad.size == 0, "1/a differs at %d inputs, e.g. a=%r: %r vs %r" % (bad.size, a[bad[0]], y[bad[0]], 1.0 / a[bad[0]])
    np.testing.assert_array_equal(q, b / a)


def test_fp32_vs_fp64_per_field_sediment_shelf_2048():
    """FP32 (compensated h/H/hc) against FP64 at the same simulation time on
    the 2048^2 sediment shelf with its N-wave inflow (1000 steps, both runs
    ending on the FP64 run's time): per-field errors relative to each
    field's scale, with tolerances about ten times the measured errors."""
    import bench
    sc = S.sediment_shelf(n=2048, dx=10.0, steps=1000)
    r = bench.precision_error(sc, 1000)
    for k, v in r["fields"].items():
        say("fp32 vs fp64 %s: max %.3e rms %.3e (scale %.3g)" % (k, v["max_rel"], v["rms_rel"], v["scale"]))
    say("totals: water %.3e sediment %.3e" % (r["water_total_rel_diff"], r["sediment_total_rel_diff"]))
    # measured: h 7.2e-7, qx 8.8e-6, qy 2.8e-6, H 1.9e-8, hc 6.6e-6; totals
    # 1.4e-8 (water: the inflow faces differ) and 1e-11 (sediment)
    tol = {"h": 7e-6, "qx": 9e-5, "qy": 3e-5, "H": 2e-7, "hc": 7e-5}
    for k, v in r["fields"].items():
        assert v["max_rel"] < tol[k], (k, v)
    assert abs(r["water_total_rel_diff"]) < 1.5e-7 and abs(r["sediment_total_rel_diff"]) < 1e-9


def test_inflow_series_times_are_checked():
    """interpolate_time divides by tCC[1] - tCC[0] and truncates to a level
    index: repeated, decreasing or non-finite times are refused."""
    sc = S.sediment_shelf(n=32, ny=16, dx=40.0)
    for bad in ([0.0, 10.0, 10.0, 30.0], [0.0, -5.0, 10.0, 20.0], [0.0, np.nan, 20.0, 30.0],
                [0.0, 10.0, 20.0, np.inf]):
        tc = np.array(bad)
        cc = np.ascontiguousarray(sc.CC[:len(tc)])
        with pytest.raises(_lib.TsuError) as e:
            Solver(sc.params, sc.scheme, sc.V, sc.fixH, cc, tc)
        assert "strictly increasing" in str(e.value)
    g = Solver.from_scenario(sc)       # the scenario's own series is accepted
    g.step(3)
    assert g.steps == 3


def test_unscaled_square_root_on_the_device():
    """The solver's square root without operand scaling (used where every
    argument of a wave is 0 or at least 2^-767) equals IEEE sqrt over that
    range: log-uniform exponents, all-ones and all-zeros significands, exact
    squares and their neighbours, and the range ends."""
    rng = np.random.default_rng(17)
    sc = S.sediment_shelf(n=8, dx=10.0, inflow=False)
    n = 400000
    a = np.ldexp(1.0 + rng.random(n), rng.integers(-767, 1023, n))
    q = np.arange(n) % 8
    a[q == 1] = np.ldexp(2.0 - 2.0 ** -52, rng.integers(-767, 1023, (q == 1).sum()))
    r = rng.random((q == 2).sum()) * 1e6 + 1.0
    a[q == 2] = r * r                                          # exact squares (when representable)
    a[q == 3] = np.nextafter(a[q == 2][:(q == 3).sum()], np.inf)
    a[q == 4] = np.nextafter(a[q == 2][:(q == 4).sum()], 0.0)
    a[q == 5] = rng.random((q == 5).sum()) * 100.0             # depths and speeds of the solver
    a[:4] = [0.0, 2.0 ** -767, np.nextafter(2.0 ** -767, 1.0), np.finfo(float).max]
    got, ref = kat_point(sc.params, sc.scheme, 7, a)[:2]
    np.testing.assert_array_equal(ref, np.sqrt(a))            # the device's sqrt() is the IEEE one
    bad = np.flatnonzero(got != np.sqrt(a))
    assert bad.size == 0, "differs at %d inputs, e.g. %r" % (bad.size, a[bad[:3]])
