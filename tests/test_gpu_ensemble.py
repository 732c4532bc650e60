"""Ensembles: every member of a batched, uneven-size ensemble advances exactly
as a single-grid run of the same scenario (and, in FP64, as the oracle)."""
import numpy as np
import pytest

from tsuner_amd import scenarios as S
from tsuner_amd.ensemble import Ensemble
from tsuner_amd.solver import Solver

pytestmark = pytest.mark.gpu


def members():
    out = [S.gaussian_source(n=n, dx=40.0, seed=seed) for n, seed in ((48, 1), (70, 2), (61, 3), (130, 4))]
    sc = S.sediment_shelf(n=90, ny=45, dx=40.0, inflow=False, seed=5)
    out.append(sc)
    return out


def test_fp64_members_equal_oracle_and_single(oracle):
    scs = members()
    ens = Ensemble(scs, precision="fp64", tile_rows=[8, 16, 5, 32, 3])
    ens.step(12)
    for i, sc in enumerate(scs):
        print("member %d %s" % (i, sc.name), flush=True)
        o = oracle.from_scenario(sc)
        o.step(12)
        W = ens.state(i)
        assert np.array_equal(W, o.state()), "member %d differs from the oracle" % i
        t, dt, n = ens.time(i)
        assert (t, dt, n) == (o.t, o.dt, 12)


def test_fp64_mixed_bedload_laws_equal_oracle(oracle):
    """A member with Grass bedload makes the whole ensemble take the general
    bedload kernel; the Ashida-Michiue members must still equal the oracle
    bit for bit (their single-grid runs take the branch-free variant)."""
    scs = [S.gaussian_source(n=48, dx=40.0, seed=1),
           S.gaussian_source(n=70, dx=40.0, seed=2, bedload_law=S.BEDLOAD_GRASS, grass_ag=1e-3),
           S.gaussian_source(n=61, dx=40.0, seed=3)]
    ens = Ensemble(scs, precision="fp64", tile_rows=[8, 16, 5])
    ens.step(12)
    for i, sc in enumerate(scs):
        print("member %d %s law %d" % (i, sc.name, sc.scheme["bedload_law"]), flush=True)
        o = oracle.from_scenario(sc)
        o.step(12)
        assert np.array_equal(ens.state(i), o.state()), "member %d differs from the oracle" % i
        g = Solver.from_scenario(sc, precision="fp64", tile_rows=16)
        g.step(12)
        assert np.array_equal(ens.state(i), g.state()), "member %d differs from its single-grid run" % i


def test_fp32_members_equal_single_solver():
    scs = members()
    ens = Ensemble(scs, precision="fp32")
    ens.step(15)
    for i, sc in enumerate(scs):
        g = Solver.from_scenario(sc, precision="fp32", tile_rows=16)
        g.step(15)
        assert np.array_equal(ens.state(i), g.state()), "member %d" % i
        assert ens.time(i) == g.time()


def test_graph_replay_and_stop_step():
    scs = members()
    a = Ensemble(scs, precision="fp32")
    b = Ensemble(scs, precision="fp32", graph_chunk=4)
    b.set_stop_step(10)
    a.step(10)
    b.step(12)          # three replays of four steps; the last two steps are no-ops
    for i in range(len(scs)):
        assert b.time(i)[2] == 10
        assert np.array_equal(a.state(i), b.state(i))
        b.mass(i)


def test_autotune_keeps_results(tmp_path, monkeypatch):
    from tsuner_amd import autotune
    monkeypatch.setenv("XDG_CACHE_HOME", str(tmp_path))
    scs = members()
    ref = Ensemble(scs, precision="fp32")
    ref.step(6)
    ens = Ensemble(scs, precision="fp32")
    choice = autotune.tune_ensemble(ens, "fp32", candidates=(4, 16), steps=2, log=print)
    assert set(choice) == {max(s.npx, s.npy) for s in scs}
    ens.step(6)
    for i in range(len(scs)):
        assert np.array_equal(ref.state(i), ens.state(i))
    assert (tmp_path / "tsuner_amd" / "tile_rows.json").exists()


def inflow_members():
    """Members with boundary series of their own: two N-wave shelves of
    different sizes driven through the left side (different series lengths
    and spacings), the reference's westport example driven through all four
    sides, and a walled member beside them."""
    from ingest_files import westport_scenario
    a = S.sediment_shelf(n=64, ny=40, dx=40.0, seed=6, t_series=60.0)   # its series ends within the test
    b = S.sediment_shelf(n=97, ny=33, dx=40.0, seed=7, t_series=2000.0)
    b.tCC = b.tCC[::2].copy()
    b.CC = np.ascontiguousarray(b.CC[::2])
    c = westport_scenario()
    d = S.gaussian_source(n=50, dx=40.0, seed=8)
    return [a, b, c, d]


def test_fp64_inflow_members_equal_oracle(oracle):
    """Per-member inflow series (VolumeSolver.cu:172-215 read at each
    member's own time) in one batched launch: every member equals the
    oracle bit for bit, through and past the end of its series."""
    scs = inflow_members()
    ens = Ensemble(scs, precision="fp64", tile_rows=[8, 5, 16, 4])
    orc = [oracle.from_scenario(sc) for sc in scs]
    for chunk in range(3):
        ens.step(60)
        for i, (sc, o) in enumerate(zip(scs, orc)):
            o.step(60)
            print("chunk %d member %d %s t=%.1f" % (chunk, i, sc.name, o.t), flush=True)
            assert np.array_equal(ens.state(i), o.state()), "member %d differs from the oracle" % i
            assert ens.time(i) == (o.t, o.dt, 60 * (chunk + 1))


def test_fp32_inflow_members_equal_single_solver():
    scs = inflow_members()
    ens = Ensemble(scs, precision="fp32", graph_chunk=10)
    ens.step(40)
    for i, sc in enumerate(scs):
        g = Solver.from_scenario(sc, precision="fp32", tile_rows=16)
        g.step(40)
        assert np.array_equal(ens.state(i), g.state()), "member %d" % i
        assert ens.time(i) == g.time()


def test_inflow_member_without_series_is_refused():
    from tsuner_amd import _lib
    sc = S.sediment_shelf(n=40, ny=20, dx=40.0)
    sc.CC = None
    sc.tCC = None
    with pytest.raises(_lib.TsuError) as e:
        Ensemble([sc], precision="fp32")
    assert "tsu_ensemble_set_inflow" in str(e.value)
