"""Ensembles of scenarios on uneven grids, advanced together on one GPU.

``Ensemble`` wraps a ``tsu_ensemble`` handle: one stage launch per RK stage
covers every tile of every member, each member keeps its own dt and time,
and steps replay from a captured HIP graph.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import check, lib, ptr


class Ensemble:
    def __init__(self, scenarios: Sequence, precision: str = "fp32", device: int = 0,
                 tile_rows: Sequence[int] = None, graph_chunk: int = 0):
        L = lib()
        self.n = len(scenarios)
        self.shapes = [(sc.npy, sc.npx) for sc in scenarios]
        P = (_lib.SimParams * self.n)(*[_lib.make_params(sc.params) for sc in scenarios])
        Sc = (_lib.Scheme * self.n)(*[_lib.make_scheme(sc.scheme) for sc in scenarios])
        self.precision = {"fp64": _lib.TSU_FP64, "fp32": _lib.TSU_FP32}[precision]
        h = C.c_void_p()
        check(L.tsu_ensemble_create(C.byref(h), self.n, P, Sc, self.precision, device), "tsu_ensemble_create")
        self._h = h
        for i, sc in enumerate(scenarios):
            V = np.ascontiguousarray(sc.V, dtype=np.float64)
            fix = np.ascontiguousarray(sc.fixH, dtype=np.int32)
            check(L.tsu_ensemble_upload(self._h, i, ptr(V), ptr(fix, C.c_int)), "tsu_ensemble_upload")
        if tile_rows is not None:
            self.set_tile_rows(tile_rows)
        if graph_chunk:
            self.set_graph_chunk(graph_chunk)
        check(L.tsu_ensemble_init_dt(self._h), "tsu_ensemble_init_dt")

    def close(self):
        if getattr(self, "_h", None):
            lib().tsu_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile_rows(self, rows: Sequence[int]):
        r = np.ascontiguousarray(rows, dtype=np.int32)
        assert r.size == self.n
        check(lib().tsu_ensemble_set_tile_rows(self._h, ptr(r, C.c_int)), "tsu_ensemble_set_tile_rows")

    def set_graph_chunk(self, chunk: int):
        check(lib().tsu_ensemble_set_graph_chunk(self._h, int(chunk)), "tsu_ensemble_set_graph_chunk")

    def set_stop_step(self, n: int):
        check(lib().tsu_ensemble_set_stop_step(self._h, int(n)), "tsu_ensemble_set_stop_step")

    def step(self, n: int = 1, sync: bool = True):
        check(lib().tsu_ensemble_step(self._h, int(n)), "tsu_ensemble_step")
        if sync:
            self.synchronize()

    def synchronize(self):
        check(lib().tsu_ensemble_synchronize(self._h), "tsu_ensemble_synchronize")

    def state(self, i: int) -> np.ndarray:
        ny, nx = self.shapes[i]
        out = np.empty((5, ny, nx), dtype=np.float64)
        check(lib().tsu_ensemble_download(self._h, int(i), ptr(out)), "tsu_ensemble_download")
        return out

    def time(self, i: int):
        t, dt, n = C.c_double(), C.c_double(), C.c_longlong()
        check(lib().tsu_ensemble_get_time(self._h, int(i), C.byref(t), C.byref(dt), C.byref(n)),
              "tsu_ensemble_get_time")
        return t.value, dt.value, n.value

    def mass(self, i: int):
        out = np.zeros(2)
        check(lib().tsu_ensemble_mass(self._h, int(i), ptr(out)), "tsu_ensemble_mass")
        return float(out[0]), float(out[1])

    def snapshot(self, restore: bool = False):
        check(lib().tsu_ensemble_snapshot(self._h, int(bool(restore))), "tsu_ensemble_snapshot")

    def time_steps(self, nsteps: int) -> float:
        ms = C.c_double()
        check(lib().tsu_ensemble_time_steps(self._h, int(nsteps), C.byref(ms)), "tsu_ensemble_time_steps")
        return ms.value
