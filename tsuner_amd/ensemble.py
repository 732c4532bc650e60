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
            if getattr(sc, "CC", None) is not None:
                cc = np.ascontiguousarray(sc.CC, dtype=np.float64)
                tc = np.ascontiguousarray(sc.tCC, dtype=np.float64)
                check(L.tsu_ensemble_set_inflow(self._h, i, ptr(cc), ptr(tc), len(tc)), "tsu_ensemble_set_inflow")
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

    def set_stop_time(self, t, clamp: bool = False):
        """Per-member stop times (a scalar applies to all); clamp shortens the
        crossing step to end exactly on the stop time."""
        arr = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (self.n,)))
        check(lib().tsu_ensemble_set_stop_time(self._h, ptr(arr), int(bool(clamp))), "tsu_ensemble_set_stop_time")

    def times(self) -> np.ndarray:
        t = np.zeros(self.n)
        check(lib().tsu_ensemble_get_times(self._h, ptr(t), None), "tsu_ensemble_get_times")
        return t

    def steps_done(self) -> np.ndarray:
        s = np.zeros(self.n, dtype=np.int64)
        check(lib().tsu_ensemble_get_times(self._h, None, s.ctypes.data_as(C.POINTER(C.c_longlong))),
              "tsu_ensemble_get_times")
        return s

    def extrema(self, i: int):
        ny, nx = self.shapes[i]
        a = [np.empty((ny, nx)) for _ in range(3)]
        check(lib().tsu_ensemble_get_extrema(self._h, int(i), ptr(a[0]), ptr(a[1]), ptr(a[2])),
              "tsu_ensemble_get_extrema")
        return a[0], a[1], a[2]

    @property
    def stream(self) -> int:
        return lib().tsu_ensemble_stream(self._h) or 0

    # ---- learned-correction hooks (device buffers given as raw addresses) --
    def nn_features(self, scale_in, wet_h: float, x_addr: int):
        s = (C.c_float * 8)(*[float(v) for v in scale_in])
        check(lib().tsu_ensemble_nn_features(self._h, s, C.c_float(wet_h), C.c_void_p(x_addr)),
              "tsu_ensemble_nn_features")

    def nn_apply(self, scale_out, wet_h: float, out_addr: int, expect_step: int, resid_addr: int = 0,
                 x_addr: int = 0, shore_h: float = 0.0):
        s = (C.c_float * 4)(*[float(v) for v in scale_out])
        check(lib().tsu_ensemble_nn_apply(self._h, s, C.c_float(wet_h), C.c_void_p(out_addr), int(expect_step),
                                          C.c_void_p(resid_addr or None), C.c_void_p(x_addr or None),
                                          C.c_float(shore_h)), "tsu_ensemble_nn_apply")

    def restrict_from(self, fine: "Ensemble"):
        check(lib().tsu_ensemble_restrict(fine._h, self._h), "tsu_ensemble_restrict")

    def nn_target(self, fine: "Ensemble", scale_out, target_addr: int):
        s = (C.c_float * 4)(*[float(v) for v in scale_out])
        check(lib().tsu_ensemble_nn_target(fine._h, self._h, s, C.c_void_p(target_addr)), "tsu_ensemble_nn_target")

    def time_steps(self, nsteps: int) -> float:
        ms = C.c_double()
        check(lib().tsu_ensemble_time_steps(self._h, int(nsteps), C.byref(ms)), "tsu_ensemble_time_steps")
        return ms.value
