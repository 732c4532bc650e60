"""Host-side mirror of the reference's run interface for the solver path.

``Solver`` wraps one ``tsu_state`` handle of libtsuner: the reference's
TSUNER() (GPU/main_TSUNER.cu:21-222) split into upload, initial dt, stepping
with device-resident dt and time, and download.  ``run_tsuner`` calls the
reference-compatible TSUNER() entry itself.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, lib, ptr


class Solver:
    """One simulation on one GPU.

    precision: "fp64" (reference arithmetic) or "fp32" (FP32 state planes).
    """

    def __init__(self, params: dict, scheme: dict, V: np.ndarray, fixH: np.ndarray,
                 CC: Optional[np.ndarray] = None, tCC: Optional[np.ndarray] = None,
                 precision: str = "fp64", device: int = 0, tile_rows: int = 0, init_dt: bool = True):
        L = lib()
        self.params, self.scheme = dict(params), dict(scheme)
        self.npx, self.npy = int(params["npx"]), int(params["npy"])
        self.nvol = self.npx * self.npy
        self._p = _lib.make_params(params)
        self._s = _lib.make_scheme(scheme)
        self.precision = {"fp64": _lib.TSU_FP64, "fp32": _lib.TSU_FP32}[precision]
        h = C.c_void_p()
        check(L.tsu_state_create(C.byref(h), C.byref(self._p), C.byref(self._s), self.precision, device),
              "tsu_state_create")
        self._h = h
        V = np.ascontiguousarray(V, dtype=np.float64)
        if V.size != 5 * self.nvol:
            raise ValueError("state must hold 5 planes of npx*npy cells")
        fix = np.ascontiguousarray(fixH, dtype=np.int32)
        check(L.tsu_state_upload(self._h, ptr(V), ptr(fix, C.c_int)), "tsu_state_upload")
        if CC is not None:
            cc = np.ascontiguousarray(CC, dtype=np.float64)
            tc = np.ascontiguousarray(tCC, dtype=np.float64)
            check(L.tsu_state_set_inflow(self._h, ptr(cc), ptr(tc), len(tc)), "tsu_state_set_inflow")
        if tile_rows:
            check(L.tsu_set_tile_rows(self._h, int(tile_rows)), "tsu_set_tile_rows")
        if init_dt:
            self.init_dt()

    @classmethod
    def from_scenario(cls, sc, **kw) -> "Solver":
        return cls(sc.params, sc.scheme, sc.V, sc.fixH, sc.CC, sc.tCC, **kw)

    def close(self):
        if getattr(self, "_h", None):
            lib().tsu_state_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def init_dt(self):
        check(lib().tsu_init_dt(self._h), "tsu_init_dt")

    def step(self, n: int = 1, sync: bool = True):
        check(lib().tsu_step(self._h, int(n)), "tsu_step")
        if sync:
            self.synchronize()

    def synchronize(self):
        check(lib().tsu_synchronize(self._h), "tsu_synchronize")

    def state(self) -> np.ndarray:
        out = np.empty((5, self.npy, self.npx), dtype=np.float64)
        check(lib().tsu_state_download(self._h, ptr(out)), "tsu_state_download")
        return out

    def time(self):
        t, dt, n = C.c_double(), C.c_double(), C.c_longlong()
        check(lib().tsu_get_time(self._h, C.byref(t), C.byref(dt), C.byref(n)), "tsu_get_time")
        return t.value, dt.value, n.value

    @property
    def t(self) -> float:
        return self.time()[0]

    @property
    def dt(self) -> float:
        return self.time()[1]

    @property
    def steps(self) -> int:
        return self.time()[2]

    def set_time(self, t: float, dt: float = 0.0):
        check(lib().tsu_set_time(self._h, float(t), float(dt)), "tsu_set_time")

    def set_stop_time(self, t: float, clamp: bool = None):
        """Steps after t are no-ops; clamp=True also shortens the crossing step
        to end exactly on t (None leaves the clamp setting unchanged)."""
        check(lib().tsu_set_stop_time(self._h, float(t)), "tsu_set_stop_time")
        if clamp is not None:
            check(lib().tsu_set_clamp(self._h, int(bool(clamp))), "tsu_set_clamp")

    def mass(self):
        out = np.zeros(2)
        check(lib().tsu_mass(self._h, ptr(out)), "tsu_mass")
        return float(out[0]), float(out[1])

    def extrema(self):
        a = [np.empty((self.npy, self.npx)) for _ in range(3)]
        check(lib().tsu_get_extrema(self._h, *[ptr(x) for x in a]), "tsu_get_extrema")
        return a

    @property
    def tile_rows(self) -> int:
        return lib().tsu_get_tile_rows(self._h)

    def set_tile_rows(self, rows: int):
        check(lib().tsu_set_tile_rows(self._h, int(rows)), "tsu_set_tile_rows")

    def set_graph_chunk(self, chunk: int):
        check(lib().tsu_set_graph_chunk(self._h, int(chunk)), "tsu_set_graph_chunk")

    def set_timing(self, on: bool = True):
        check(lib().tsu_set_timing(self._h, int(bool(on))), "tsu_set_timing")

    def timing(self, reset: bool = False):
        """(stage_ms[2], launches[2]) of the event-timed stage kernels."""
        ms = np.zeros(2)
        n = (C.c_longlong * 2)()
        check(lib().tsu_get_timing(self._h, ptr(ms), n, int(bool(reset))), "tsu_get_timing")
        return ms.copy(), [n[0], n[1]]

    @property
    def stream(self) -> int:
        return lib().tsu_stream(self._h)


def run_tsuner(params: dict, V: np.ndarray, fixH: np.ndarray, prefix: Optional[str],
               CC: Optional[np.ndarray] = None, tCC: Optional[np.ndarray] = None, device: int = 0):
    """Calls the reference-compatible TSUNER() entry (TSUNER.h:39) with AoS
    momentum as the reference's caller passes it; returns (code, seconds)."""
    L = lib()
    nv = int(params["npx"]) * int(params["npy"])
    V = np.ascontiguousarray(V, dtype=np.float64).reshape(5, nv)
    h = np.ascontiguousarray(V[0])
    H = np.ascontiguousarray(V[3])
    q = (_lib.Double2 * nv)()
    qa = np.frombuffer(q, dtype=np.float64).reshape(nv, 2)
    qa[:, 0] = V[1]
    qa[:, 1] = V[2]
    fix = np.ascontiguousarray(fixH, dtype=np.int32).reshape(nv)
    t = C.c_double()
    if CC is not None:
        cc = np.ascontiguousarray(CC, dtype=np.float64)
        tc = np.ascontiguousarray(tCC, dtype=np.float64)
        ccp, tcp, T = ptr(cc), ptr(tc), len(tc)
    else:
        ccp, tcp, T = None, None, 0
    code = L.TSUNER(ptr(h), q, ptr(H), ptr(fix, C.c_int), _lib.make_params(params),
                    prefix.encode() if prefix else None, None, ccp, tcp, T, None, None, C.byref(t), device,
                    None, None, 0)
    return code, t.value
