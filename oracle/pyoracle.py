"""ctypes binding of the CPU oracle (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker / timed CPU baseline; the product path never imports this file.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TSU_ORACLE_LIB names another build of the same sources (the mutation
# self-test of tests/test_oracle_mutations.py loads deliberately broken
# copies through it, to show that the pins can tell).
LIB_PATH = os.environ.get("TSU_ORACLE_LIB") or os.path.join(HERE, "libtsu_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcSimParams(C.Structure):
    """Layout of the reference SimParams (TSUNER.h:10-30)."""
    _fields_ = [("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
                ("total_time", C.c_double), ("save_time", C.c_double),
                ("npx", C.c_int), ("npy", C.c_int), ("nvol", C.c_int),
                ("CFL", C.c_double), ("Q", C.c_double), ("R", C.c_double), ("porosity", C.c_double),
                ("DS", C.c_double), ("FR", C.c_double), ("Opar", C.c_double), ("OC", C.c_double),
                ("Dsg", C.c_double), ("HQB", C.c_double), ("vmax", C.c_double), ("epsilon_h", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double)]


class OrcScheme(C.Structure):
    _fields_ = [("order", C.c_int), ("bedload", C.c_int), ("suspension", C.c_int),
                ("bc", C.c_int * 4), ("limiter", C.c_int), ("bedload_law", C.c_int),
                ("friction_law", C.c_int), ("dt_narrow", C.c_int), ("extrema", C.c_int),
                ("grass_ag", C.c_double), ("manning_n", C.c_double)]


def make_params(p: dict) -> OrcSimParams:
    sp = OrcSimParams()
    for name, _ in OrcSimParams._fields_:
        setattr(sp, name, p[name])
    return sp


def make_scheme(s: dict) -> OrcScheme:
    sc = OrcScheme()
    for name, _ in OrcScheme._fields_:
        if name == "bc":
            for i in range(4):
                sc.bc[i] = int(s["bc"][i])
        else:
            setattr(sc, name, s[name])
    return sc


_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcSimParams), C.POINTER(OrcScheme), _dp, _ip, _dp, _dp, C.c_long]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_stage.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_state.argtypes = [C.c_void_p, _dp]
        L.orc_get_state0.argtypes = [C.c_void_p, _dp]
        L.orc_get_extrema.argtypes = [C.c_void_p, _dp, _dp, _dp]
        for fn in ("orc_time", "orc_dt"):
            getattr(L, fn).restype = C.c_double
            getattr(L, fn).argtypes = [C.c_void_p]
        L.orc_steps.restype = C.c_long
        L.orc_steps.argtypes = [C.c_void_p]
        L.orc_set_time.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_limiter.restype = C.c_double
        L.orc_limiter.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_velocity.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.orc_state_filter.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_bedload_flux.restype = C.c_double
        L.orc_bedload_flux.argtypes = [C.POINTER(OrcSimParams), C.POINTER(OrcScheme), C.c_double, C.c_double, C.c_double]
        L.orc_edge_flux.argtypes = [C.POINTER(OrcSimParams), C.POINTER(OrcScheme), _dp, _dp, C.c_int, C.c_int,
                                    C.c_int, _dp, _dp]
        L.orc_cell_dt.restype = C.c_double
        L.orc_cell_dt.argtypes = [C.POINTER(OrcSimParams), C.c_double, C.c_double, C.c_double]
        L.orc_min_reduction.restype = C.c_double
        L.orc_min_reduction.argtypes = [_dp, C.c_long, C.c_int]
        L.orc_erosion_deposition.argtypes = [C.POINTER(OrcSimParams), C.POINTER(OrcScheme), C.c_double, C.c_double,
                                             C.c_double, C.c_double, _dp]
        L.orc_derive_params.argtypes = [C.POINTER(OrcSimParams)]
        L.orc_cell_faces.restype = C.c_double
        L.orc_cell_faces.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_shared_quotient_mismatches.restype = C.c_long
        L.orc_shared_quotient_mismatches.argtypes = [C.c_long, C.c_ulonglong]
        L.orc_shared_quotient_tiny_mismatches.restype = C.c_long
        L.orc_shared_quotient_tiny_mismatches.argtypes = [C.c_long, C.c_ulonglong, C.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


class OracleSim:
    """Reference-structured FP64 CPU solver (the checker)."""

    def __init__(self, params: dict, scheme: dict, V: np.ndarray, fixH: np.ndarray,
                 CC: np.ndarray = None, tCC: np.ndarray = None):
        L = lib()
        self._p = make_params(params)
        self._s = make_scheme(scheme)
        self.npx, self.npy = int(params["npx"]), int(params["npy"])
        self.nvol = self.npx * self.npy
        V = np.ascontiguousarray(V, dtype=np.float64)
        assert V.size == 5 * self.nvol
        fix = np.ascontiguousarray(fixH, dtype=np.int32)
        if CC is not None:
            self._CC = np.ascontiguousarray(CC, dtype=np.float64)
            self._tCC = np.ascontiguousarray(tCC, dtype=np.float64)
            ccp, tp, T = _ptr(self._CC), _ptr(self._tCC), len(self._tCC)
        else:
            ccp, tp, T = None, None, 0
        self._h = L.orc_create(C.byref(self._p), C.byref(self._s), _ptr(V), _ptr(fix, C.c_int), ccp, tp, T)
        if not self._h:
            raise MemoryError("oracle allocation failed")

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        lib().orc_step(self._h, int(n))

    def stage(self, k: int):
        lib().orc_stage(self._h, int(k))

    def state(self) -> np.ndarray:
        out = np.empty((5, self.npy, self.npx), dtype=np.float64)
        lib().orc_get_state(self._h, _ptr(out))
        return out

    def state0(self) -> np.ndarray:
        out = np.empty((5, self.npy, self.npx), dtype=np.float64)
        lib().orc_get_state0(self._h, _ptr(out))
        return out

    def extrema(self):
        a = [np.empty((self.npy, self.npx), dtype=np.float64) for _ in range(3)]
        lib().orc_get_extrema(self._h, *[_ptr(x) for x in a])
        return a

    @property
    def t(self) -> float:
        return lib().orc_time(self._h)

    @property
    def dt(self) -> float:
        return lib().orc_dt(self._h)

    @property
    def steps(self) -> int:
        return lib().orc_steps(self._h)

    def set_time(self, t: float, dt: float):
        lib().orc_set_time(self._h, t, dt)


def from_scenario(sc) -> OracleSim:
    return OracleSim(sc.params, sc.scheme, sc.V, sc.fixH, sc.CC, sc.tCC)


# point functions ---------------------------------------------------------

def limiter(law: int, a: float, b: float) -> float:
    return lib().orc_limiter(law, a, b)


def velocity(h, qx, qy):
    ux, uy = C.c_double(), C.c_double()
    lib().orc_velocity(h, qx, qy, C.byref(ux), C.byref(uy))
    return ux.value, uy.value


def state_filter(h, qx, qy, hc):
    a = [C.c_double(v) for v in (h, qx, qy, hc)]
    lib().orc_state_filter(*[C.byref(x) for x in a])
    return tuple(x.value for x in a)


def edge_flux(params: dict, scheme: dict, left, right, normal, fixed: int):
    p, s = make_params(params), make_scheme(scheme)
    l = np.ascontiguousarray(left, dtype=np.float64)
    r = np.ascontiguousarray(right, dtype=np.float64)
    dm, dp = np.zeros(5), np.zeros(5)
    lib().orc_edge_flux(C.byref(p), C.byref(s), _ptr(l), _ptr(r), int(normal[0]), int(normal[1]), int(fixed),
                        _ptr(dm), _ptr(dp))
    return dm, dp


def bedload_flux(params: dict, scheme: dict, h, un, ut) -> float:
    return lib().orc_bedload_flux(C.byref(make_params(params)), C.byref(make_scheme(scheme)), h, un, ut)


def cell_dt(params: dict, h, qx, qy) -> float:
    return lib().orc_cell_dt(C.byref(make_params(params)), h, qx, qy)


def min_reduction(dtv: np.ndarray, narrow: int = 1) -> float:
    a = np.ascontiguousarray(dtv, dtype=np.float64)
    return lib().orc_min_reduction(_ptr(a), a.size, narrow)


def cell_faces(limiter: int, minus, cell, plus):
    """Filtered faces (Fm, Fp) of the middle cell along one axis and its
    limited surface slope deta; each argument is (h, q_axis, q_other, H, hc)."""
    c3 = np.ascontiguousarray(np.concatenate([minus, cell, plus]), dtype=np.float64)
    fm, fp = np.zeros(5), np.zeros(5)
    deta = lib().orc_cell_faces(int(limiter), _ptr(c3), _ptr(fm), _ptr(fp))
    return fm, fp, deta


def erosion_rate(params: dict, scheme: dict, h, qx, qy, hc) -> float:
    out = C.c_double()
    lib().orc_erosion_deposition(C.byref(make_params(params)), C.byref(make_scheme(scheme)), h, qx, qy, hc,
                                 C.byref(out))
    return out.value
