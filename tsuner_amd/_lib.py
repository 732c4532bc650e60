"""ctypes binding of libtsuner.so (the C ABI declared in include/tsuner.h).

There is no fallback: if the HIP library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TSUNER_AMD_LIB selects another build of the library (e.g. an occupancy
# variant from build.build_variant); default: the in-tree build.
LIB_PATH = os.environ.get("TSUNER_AMD_LIB") or os.path.join(HERE, "libtsuner.so")

TSU_FP64, TSU_FP32 = 0, 1
STATUS = {0: "ok", 1: "device memory", 2: "host", 3: "argument", 4: "hip", 5: "unsupported"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class SimParams(C.Structure):
    """Field order and types of the reference SimParams (TSUNER.h:10-30)."""
    _fields_ = [("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
                ("total_time", C.c_double), ("save_time", C.c_double),
                ("npx", C.c_int), ("npy", C.c_int), ("nvol", C.c_int),
                ("CFL", C.c_double), ("Q", C.c_double), ("R", C.c_double), ("porosity", C.c_double),
                ("DS", C.c_double), ("FR", C.c_double), ("Opar", C.c_double), ("OC", C.c_double),
                ("Dsg", C.c_double), ("HQB", C.c_double), ("vmax", C.c_double), ("epsilon_h", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double)]


class Scheme(C.Structure):
    _fields_ = [("order", C.c_int), ("bedload", C.c_int), ("suspension", C.c_int),
                ("bc", C.c_int * 4), ("limiter", C.c_int), ("bedload_law", C.c_int),
                ("friction_law", C.c_int), ("dt_narrow", C.c_int), ("extrema", C.c_int),
                ("grass_ag", C.c_double), ("manning_n", C.c_double)]


class Double2(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double)]


def make_params(p: dict) -> SimParams:
    sp = SimParams()
    for name, _ in SimParams._fields_:
        setattr(sp, name, p[name])
    return sp


def make_scheme(s: dict) -> Scheme:
    sc = Scheme()
    for name, _ in Scheme._fields_:
        if name == "bc":
            for i in range(4):
                sc.bc[i] = int(s["bc"][i])
        else:
            setattr(sc, name, s[name])
    return sc


class TsuError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__("%s failed (%s): %s" % (where, STATUS.get(code, code), msg))
        self.code = code


_lib = None

# name: (restype, argtypes)
_PROTOS = {
    "TSUNER": (C.c_int, [_dp, C.POINTER(Double2), _dp, _ip, SimParams, C.c_char_p, _dp, _dp, _dp, C.c_ulong,
                         _dp, C.POINTER(C.c_float), _dp, C.c_int, _dp, C.POINTER(C.c_float), C.c_int]),
    "checkCUDAsupport": (C.c_int, [C.c_int]),
    "tsu_device_check": (C.c_int, [C.c_int]),
    "initializeGPU": (None, []),
    "tsu_run": (C.c_int, [_dp, _ip, C.POINTER(SimParams), C.POINTER(Scheme), C.c_int, C.c_char_p, _dp, _dp,
                          C.c_ulong, _dp, C.c_int]),
    "tsu_state_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(SimParams), C.POINTER(Scheme), C.c_int,
                                   C.c_int]),
    "tsu_state_destroy": (C.c_int, [C.c_void_p]),
    "tsu_state_upload": (C.c_int, [C.c_void_p, _dp, _ip]),
    "tsu_state_download": (C.c_int, [C.c_void_p, _dp]),
    "tsu_state_set_inflow": (C.c_int, [C.c_void_p, _dp, _dp, C.c_ulong]),
    "tsu_init_dt": (C.c_int, [C.c_void_p]),
    "tsu_step": (C.c_int, [C.c_void_p, C.c_int]),
    "tsu_synchronize": (C.c_int, [C.c_void_p]),
    "tsu_get_time": (C.c_int, [C.c_void_p, _dp, _dp, C.POINTER(C.c_longlong)]),
    "tsu_set_time": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "tsu_set_stop_time": (C.c_int, [C.c_void_p, C.c_double]),
    "tsu_mass": (C.c_int, [C.c_void_p, _dp]),
    "tsu_get_extrema": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "tsu_set_tile_rows": (C.c_int, [C.c_void_p, C.c_int]),
    "tsu_get_tile_rows": (C.c_int, [C.c_void_p]),
    "tsu_set_graph_chunk": (C.c_int, [C.c_void_p, C.c_int]),
    "tsu_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "tsu_get_timing": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_longlong), C.c_int]),
    "tsu_stream": (C.c_void_p, [C.c_void_p]),
    "tsu_device_state": (C.c_void_p, [C.c_void_p]),
    "tsu_ensemble_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(SimParams), C.POINTER(Scheme),
                                      C.c_int, C.c_int]),
    "tsu_ensemble_destroy": (C.c_int, [C.c_void_p]),
    "tsu_ensemble_upload": (C.c_int, [C.c_void_p, C.c_int, _dp, _ip]),
    "tsu_ensemble_download": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "tsu_ensemble_init_dt": (C.c_int, [C.c_void_p]),
    "tsu_ensemble_set_inflow": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, C.c_ulong]),
    "tsu_ensemble_set_stop_step": (C.c_int, [C.c_void_p, C.c_longlong]),
    "tsu_ensemble_step": (C.c_int, [C.c_void_p, C.c_int]),
    "tsu_ensemble_synchronize": (C.c_int, [C.c_void_p]),
    "tsu_ensemble_get_time": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, C.POINTER(C.c_longlong)]),
    "tsu_ensemble_mass": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "tsu_ensemble_set_tile_rows": (C.c_int, [C.c_void_p, _ip]),
    "tsu_ensemble_set_graph_chunk": (C.c_int, [C.c_void_p, C.c_int]),
    "tsu_ensemble_snapshot": (C.c_int, [C.c_void_p, C.c_int]),
    "tsu_ensemble_time_steps": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "tsu_set_clamp": (C.c_int, [C.c_void_p, C.c_int]),
    "tsu_ensemble_set_stop_time": (C.c_int, [C.c_void_p, _dp, C.c_int]),
    "tsu_ensemble_get_extrema": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "tsu_ensemble_stream": (C.c_void_p, [C.c_void_p]),
    "tsu_ensemble_get_times": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_longlong)]),
    "tsu_nn_conv": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                              C.c_int, C.c_void_p]),
    "tsu_nn_wgrad_workspace": (C.c_longlong, [C.c_int]),
    "tsu_nn_wgrad": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                               C.c_int, C.c_void_p, C.c_longlong, C.c_void_p]),
    "tsu_nn_wgrad_noalloc": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_void_p]),
    "tsu_nn_mse": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_longlong, C.c_float,
                             C.c_void_p]),
    "tsu_nn_adam": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_float, C.c_float,
                              C.c_float, C.c_float, C.c_void_p]),
    "tsu_nn_pack": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "tsu_ensemble_nn_features": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_float, C.c_void_p]),
    "tsu_ensemble_nn_apply": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_float, C.c_void_p, C.c_longlong,
                                        C.c_void_p, C.c_void_p, C.c_float]),
    "tsu_ensemble_restrict": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tsu_ensemble_nn_target": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.c_void_p]),
    "tsu_bandwidth_probe": (C.c_int, [C.c_int, C.c_longlong, C.c_int, _dp]),
    "tsu_parse_data_dat": (C.c_int, [C.c_char_p, C.POINTER(SimParams), C.c_char_p, C.c_int, C.c_char_p, C.c_int,
                                     _dp]),
    "tsu_input_time_levels": (C.c_int, [C.c_char_p, C.POINTER(C.c_ulong)]),
    "tsu_load_input_files": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.POINTER(SimParams), _dp,
                                       C.c_void_p, _dp, _ip, _dp, _dp, C.c_ulong]),
    "tsu_last_error": (C.c_char_p, []),
    "tsu_version": (C.c_char_p, []),
    "tsu_kat_edge_flux": (C.c_int, [C.POINTER(SimParams), C.POINTER(Scheme), C.c_int, _dp, _dp, _ip, _ip, _dp,
                                    _dp, C.c_int]),
    "tsu_kat_edge_flux_pairs": (C.c_int, [C.POINTER(SimParams), C.POINTER(Scheme), C.c_int, _dp, _dp, _ip, _ip,
                                          _dp, _dp, C.c_int]),
    "tsu_kat_point": (C.c_int, [C.POINTER(SimParams), C.POINTER(Scheme), C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                _dp, _dp, _dp, _dp, C.c_int]),
    "tsu_write_output": (C.c_int, [C.c_char_p, C.POINTER(SimParams), C.c_int, _dp, _dp, _dp, _dp, _dp]),
}


def lib():
    """Load libtsuner.so; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtsuner.so is not built (run __graft_entry__.build() or "
                               "python tsuner_amd/build.py); there is no CPU fallback")
        # torch ships its own HIP runtime under the same soname
        # (libamdhip64.so.7).  Loaded first, it serves libtsuner too; loaded
        # after libtsuner it would bind to the system runtime and find no GPU.
        # So when torch is installed it is imported before the library.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int, where: str):
    if code != 0:
        raise TsuError(code, where, lib().tsu_last_error().decode(errors="replace"))


def ptr(a: np.ndarray, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))
