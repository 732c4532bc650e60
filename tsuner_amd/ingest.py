"""Real-data input: the reference's Read_files = 1 loader
(Problem_loader.cxx:355-624) through libtsuner's tsu_load_input_files.

A bathymetry grid and an NCTR free-surface/velocity series in
longitude/latitude (NetCDF classic or NetCDF-4) become a Cartesian Scenario:
initial state, fixed-bed flags and the boundary series CC/tCC, with inflow on
all four sides while t < tCC[T-2] (VolumeSolver.cu:172-215).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from . import scenarios as S


def load_input_files(bathy_file: str, eta_file: str, params: dict, fixed_bed_file: Optional[str] = None,
                     sea_level_correction: float = 0.0) -> dict:
    """Runs tsu_load_input_files on params['npx'] x params['npy'] cells.
    Returns h, H (npy, npx), fixH int32, CC (T, 2*npx+2*npy, 4), tCC (T,)
    and params updated with the Cartesian extent, dx, dy and nvol."""
    L = _lib.lib()
    T = C.c_ulong(0)
    _lib.check(L.tsu_input_time_levels(eta_file.encode(), C.byref(T)), "tsu_input_time_levels")
    npx, npy = int(params["npx"]), int(params["npy"])
    n = npx * npy
    h = np.zeros(n)
    q = np.zeros(2 * n)
    H = np.zeros(n)
    fix = np.zeros(n, np.int32)
    CC = np.zeros((T.value, 2 * npx + 2 * npy, 4))
    tCC = np.zeros(T.value)
    sp = _lib.make_params(params)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    code = L.tsu_load_input_files(bathy_file.encode(), eta_file.encode(),
                                  (fixed_bed_file or "0").encode(), float(sea_level_correction), C.byref(sp),
                                  dp(h), q.ctypes.data_as(C.c_void_p), dp(H),
                                  fix.ctypes.data_as(C.POINTER(C.c_int)), dp(CC), dp(tCC), T.value)
    _lib.check(code, "tsu_load_input_files")
    p = dict(params)
    for k in ("xmin", "xmax", "ymin", "ymax", "dx", "dy", "nvol"):
        p[k] = getattr(sp, k)
    return dict(h=h.reshape(npy, npx), q=q.reshape(npy, npx, 2), H=H.reshape(npy, npx),
                fixH=fix.reshape(npy, npx), CC=CC, tCC=tCC, params=p)


def scenario_from_files(bathy_file: str, eta_file: str, params: dict, fixed_bed_file: Optional[str] = None,
                        sea_level_correction: float = 0.0, name: str = "input_files", **scheme_kw) -> S.Scenario:
    """A Scenario in the reference's Read_files = 1 configuration: the loaded
    state, inflow series on all sides, the default scheme otherwise."""
    d = load_input_files(bathy_file, eta_file, params, fixed_bed_file, sea_level_correction)
    V = S.pack_state(d["h"], d["q"][..., 0], d["q"][..., 1], d["H"])
    scheme = S.default_scheme(bc=[S.BC_INFLOW] * 4)
    scheme.update(scheme_kw)
    return S.Scenario(name=name, params=d["params"], scheme=scheme, V=V, fixH=d["fixH"], CC=d["CC"],
                      tCC=d["tCC"])
