"""CPU restatement of the reference's real-data input path (test
infrastructure only: the checker for tsu_load_input_files).

Restates, in numpy float64 with the reference's order of operations:
  Problem_loader.cxx:397-610   the Read_files = 1 branch of LoadProblemData
  Problem_loader.cxx:57-175    the grid readers (dimension / variable names)
  interpolations.cxx:46-95     Lagrange bases, degree -> metre transforms
  interpolations.cxx:368-774   interpolate_CC (boundary series)
  interpolations.cxx:779-877   interpolate_spherics, fix_eta
  interpolations.cxx:879-1114  spheric2cartesian, spheric2cartesian_CC

Files are read with scipy.io.netcdf_file (NetCDF classic) and, for NetCDF-4
files, with the HDF5 C library through ctypes (HDF5 1.10, the third-party
library the reference reaches through libnetcdf's nc_open).

Parity pins (tests/test_ingest.py): biquadratic reproduction of the
boundary-series interpolation, the closed form ic + xt of the global-basis
resampling on a linear ramp, partition of unity, the degree -> metre closed
forms, fill-value and sign rules, and agreement with the library on the
reference's own example files.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

PI = 3.14159265358979323846   # Sim_parameters.hxx:30
RE = 6371e3                   # Sim_parameters.hxx:37
UNITS = 0.01                  # Sim_parameters.hxx:42


# ------------------------------------------------------------------ readers

class _H5:
    """Minimal ctypes binding of the HDF5 1.10 C API (hid_t = int64)."""
    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            for name in (os.environ.get("TSU_HDF5_LIB"), "libhdf5.so", "/opt/conda/lib/libhdf5.so"):
                if not name:
                    continue
                try:
                    L = C.CDLL(name)
                    break
                except OSError:
                    continue
            else:
                raise RuntimeError("libhdf5 not loadable")
            i64, u64p = C.c_int64, C.POINTER(C.c_ulonglong)
            L.H5open.restype = C.c_int
            L.H5Fopen.restype, L.H5Fopen.argtypes = i64, [C.c_char_p, C.c_uint, i64]
            L.H5Fclose.argtypes = [i64]
            L.H5Lexists.restype, L.H5Lexists.argtypes = C.c_int, [i64, C.c_char_p, i64]
            L.H5Dopen2.restype, L.H5Dopen2.argtypes = i64, [i64, C.c_char_p, i64]
            L.H5Dclose.argtypes = [i64]
            L.H5Dget_space.restype, L.H5Dget_space.argtypes = i64, [i64]
            L.H5Sget_simple_extent_ndims.argtypes = [i64]
            L.H5Sget_simple_extent_dims.argtypes = [i64, u64p, u64p]
            L.H5Sclose.argtypes = [i64]
            L.H5Dread.argtypes = [i64, i64, i64, i64, i64, C.c_void_p]
            L.H5Eset_auto2.argtypes = [i64, C.c_void_p, C.c_void_p]
            L.H5open()
            L.H5Eset_auto2(0, None, None)
            cls.native_double = C.c_int64.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
            cls._lib = L
        return cls._lib


class Grid:
    """nc_open-like access: names of dimensions and variables, whole reads as
    float64 (Problem_loader.cxx:57-175)."""

    def __init__(self, path: str):
        with open(path, "rb") as f:
            magic = f.read(8)
        self.path = path
        self.hdf5 = magic == b"\x89HDF\r\n\x1a\n"
        if self.hdf5:
            self.L = _H5.lib()
            self.fid = self.L.H5Fopen(path.encode(), 0, 0)
            if self.fid < 0:
                raise OSError("HDF5 cannot open " + path)
        else:
            from scipy.io import netcdf_file
            self.nc = netcdf_file(path, "r", mmap=False)

    def close(self):
        if self.hdf5:
            self.L.H5Fclose(self.fid)
        else:
            self.nc.close()

    def has(self, name: str) -> bool:
        if self.hdf5:
            return self.L.H5Lexists(self.fid, name.encode(), 0) > 0
        return name in self.nc.variables

    def first(self, names):
        for n in names:
            if self.has(n):
                return n
        raise KeyError("%s: none of %s" % (self.path, names))

    def read(self, name: str) -> np.ndarray:
        if not self.hdf5:
            return np.array(self.nc.variables[name].data, dtype=np.float64)
        L = self.L
        d = L.H5Dopen2(self.fid, name.encode(), 0)
        s = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(s)
        dims = (C.c_ulonglong * max(nd, 1))()
        L.H5Sget_simple_extent_dims(s, dims, None)
        out = np.empty(tuple(dims[:nd]), np.float64)
        rc = L.H5Dread(d, _H5.native_double, 0, 0, 0, out.ctypes.data)
        L.H5Sclose(s)
        L.H5Dclose(d)
        if rc < 0:
            raise OSError("HDF5 read of %s failed" % name)
        return out


# ------------------------------------------------------------------ bases

def lag_global(s):
    """interpolations.cxx:46-54: L0, L1, L2 on the nodes -1, 0, 1."""
    return s / 2.0 * (s - 1), 1 - s * s, s / 2.0 * (s + 1)


def lag_nodes(x, x0, x1, x2):
    """interpolations.cxx:67-75."""
    return (((x - x1) * (x - x2)) / ((x0 - x1) * (x0 - x2)),
            ((x - x0) * (x - x2)) / ((x1 - x0) * (x1 - x2)),
            ((x - x0) * (x - x1)) / ((x2 - x0) * (x2 - x1)))


def _trunc(v):
    return np.trunc(v).astype(np.int64)


def _collapse(nx, ny, ig, jg):
    """Neighbour positions of interpolations.cxx:818-850 (and 921-952)."""
    pos = jg * nx + ig
    posL, posR, posU, posD = pos - 1, pos + 1, pos + nx, pos - nx
    posLU, posLD, posRU, posRD = posU - 1, posD - 1, posU + 1, posD + 1
    e = ig == 0
    posL, posLU, posLD = np.where(e, pos, posL), np.where(e, pos, posLU), np.where(e, pos, posLD)
    e = ig == nx - 1
    posR, posRD, posRU = np.where(e, pos, posR), np.where(e, pos, posRD), np.where(e, pos, posRU)
    e = jg == 0
    posD, posLD, posRD = np.where(e, pos, posD), np.where(e, pos, posLD), np.where(e, pos, posRD)
    e = jg == ny - 1
    posU, posLU, posRU = np.where(e, pos, posU), np.where(e, pos, posLU), np.where(e, pos, posRU)
    return posLD, posD, posRD, posL, pos, posR, posLU, posU, posRU


def _sum9(fun, P, xt, yt):
    posLD, posD, posRD, posL, pos, posR, posLU, posU, posRU = P
    L0x, L1x, L2x = lag_global(xt)
    L0y, L1y, L2y = lag_global(yt)
    N1, N2, N3 = L0x * L0y, L1x * L0y, L2x * L0y
    N4, N5, N6 = L0x * L1y, L1x * L1y, L2x * L1y
    N7, N8, N9 = L0x * L2y, L1x * L2y, L2x * L2y
    return (fun[posLD] * N1 + fun[posD] * N2 + fun[posRD] * N3 + fun[posL] * N4 + fun[pos] * N5
            + fun[posR] * N6 + fun[posLU] * N7 + fun[posU] * N8 + fun[posRU] * N9)


def interpolate_spherics(xmin, xmax, ymin, ymax, xmin_fun, ymin_fun, npx_fun, npy_fun, dx_fun, dy_fun, npx, npy,
                         fun):
    """interpolations.cxx:779-865."""
    fun = np.asarray(fun, np.float64).ravel()
    dx = (xmax - xmin) / (npx)
    dy = (ymax - ymin) / (npy)
    i = np.arange(npx)[None, :]
    j = np.arange(npy)[:, None]
    x = xmin + 0.5 * dx + i * dx
    y = ymin + 0.5 * dy + j * dy
    xt = 1.0 / (xmax - xmin) * (2 * x - (xmax + xmin))
    yt = 1.0 / (ymax - ymin) * (2 * y - (ymax + ymin))
    i_grid = np.clip(_trunc((x - xmin_fun - 0.5 * dx_fun) / dx_fun), 0, npx_fun - 1)
    j_grid = np.clip(_trunc((y - ymin_fun - 0.5 * dy_fun) / dy_fun), 0, npy_fun - 1)
    i_grid, j_grid = np.broadcast_arrays(i_grid, j_grid)
    xt, yt = np.broadcast_arrays(xt, yt)
    return _sum9(fun, _collapse(npx_fun, npy_fun, i_grid, j_grid), xt, yt).reshape(npy, npx)


def fix_eta(p, H, SLR):
    """interpolations.cxx:867-877."""
    p = p - SLR
    p = np.where(p < -1e20, H, p)
    return np.where(p > 1e20, H, p)


def _series_index(v, n):
    """int(v), then `if (i > n - 1)` against an unsigned long n: a negative
    index compares as huge and becomes n - 1 (interpolations.cxx:467-470)."""
    k = _trunc(v)
    return np.where((k < 0) | (k > n - 1), n - 1, k)


def _series_values(fh, fhu, fhv, nx, ny, t, x, y, xmin_grid, ymin_grid, dx_grid, dy_grid):
    """One boundary of interpolations.cxx:456-541 at time level t, points (x, y)."""
    i_grid = _series_index((x - xmin_grid - 0.5 * dx_grid) / dx_grid, nx)
    j_grid = _series_index((y - ymin_grid - 0.5 * dy_grid) / dy_grid, ny)
    pos = t * nx * ny + nx * j_grid + i_grid
    posL = pos - 1 * (i_grid > 0)
    posR = pos + 1 * (i_grid < nx - 1)
    posU = pos + nx * (j_grid < ny - 1)
    posD = pos - nx * (j_grid > 0)
    posLD = posD - 1 * (i_grid > 0)
    posRD = posD + 1 * (i_grid < nx - 1)
    posLU = posU - 1 * (i_grid > 0)
    posRU = posU + 1 * (i_grid < nx - 1)
    x1 = xmin_grid + 0.5 * dx_grid + i_grid * dx_grid
    x0 = x1 - dx_grid
    x2 = x1 + dx_grid
    y1 = ymin_grid + 0.5 * dy_grid + j_grid * dy_grid
    y0 = y1 - dy_grid
    y2 = y1 + dy_grid
    Lx0, Lx1, Lx2 = lag_nodes(x, x0, x1, x2)
    Ly0, Ly1, Ly2 = lag_nodes(y, y0, y1, y2)
    N0, N1, N2 = Lx0 * Ly0, Lx1 * Ly0, Lx2 * Ly0
    N3, N4, N5 = Lx0 * Ly1, Lx1 * Ly1, Lx2 * Ly1
    N6, N7, N8 = Lx0 * Ly2, Lx1 * Ly2, Lx2 * Ly2
    out = []
    for f in (fh, fhu, fhv):
        out.append(N0 * f[posLD] + N1 * f[posD] + N2 * f[posRD] + N3 * f[posL] + N4 * f[pos] + N5 * f[posR]
                   + N6 * f[posLU] + N7 * f[posU] + N8 * f[posRU])
    return out


def _finish(vals, Hb, SLR):
    h, u, v = vals
    h = h * UNITS
    u = u * UNITS
    v = v * UNITS
    h = h + SLR
    h = np.where(h < -1e20, Hb, h)
    u = np.where(u < -1e20, 0.0, u)
    v = np.where(v < -1e20, 0.0, v)
    h = np.where(h > 1e20, Hb, h)
    u = np.where(u > 1e20, 0.0, u)
    v = np.where(v > 1e20, 0.0, v)
    return np.stack([h, u, v, np.broadcast_to(Hb, h.shape)], axis=-1)


def interpolate_CC(fh, fhu, fhv, T, npx, npy, xmin, xmax, ymin, ymax, xmin_grid, ymin_grid, dx_grid, dy_grid, H,
                   SLR):
    """interpolations.cxx:368-774; fh/fhu/fhv (T, ny, nx); H (npy, npx).
    Returns CC (T, 2*npx+2*npy, 4)."""
    _, ny, nx = fh.shape
    fh, fhu, fhv = (np.asarray(a, np.float64).ravel() for a in (fh, fhu, fhv))
    Hf = np.asarray(H, np.float64).ravel()
    dx = (xmax - xmin) / npx
    dy = (ymax - ymin) / npy
    CC = np.zeros((T, 2 * npx + 2 * npy, 4))
    i = np.arange(npx)
    j = np.arange(npy)
    x = xmin + 0.5 * dx + i * dx
    y = ymin + 0.5 * dy + j * dy
    g = (xmin_grid, ymin_grid, dx_grid, dy_grid)
    for t in range(T):
        ymn = np.full(npx, float(ymin))
        ymx = np.full(npx, float(ymax))
        xmn = np.full(npy, float(xmin))
        xmx = np.full(npy, float(xmax))
        CC[t, i] = _finish(_series_values(fh, fhu, fhv, nx, ny, t, x, ymn, *g), Hf[i], SLR)
        CC[t, npx + npy + i] = _finish(_series_values(fh, fhu, fhv, nx, ny, t, x, ymx, *g),
                                       Hf[(npy - 1) * npx + i], SLR)
        CC[t, 2 * npx + npy + j] = _finish(_series_values(fh, fhu, fhv, nx, ny, t, xmn, y, *g), Hf[j * npx], SLR)
        CC[t, npx + j] = _finish(_series_values(fh, fhu, fhv, nx, ny, t, xmx, y, *g), Hf[j * npx + npx - 1], SLR)
    return CC


def dy_transform():
    """interpolations.cxx:79-81: metres per degree of latitude."""
    return (2 * PI * RE) / 360


def dx_transform(ymin, dy, tam, xmax, xmin, dx_cartesianas):
    """interpolations.cxx:83-93: per-row metres per cell and the widest row."""
    dx = np.empty(tam)
    dmax = -1.0
    for i in range(tam):
        lat = ymin + 0.5 * dy + i * dy
        dx[i] = (2 * PI * RE * math.cos(lat * PI / 180)) / 360
        dmax = max(dmax, dx[i] * (xmax - xmin))
        dx[i] *= dx_cartesianas
    return dx, dmax


def spheric2cartesian(xmin, xmax, ymin, ymax, dx, dy, npx, npy, fun):
    """interpolations.cxx:879-968; dx per row (metres per cell), dy scalar."""
    fun = np.asarray(fun, np.float64).ravel()
    dxG = (xmax - xmin) / npx
    dyG = (ymax - ymin) / npy
    i = np.arange(npx)[None, :]
    j = np.arange(npy)[:, None]
    x = xmin + 0.5 * dxG + i * dxG
    y = ymin + 0.5 * dyG + j * dyG
    xt = 1.0 / (xmax - xmin) * (2 * x - (xmax + xmin))
    yt = 1.0 / (ymax - ymin) * (2 * y - (ymax + ymin))
    j_grid = np.clip(_trunc((y - ymin - 0.5 * dy) / dy), 0, npy - 1)
    dxr = dx[j_grid]
    xminE = 0 + (xmax - dxr * npx) / 2
    i_grid = np.clip(_trunc((x - xminE - 0.5 * dxr) / dxr), 0, npx - 1)
    i_grid, j_grid = np.broadcast_arrays(i_grid, j_grid)
    xt, yt = np.broadcast_arrays(xt, yt)
    return _sum9(fun, _collapse(npx, npy, i_grid, j_grid), xt, yt).reshape(npy, npx)


def spheric2cartesian_CC(xmin, xmax, ymin, ymax, dxE, dyE, npx, npy, CC):
    """interpolations.cxx:970-1114; CC (T, 2*npx+2*npy, 4), returned resampled."""
    T = CC.shape[0]
    out = np.empty_like(CC)
    dx = (xmax - xmin) / npx
    dy = (ymax - ymin) / npy
    i = np.arange(npx)
    j = np.arange(npy)
    x = xmin + 0.5 * dx + i * dx
    y = ymin + 0.5 * dy + j * dy
    xt = 1.0 / (xmax - xmin) * (2 * x - (xmax + xmin))
    yt = 1.0 / (ymax - ymin) * (2 * y - (ymax + ymin))
    N0x, N1x, N2x = lag_global(xt)
    N0y, N1y, N2y = lag_global(yt)
    ig_b = np.clip(_trunc((x - (0 + (xmax - dxE[0] * npx) / 2) - 0.5 * dxE[0]) / dxE[0]), 0, npx - 1)
    ig_t = np.clip(_trunc((x - (0 + (xmax - dxE[npy - 1] * npx) / 2) - 0.5 * dxE[npy - 1]) / dxE[npy - 1]), 0,
                   npx - 1)
    jg = np.clip(_trunc((y - (0 + (ymax - dyE * npy) / 2) - 0.5 * dyE) / dyE), 0, npy - 1)
    for t in range(T):
        c = CC[t]
        # bottom, top: slot = side offset + index
        for off, ig in ((0, ig_b), (npx + npy, ig_t)):
            s = off + ig
            s0 = s - 1 * (ig > 0)
            s1 = s + 1 * (ig < npx - 1)
            out[t, off + i] = N0x[:, None] * c[s0] + N1x[:, None] * c[s] + N2x[:, None] * c[s1]
        # left: upper neighbour while j_grid < npy - 1; right: while j_grid < npx - 1
        for off, lim in ((2 * npx + npy, npy), (npx, npx)):
            s = off + jg
            s0 = s - 1 * (jg > 0)
            s1 = s + 1 * (jg < lim - 1)
            out[t, off + j] = N0y[:, None] * c[s0] + N1y[:, None] * c[s] + N2y[:, None] * c[s1]
    return out


# ------------------------------------------------------------------ loader

def _coords(g: Grid):
    return g.read(g.first(["lon", "x", "LON"])), g.read(g.first(["lat", "y", "LAT"]))


def load_problem_files(bathy_file, eta_file, fixed_file, SLR, npx, npy):
    """Problem_loader.cxx:397-610 (SPL = 0, no tides).  Returns dict with
    h, H (npy, npx), fixH int32, CC (T, 2*npx+2*npy, 4), tCC, and the
    Cartesian xmin, xmax, ymin, ymax, dx, dy."""
    g = Grid(bathy_file)
    xp, yp = _coords(g)
    bathy = g.read("bathy")
    g.close()
    nxg, nyg = len(xp), len(yp)
    dx_grid = (xp[nxg - 1] - xp[0]) / (nxg - 1)
    dy_grid = (yp[nyg - 1] - yp[0]) / (nyg - 1)
    xmin = xp[0] - 0.5 * dx_grid
    xmax = xp[nxg - 1] + 0.5 * dx_grid
    ymin = yp[0] - 0.5 * dy_grid
    ymax = yp[nyg - 1] + 0.5 * dy_grid
    bathy_i = interpolate_spherics(xmin, xmax, ymin, ymax, xmin, ymin, nxg, nyg, dx_grid, dy_grid, npx, npy, bathy)

    read_fix = fixed_file not in (None, "", "0")
    if read_fix:
        g = Grid(fixed_file)
        fxp, fyp = _coords(g)
        fx = g.read("bathy")
        g.close()
        fix_i = interpolate_spherics(xmin, xmax, ymin, ymax, fxp[0], fyp[0], len(fxp), len(fyp),
                                     (fxp[-1] - fxp[0]) / (len(fxp) - 1), (fyp[-1] - fyp[0]) / (len(fyp) - 1),
                                     npx, npy, fx)

    g = Grid(eta_file)
    xci, yci = _coords(g)
    fh = g.read(g.first(["ha_lc", "HA", "ha"]))
    g.close()
    T = fh.shape[0]
    nxc, nyc = len(xci), len(yci)
    dx_CI = (xci[nxc - 1] - xci[0]) / nxc
    dy_CI = (yci[nyc - 1] - yci[0]) / nyc
    eta_i = interpolate_spherics(xmin, xmax, ymin, ymax, xci[0], yci[0], nxc, nyc, dx_CI, dy_CI, npx, npy, fh[0])
    eta_i = fix_eta(eta_i, bathy_i, SLR)

    stem = list(eta_file)
    stem[-4] = "u"
    g = Grid("".join(stem))
    fhu = g.read(g.first(["ua_lc", "UA", "ua"]))
    tCC = g.read(g.first(["time", "TIME"]))[:T]
    g.close()
    stem[-4] = "v"
    g = Grid("".join(stem))
    fhv = g.read(g.first(["va_lc", "VA", "va"]))
    g.close()
    CC = interpolate_CC(fh, fhu, fhv, T, npx, npy, xmin, xmax, ymin, ymax, xci[0], yci[0], dx_CI, dy_CI, bathy_i,
                        SLR)

    dx = (xmax - xmin) / (npx)
    dy = (ymax - ymin) / npy
    dy_unidad = dy_transform()
    dxE, dmax = dx_transform(ymin, dy, npy, xmax, xmin, dx)
    xmin = 0
    xmax = dmax
    ymax = dy_unidad * (ymax - ymin)
    ymin = 0
    dyE = dy * dy_unidad
    H = spheric2cartesian(xmin, xmax, ymin, ymax, dxE, dyE, npx, npy, bathy_i)
    fix = spheric2cartesian(xmin, xmax, ymin, ymax, dxE, dyE, npx, npy, fix_i) if read_fix else None
    eta = spheric2cartesian(xmin, xmax, ymin, ymax, dxE, dyE, npx, npy, eta_i)
    CC = spheric2cartesian_CC(xmin, xmax, ymin, ymax, dxE, dyE, npx, npy, CC)
    dx = (xmax - xmin) / (npx)
    dy = (ymax - ymin) / npy
    h = eta * UNITS + H + SLR
    h = h * (h >= 0.0)
    fixH = (fix >= 0.5).astype(np.int32) if read_fix else np.ones((npy, npx), np.int32)
    return dict(h=h, H=H, fixH=fixH, CC=CC, tCC=tCC, xmin=float(xmin), xmax=float(xmax), ymin=float(ymin),
                ymax=float(ymax), dx=float(dx), dy=float(dy))
