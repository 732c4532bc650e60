"""Synthetic input files of the real-data path, written with
scipy.io.netcdf_file (NetCDF classic): a bathymetry grid and an NCTR-style
<stem>h/u/v series in longitude/latitude, with the reference's variable
names (Problem_loader.cxx:57-175, interpolations.cxx:386-460)."""
from __future__ import annotations

import os

import numpy as np


def write_bathy(path, lon, lat, bathy, var_dtype="f4", coord_names=("lon", "lat"), version=1):
    from scipy.io import netcdf_file
    with netcdf_file(str(path), "w", version=version) as f:
        f.createDimension(coord_names[0], len(lon))
        f.createDimension(coord_names[1], len(lat))
        x = f.createVariable(coord_names[0], "f8", (coord_names[0],))
        y = f.createVariable(coord_names[1], "f8", (coord_names[1],))
        x[:] = lon
        y[:] = lat
        b = f.createVariable("bathy", var_dtype, (coord_names[1], coord_names[0]))
        b[:] = bathy
        b.units = "meters"
        b.positive = "down"


def write_series(stem, lon, lat, times, h, u, v, upper=True, version=1):
    """stem + 'h.nc', 'u.nc', 'v.nc'; h, u, v (T, nlat, nlon) in cm and cm/s.
    upper=True uses the NCTR names (HA, LON, LAT, TIME, unlimited time),
    False the lower-case ones (ha_lc, lon, lat, time, fixed time)."""
    from scipy.io import netcdf_file
    names = dict(lon="LON", lat="LAT", time="TIME", h="HA", u="UA", v="VA") if upper else \
        dict(lon="lon", lat="lat", time="time", h="ha_lc", u="ua_lc", v="va_lc")
    paths = []
    for c, data in (("h", h), ("u", u), ("v", v)):
        path = str(stem) + c + ".nc"
        with netcdf_file(path, "w", version=version) as f:
            f.createDimension(names["time"], None if upper else len(times))
            f.createDimension(names["lat"], len(lat))
            f.createDimension(names["lon"], len(lon))
            t = f.createVariable(names["time"], "f4", (names["time"],))
            x = f.createVariable(names["lon"], "f4", (names["lon"],))
            y = f.createVariable(names["lat"], "f4", (names["lat"],))
            x[:] = lon
            y[:] = lat
            d = f.createVariable(names[c], "f4", (names["time"], names["lat"], names["lon"]))
            d[:] = data
            t[:] = times
        paths.append(path)
    return paths


# NetCDF classic files written from the published format description, in
# any of its versions (scipy writes only 1 and 2): big-endian, magic "CDF" +
# version byte; counts, lengths, dimension ids and vsize 32-bit in versions
# 1 and 2, 64-bit in version 5 (the "64-bit data" format, CDF-5); offsets
# 32-bit in version 1, 64-bit in 2 and 5; tags and type codes 32-bit; names
# and values padded to 4 bytes; record variables interleaved record by
# record (unpadded when there is only one).
_NC_TYPES = {"i1": 1, "S1": 2, "i2": 3, "i4": 4, "f4": 5, "f8": 6,
             "u1": 7, "u2": 8, "u4": 9, "i8": 10, "u8": 11}
# default fill values, which pad variable data to 4 bytes
_NC_FILL = {"i1": -127, "S1": 0, "i2": -32767, "u1": 255, "u2": 65535}


def write_classic(path, version, dims, variables, gatts=None):
    """dims: [(name, length or None for the record dimension)];
    variables: [(name, (dim names), dtype string, array, {attr: value})];
    gatts: {name: str or array}."""
    import struct
    nfmt = ">q" if version == 5 else ">i"
    ofmt = ">i" if version == 1 else ">q"

    def i4(v): return struct.pack(">i", v)
    def nn(v): return struct.pack(nfmt, v)

    def pad(b): return b + b"\0" * (-len(b) % 4)

    def pad_data(b, dt):
        n = (-len(b) % 4) // np.dtype(dt).itemsize
        return b + np.full(n, _NC_FILL.get(dt, 0), ">" + dt).tobytes()

    def name(n):
        e = n.encode()
        return nn(len(e)) + pad(e)

    def atts(d):
        if not d:
            return i4(0) + nn(0)
        out = i4(0x0C) + nn(len(d))
        for k, v in d.items():
            if isinstance(v, str):
                raw, t, n = v.encode(), 2, len(v.encode())
            else:
                a = np.asarray(v)
                dt = a.dtype.str[1:]
                raw, t, n = a.astype(">" + dt).tobytes(), _NC_TYPES[dt], a.size
            out += name(k) + i4(t) + nn(n) + pad(raw)
        return out

    dlen = {n: l for n, l in dims}
    rec_dim = next((n for n, l in dims if l is None), None)
    is_rec = [bool(v[1]) and v[1][0] == rec_dim for v in variables]
    recs = [v for v, r in zip(variables, is_rec) if r]
    numrecs = max((len(v[3]) for v in recs), default=0)
    dim_list = (i4(0x0A) + nn(len(dims)) + b"".join(name(n) + nn(l or 0) for n, l in dims)) if dims \
        else i4(0) + nn(0)

    def slab(v, rec):   # bytes of the whole variable, or of one record
        shape = [dlen[d] for d in v[1][(1 if rec else 0):]]
        return int(np.prod(shape, dtype=np.int64)) * np.dtype(v[2]).itemsize

    def vsize(v, rec): return slab(v, rec) + (-slab(v, rec) % 4)

    def header(begins):
        out = b"CDF" + bytes([version]) + nn(numrecs) + dim_list + atts(gatts)
        out += (i4(0x0B) + nn(len(variables))) if variables else i4(0) + nn(0)
        for v, rec, beg in zip(variables, is_rec, begins):
            out += name(v[0]) + nn(len(v[1])) + b"".join(nn([n for n, _ in dims].index(d)) for d in v[1])
            out += atts(v[4]) + i4(_NC_TYPES[v[2]]) + nn(vsize(v, rec)) + struct.pack(ofmt, beg)
        return out

    off = len(header([0] * len(variables)))
    begins = []
    for v, rec in zip(variables, is_rec):
        begins.append(None if rec else off)
        off += 0 if rec else vsize(v, False)
    rec_start = off
    for i, (v, rec) in enumerate(zip(variables, is_rec)):
        if rec:
            begins[i] = off
            off += vsize(v, True)
    data = bytearray(header(begins))
    for v, rec in zip(variables, is_rec):
        if not rec:
            data += pad_data(np.ascontiguousarray(v[3]).astype(">" + v[2]).tobytes(), v[2])
    for r in range(numrecs):
        for v in recs:
            raw = np.ascontiguousarray(v[3][r]).astype(">" + v[2]).tobytes()
            data += raw if len(recs) == 1 else pad_data(raw, v[2])
    with open(path, "wb") as f:
        f.write(bytes(data))


def scipy_order(dims, variables):
    """The variables in the order scipy's writer lays them out: non-record
    variables by shape, largest first (stable), then the record variables."""
    rec = next((n for n, l in dims if l is None), None)
    dl = dict(dims)
    key = lambda v: (-1,) if v[1] and v[1][0] == rec else tuple(dl[d] for d in v[1])
    return sorted(variables, key=key, reverse=True)


def write_bathy_classic(path, lon, lat, bathy, var_dtype="f4", coord_names=("lon", "lat"), version=5):
    """write_bathy's file, here in any classic version (default CDF-5)."""
    x, y = coord_names
    dims = [(x, len(lon)), (y, len(lat))]
    write_classic(path, version, dims, scipy_order(dims, [
        (x, (x,), "f8", np.asarray(lon, np.float64), {}),
        (y, (y,), "f8", np.asarray(lat, np.float64), {}),
        ("bathy", (y, x), var_dtype, np.asarray(bathy), {"units": "meters", "positive": "down"})]))


def write_series_classic(stem, lon, lat, times, h, u, v, upper=True, version=5, gatts=None):
    """write_series' files, here in any classic version (default CDF-5)."""
    names = dict(lon="LON", lat="LAT", time="TIME", h="HA", u="UA", v="VA") if upper else \
        dict(lon="lon", lat="lat", time="time", h="ha_lc", u="ua_lc", v="va_lc")
    paths = []
    for c, data in (("h", h), ("u", u), ("v", v)):
        path = str(stem) + c + ".nc"
        T, X, Y = names["time"], names["lon"], names["lat"]
        dims = [(T, None if upper else len(times)), (Y, len(lat)), (X, len(lon))]
        write_classic(path, version, dims, scipy_order(dims, [
            (T, (T,), "f4", np.asarray(times, np.float32), {}),
            (X, (X,), "f4", np.asarray(lon, np.float32), {}),
            (Y, (Y,), "f4", np.asarray(lat, np.float32), {}),
            (names[c], (T, Y, X), "f4", np.asarray(data, np.float32), {})]), gatts=gatts)
        paths.append(path)
    return paths


def coastal_case(tmp, nlon=90, nlat=70, T=8, seed=0, fill=False, upper=True):
    """A sloping coast near 46.9 N: depth 40 m offshore (west) to 8 m of land
    (east) with seeded bumps; a free-surface series of a long wave entering
    from the west, sampled on a coarser forcing grid covering the bathymetry
    with a margin.  Returns (bathy path, h-series path)."""
    rng = np.random.default_rng(seed)
    lon = 235.80 + np.arange(nlon) * (0.10 / (nlon - 1))
    lat = 46.85 + np.arange(nlat) * (0.08 / (nlat - 1))
    X, Y = np.meshgrid((lon - lon[0]) / (lon[-1] - lon[0]), (lat - lat[0]) / (lat[-1] - lat[0]))
    bathy = 40.0 - 48.0 * X + 2.0 * np.sin(7 * Y + 3 * X) + rng.normal(0, 0.3, X.shape)
    bpath = os.path.join(str(tmp), "coast_bathy.nc")
    write_bathy(bpath, lon, lat, bathy.astype(np.float32))
    flon = 235.70 + np.arange(48) * (0.30 / 47)
    flat = 46.80 + np.arange(40) * (0.18 / 39)
    times = 30.0 * (1 + np.arange(T))
    FX, FY = np.meshgrid(flon, flat)
    k = 2 * np.pi / 0.12
    h = np.stack([60.0 * np.sin(k * (FX - 235.7) - 0.02 * t) * np.exp(-((FY - 46.9) / 0.2) ** 2) for t in times])
    u = np.stack([25.0 * np.sin(k * (FX - 235.7) - 0.02 * t) for t in times])
    v = np.stack([4.0 * np.cos(k * (FY - 46.8) - 0.02 * t) + 0 * FX for t in times])
    if fill:
        h[:, :3, :3] = 1e35
        u[:, -2:, :] = -1e35
    stem = os.path.join(str(tmp), "coast_runup_")
    write_series(stem, flon, flat, times, h.astype(np.float32), u.astype(np.float32), v.astype(np.float32),
                 upper=upper)
    return bpath, stem + "h.nc"


WESTPORT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "westport_100x100.npz")


def westport_scenario():
    """The reference's example case (data.dat with the westport files, 100 x
    100 cells) from the committed fixture: Read_files = 1 configuration,
    inflow series on all sides."""
    from tsuner_amd import scenarios as S
    z = np.load(WESTPORT)
    dd = dict(zip([str(k) for k in z["data_dat_keys"]], z["data_dat"]))
    npx, npy = int(dd["npx"]), int(dd["npy"])
    xmin, xmax, ymin, ymax, dx, dy = z["extent"]
    p = S.base_params(npx, npy, dx, dy, friction=dd["FR"], R=dd["R"], DS=dd["DS"], porosity=dd["porosity"],
                      OC=dd["OC"], Dsg=dd["Dsg"], CFL=dd["CFL"], epsilon_h=dd["epsilon_h"],
                      total_time=dd["total_time"], save_time=dd["save_time"])
    p.update(xmin=xmin, xmax=xmax, ymin=ymin, ymax=ymax, dx=dx, dy=dy)
    V = S.pack_state(z["h"], 0 * z["h"], 0 * z["h"], z["H"])
    return S.Scenario(name="westport_100x100", params=p, scheme=S.default_scheme(bc=[S.BC_INFLOW] * 4), V=V,
                      fixH=z["fixH"], CC=z["CC"], tCC=z["tCC"])


def write_series_hdf5(stem, lon, lat, times, h, u, v):
    """stem + 'h.nc', 'u.nc', 'v.nc' as HDF5 files laid out like the NCTR
    NetCDF-4 series (root datasets LON, LAT, TIME and HA / UA / VA of shape
    (T, nlat, nlon), 32-bit floats), written through the HDF5 C library."""
    import ctypes as C
    L = None
    for name in (os.environ.get("TSU_HDF5_LIB"), "libhdf5.so", "/opt/conda/lib/libhdf5.so"):
        if name:
            try:
                L = C.CDLL(name)
                break
            except OSError:
                pass
    if L is None:
        raise RuntimeError("libhdf5 not loadable")
    i64 = C.c_int64
    L.H5open()
    L.H5Fcreate.restype, L.H5Fcreate.argtypes = i64, [C.c_char_p, C.c_uint, i64, i64]
    L.H5Screate_simple.restype, L.H5Screate_simple.argtypes = i64, [C.c_int, C.POINTER(C.c_ulonglong), C.c_void_p]
    L.H5Dcreate2.restype, L.H5Dcreate2.argtypes = i64, [i64, C.c_char_p, i64, i64, i64, i64, i64]
    L.H5Dwrite.argtypes = [i64, i64, i64, i64, i64, C.c_void_p]
    for f in ("H5Dclose", "H5Sclose", "H5Fclose"):
        getattr(L, f).argtypes = [i64]
    f32 = C.c_int64.in_dll(L, "H5T_IEEE_F32LE_g").value
    nat = C.c_int64.in_dll(L, "H5T_NATIVE_FLOAT_g").value
    paths = []
    for c, data, var in (("h", h, "HA"), ("u", u, "UA"), ("v", v, "VA")):
        path = str(stem) + c + ".nc"
        fid = L.H5Fcreate(path.encode(), 2, 0, 0)   # H5F_ACC_TRUNC
        for name, arr in (("LON", lon), ("LAT", lat), ("TIME", times), (var, data)):
            a = np.ascontiguousarray(arr, np.float32)
            dims = (C.c_ulonglong * a.ndim)(*a.shape)
            sp = L.H5Screate_simple(a.ndim, dims, None)
            ds = L.H5Dcreate2(fid, name.encode(), f32, sp, 0, 0, 0)
            assert ds >= 0 and L.H5Dwrite(ds, nat, 0, 0, 0, a.ctypes.data) >= 0
            L.H5Dclose(ds)
            L.H5Sclose(sp)
        L.H5Fclose(fid)
        paths.append(path)
    return paths
