"""Real-data input (Problem_loader.cxx:355-624, interpolations.cxx): the
library's tsu_load_input_files against the oracle restatement (bit-exact),
the oracle pinned by closed forms, and the reference's own example files
(when the reference tree is present)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import pyingest as OI
from tsuner_amd import _lib, ingest
from tsuner_amd import scenarios as S

from ingest_files import (coastal_case, write_bathy, write_bathy_classic, write_classic, write_series,
                          write_series_classic, write_series_hdf5)

EXAMPLE = "/root/reference/example"


def params(npx, npy):
    return S.base_params(npx, npy, 1.0, 1.0)


def compare(lib_out, orc):
    for k in ("h", "H", "CC", "tCC"):
        np.testing.assert_array_equal(lib_out[k], orc[k], err_msg=k)
    np.testing.assert_array_equal(lib_out["fixH"], orc["fixH"])
    for k in ("xmin", "xmax", "ymin", "ymax", "dx", "dy"):
        assert lib_out["params"][k] == orc[k], k


# ------------------------------------------------------------- oracle pins ----

def test_series_interpolation_reproduces_biquadratics():
    """interpolate_CC's 3x3 Lagrange stencil is exact on span{1,x,x^2} x
    span{1,y,y^2}; an asymmetric polynomial catches a transposed axis, a
    swapped node or a weight paired with the wrong neighbour."""
    nx, ny = 30, 20
    x0, y0, dxg, dyg = 10.0, -4.0, 0.5, 0.25
    xc = x0 + 0.5 * dxg + np.arange(nx) * dxg
    yc = y0 + 0.5 * dyg + np.arange(ny) * dyg
    X, Y = np.meshgrid(xc, yc)
    P = lambda x, y: 1.5 + 0.7 * x - 0.3 * y + 0.11 * x * x - 0.05 * x * y + 0.21 * y * y + 0.013 * x * x * y * y
    f = P(X, Y)[None].ravel()
    rng = np.random.default_rng(1)
    x = rng.uniform(xc[1], xc[-2], 200)
    y = rng.uniform(yc[1], yc[-2], 200)
    vals = OI._series_values(f, f, f, nx, ny, 0, x, y, x0, y0, dxg, dyg)
    np.testing.assert_allclose(vals[0], P(x, y), rtol=1e-12)


def test_global_basis_resampling_closed_forms():
    """interpolate_spherics weights the 3x3 stencil by L(xt), L(yt) at the
    target's GLOBAL coordinate: a ramp f = i gives ig + xt, f = j gives
    jg + yt, a constant stays constant, and the domain centre takes the
    centre cell alone."""
    nxf, nyf, npx, npy = 40, 30, 23, 17
    xmin, xmax, ymin, ymax = 2.0, 6.0, -1.0, 2.0
    dxf, dyf = (xmax - xmin) / nxf, (ymax - ymin) / nyf
    I, J = np.meshgrid(np.arange(nxf, dtype=float), np.arange(nyf, dtype=float))
    ri = OI.interpolate_spherics(xmin, xmax, ymin, ymax, xmin, ymin, nxf, nyf, dxf, dyf, npx, npy, I)
    rj = OI.interpolate_spherics(xmin, xmax, ymin, ymax, xmin, ymin, nxf, nyf, dxf, dyf, npx, npy, J)
    dx, dy = (xmax - xmin) / npx, (ymax - ymin) / npy
    x = xmin + 0.5 * dx + np.arange(npx) * dx
    y = ymin + 0.5 * dy + np.arange(npy) * dy
    ig = np.clip(np.trunc((x - xmin - 0.5 * dxf) / dxf), 0, nxf - 1)
    jg = np.clip(np.trunc((y - ymin - 0.5 * dyf) / dyf), 0, nyf - 1)
    xt = (2 * x - (xmax + xmin)) / (xmax - xmin)
    yt = (2 * y - (ymax + ymin)) / (ymax - ymin)
    inner_x = (ig > 0) & (ig < nxf - 1)
    inner_y = (jg > 0) & (jg < nyf - 1)
    inner = inner_y[:, None] & inner_x[None, :]   # the collapse rule alters edge stencils on both axes
    assert inner.sum() > 200
    np.testing.assert_allclose(ri[inner], np.broadcast_to(ig + xt, ri.shape)[inner], atol=1e-12)
    np.testing.assert_allclose(rj[inner], np.broadcast_to((jg + yt)[:, None], rj.shape)[inner], atol=1e-12)
    c = OI.interpolate_spherics(xmin, xmax, ymin, ymax, xmin, ymin, nxf, nyf, dxf, dyf, npx, npy,
                                np.full((nyf, nxf), 7.25))
    np.testing.assert_allclose(c, 7.25, rtol=4e-16)
    F = np.random.default_rng(2).normal(size=(nyf, nxf))
    m = OI.interpolate_spherics(xmin, xmax, ymin, ymax, xmin, ymin, nxf, nyf, dxf, dyf, 21, 15, F)
    assert m[7, 10] == F[int((ymin + 7.5 * (ymax - ymin) / 15 - ymin - 0.5 * dyf) / dyf),
                         int((xmin + 10.5 * (xmax - xmin) / 21 - xmin - 0.5 * dxf) / dxf)]


def test_degree_to_metre_closed_forms():
    """interpolations.cxx:79-93: 2*pi*R/360 metres per degree of latitude,
    cos(latitude) of it per degree of longitude; the Cartesian width is the
    widest (southernmost) row."""
    assert abs(OI.dy_transform() - 111194.92664455873) < 1e-6
    dxE, dmax = OI.dx_transform(46.0, 0.5, 4, 236.0, 235.0, 0.01)
    for i in range(4):
        lat = 46.0 + 0.25 + 0.5 * i
        assert abs(dxE[i] - 111194.92664455873 * math.cos(math.radians(lat)) * 0.01) < 1e-9
    assert abs(dmax - 111194.92664455873 * math.cos(math.radians(46.25))) < 1e-6


def test_series_negative_index_takes_last_column():
    """The upper clamp compares the int index with an unsigned size, so a
    point west of the forcing grid reads its LAST column
    (interpolations.cxx:467-470); f = column index makes it visible."""
    nx, ny = 12, 10
    x0, y0, d = 5.0, 0.0, 1.0
    I, _ = np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float))
    f = I.ravel()
    x = np.array([2.0])   # (2 - 5 - 0.5) / 1 -> -3
    y = np.array([4.3])
    val = OI._series_values(f, f, f, nx, ny, 0, x, y, x0, y0, d, d)[0][0]
    x1 = x0 + 0.5 * d + (nx - 1) * d
    w = OI.lag_nodes(2.0, x1 - d, x1, x1 + d)
    expect = w[0] * (nx - 2) + w[1] * (nx - 1) + w[2] * (nx - 1)
    assert abs(val - expect) < 1e-9 * abs(expect)


def test_fill_values_and_units(tmp_path):
    """Forcing values beyond +-1e20 become the boundary depth (h slot) or 0
    (u, v); a filled initial surface becomes the depth in cm units before the
    0.01 scaling (interpolations.cxx:560-565, 867-877); h = eta*0.01 + H +
    shift clipped at 0 (Problem_loader.cxx:586-588), eta having lost the
    shift once already in its own units."""
    lon = 235.0 + 0.01 * np.arange(8)
    lat = 46.0 + 0.01 * np.arange(6)
    bathy = np.full((6, 8), 12.0, np.float32)
    write_bathy(tmp_path / "b.nc", lon, lat, bathy)
    flon = 234.9 + 0.02 * np.arange(16)
    flat = 45.9 + 0.02 * np.arange(14)
    T = 3
    h = np.full((T, 14, 16), 50.0, np.float32)
    u = np.full((T, 14, 16), 20.0, np.float32)
    v = np.full((T, 14, 16), -10.0, np.float32)
    h[1] = 1e35
    u[2] = -1e35
    write_series(str(tmp_path / "s_"), flon, flat, [10, 20, 30], h, u, v)
    slr = 0.25
    o = OI.load_problem_files(str(tmp_path / "b.nc"), str(tmp_path / "s_h.nc"), "0", slr, 8, 6)
    CC = o["CC"]
    np.testing.assert_allclose(CC[0, :, 0], 0.5 + slr, rtol=1e-14)
    np.testing.assert_allclose(CC[0, :, 1], 0.2, rtol=1e-14)
    np.testing.assert_allclose(CC[0, :, 2], -0.1, rtol=1e-14)
    np.testing.assert_allclose(CC[:, :, 3], 12.0, rtol=1e-14)
    np.testing.assert_array_equal(CC[1, :, 0], CC[1, :, 3])
    np.testing.assert_array_equal(CC[2, :, 1], 0.0)
    # the shift is taken off eta while still in cm (interpolations.cxx:871),
    # then added in metres (Problem_loader.cxx:586)
    np.testing.assert_allclose(o["h"], (50.0 - slr) * 0.01 + 12.0 + slr, rtol=1e-14)
    h[0] = 1e35
    write_series(str(tmp_path / "s_"), flon, flat, [10, 20, 30], h, u, v)
    o = OI.load_problem_files(str(tmp_path / "b.nc"), str(tmp_path / "s_h.nc"), "0", slr, 8, 6)
    np.testing.assert_allclose(o["h"], 12.0 * 0.01 + 12.0 + slr, rtol=1e-14)
    lib_out = ingest.load_input_files(str(tmp_path / "b.nc"), str(tmp_path / "s_h.nc"), params(8, 6),
                                      sea_level_correction=slr)
    compare(lib_out, o)


# ------------------------------------------------------ library vs oracle ----

@pytest.mark.parametrize("npx,npy,upper,fill", [(64, 48, True, False), (37, 53, False, True), (96, 96, True, True)])
def test_library_matches_oracle_on_synthetic_files(tmp_path, npx, npy, upper, fill):
    b, e = coastal_case(tmp_path, fill=fill, upper=upper, seed=npx)
    lib_out = ingest.load_input_files(b, e, params(npx, npy), sea_level_correction=0.3)
    orc = OI.load_problem_files(b, e, "0", 0.3, npx, npy)
    compare(lib_out, orc)
    assert np.all(lib_out["h"] >= 0) and np.all(np.isfinite(lib_out["CC"]))


def _hdf5_available():
    try:
        OI._H5.lib()
        return True
    except Exception:
        return False


@pytest.mark.skipif(not _hdf5_available(), reason="libhdf5 not loadable")
def test_library_matches_oracle_on_hdf5_series(tmp_path):
    """The NetCDF-4 (HDF5) path of the reader, on files written here: the
    same series as NetCDF classic and as HDF5 load to identical arrays."""
    b, e = coastal_case(tmp_path, seed=11, T=5)
    from scipy.io import netcdf_file
    with netcdf_file(e, "r", mmap=False) as f:
        lon, lat, t = (np.array(f.variables[k].data) for k in ("LON", "LAT", "TIME"))
        h = np.array(f.variables["HA"].data)
    with netcdf_file(e.replace("h.nc", "u.nc"), "r", mmap=False) as f:
        u = np.array(f.variables["UA"].data)
    with netcdf_file(e.replace("h.nc", "v.nc"), "r", mmap=False) as f:
        v = np.array(f.variables["VA"].data)
    paths = write_series_hdf5(str(tmp_path / "h5_"), lon, lat, t, h, u, v)
    with open(paths[0], "rb") as f:
        assert f.read(4) == b"\x89HDF"
    classic = ingest.load_input_files(b, e, params(70, 55), sea_level_correction=0.05)
    h5 = ingest.load_input_files(b, paths[0], params(70, 55), sea_level_correction=0.05)
    compare(h5, OI.load_problem_files(b, paths[0], "0", 0.05, 70, 55))
    for k in ("h", "H", "CC", "tCC"):
        np.testing.assert_array_equal(h5[k], classic[k], err_msg=k)


def test_classic_writer_matches_scipy(tmp_path):
    """The test suite's own classic-format writer (the only one here that
    writes CDF-5) gives scipy's bytes for versions 1 and 2: header layout,
    variable order, padding and record interleaving are those of an
    independent writer, and version 5 changes only the magic byte and the
    field widths."""
    rng = np.random.default_rng(4)
    lon, lat = np.linspace(235.8, 235.9, 7), np.linspace(46.8, 46.9, 5)
    t = np.array([30.0, 60.0, 90.0])
    h = rng.normal(size=(3, 5, 7)).astype(np.float32)
    for version in (1, 2):
        for dt in ("f4", "f8", "i1", "i2", "i4"):
            data = (rng.normal(size=(5, 7)) * 50).astype(dt)
            write_bathy(tmp_path / "a.nc", lon, lat, data, var_dtype=dt, coord_names=("x", "y"), version=version)
            write_bathy_classic(tmp_path / "b.nc", lon, lat, data, var_dtype=dt, coord_names=("x", "y"),
                                version=version)
            assert (tmp_path / "a.nc").read_bytes() == (tmp_path / "b.nc").read_bytes(), (version, dt)
        for upper in (True, False):
            p1 = write_series(tmp_path / "s1_", lon, lat, t, h, h + 1, h + 2, upper=upper, version=version)
            p2 = write_series_classic(tmp_path / "s2_", lon, lat, t, h, h + 1, h + 2, upper=upper, version=version)
            for x, y in zip(p1, p2):
                assert open(x, "rb").read() == open(y, "rb").read(), (version, upper, x)


def _read_classic(path, names):
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as f:
        return [np.array(f.variables[k].data) for k in names]


@pytest.mark.parametrize("upper", [True, False])
def test_cdf5_files_load_like_classic(tmp_path, upper):
    """The 64-bit data format (CDF-5: 64-bit counts, lengths and ids, record
    variables, global attributes of 64-bit integers) reads to the same arrays
    as the classic files it copies."""
    b, e = coastal_case(tmp_path, seed=21, T=5, upper=upper)
    lon, lat, bathy = _read_classic(b, ("lon", "lat", "bathy"))
    b5 = str(tmp_path / "b5.nc")
    write_bathy_classic(b5, lon, lat, bathy)
    k = ("LON", "LAT", "TIME") if upper else ("lon", "lat", "time")
    v = ("HA", "UA", "VA") if upper else ("ha_lc", "ua_lc", "va_lc")
    slon, slat, t = _read_classic(e, k)
    series = [_read_classic(e.replace("h.nc", c + ".nc"), (n,))[0] for c, n in zip("huv", v)]
    paths = write_series_classic(str(tmp_path / "c5_"), slon, slat, t, *series, upper=upper,
                                 gatts={"title": "CDF-5 copy", "levels": np.arange(3, dtype=np.int64)})
    for p in (b5, paths[0]):
        with open(p, "rb") as f:
            assert f.read(4) == b"CDF\x05"
    L = _lib.lib()
    T = C.c_ulong(0)
    assert L.tsu_input_time_levels(paths[0].encode(), C.byref(T)) == 0 and T.value == 5
    ref = ingest.load_input_files(b, e, params(60, 45), sea_level_correction=0.1)
    cdf5 = ingest.load_input_files(b5, paths[0], params(60, 45), sea_level_correction=0.1)
    for key in ("h", "H", "CC", "tCC", "fixH"):
        np.testing.assert_array_equal(cdf5[key], ref[key], err_msg=key)


@pytest.mark.parametrize("dt5,dt1", [("i8", "f8"), ("u2", "f8"), ("u4", "f8"), ("u8", "f8")])
def test_cdf5_integer_types(tmp_path, dt5, dt1):
    """The integer types that only version 5 has: a bathymetry of whole
    metres stored as such loads like the same values as doubles in a
    version-1 file, and a fixed-bed map of unsigned bytes like short
    integers."""
    b, e = coastal_case(tmp_path, seed=8, T=3)
    lon, lat, bathy = _read_classic(b, ("lon", "lat", "bathy"))
    whole = np.round(bathy + 30.0).astype(np.float64)   # non-negative whole metres, for the unsigned types
    write_bathy(tmp_path / "b1.nc", lon, lat, whole, var_dtype=dt1)
    write_bathy_classic(tmp_path / "b5.nc", lon, lat, whole.astype(dt5))
    mlon = 235.78 + 0.004 * np.arange(40)
    mlat = 46.84 + 0.003 * np.arange(35)
    mask = np.ones((35, 40), np.int16)
    mask[8:22, 6:31] = 0
    write_bathy(tmp_path / "f1.nc", mlon, mlat, mask, var_dtype="i2", coord_names=("x", "y"), version=2)
    write_bathy_classic(tmp_path / "f5.nc", mlon, mlat, mask.astype(np.uint8), var_dtype="u1",
                        coord_names=("x", "y"))
    one = ingest.load_input_files(str(tmp_path / "b1.nc"), e, params(48, 36),
                                  fixed_bed_file=str(tmp_path / "f1.nc"))
    five = ingest.load_input_files(str(tmp_path / "b5.nc"), e, params(48, 36),
                                   fixed_bed_file=str(tmp_path / "f5.nc"))
    for key in ("h", "H", "CC", "tCC", "fixH"):
        np.testing.assert_array_equal(five[key], one[key], err_msg=key)
    assert 0 < (five["fixH"] == 0).sum() < five["fixH"].size


def test_version5_types_rejected_in_classic_versions(tmp_path):
    """Type codes 7-11 exist only in version 5: a version-1 or -2 header that
    uses one is malformed."""
    L = _lib.lib()
    T = C.c_ulong(0)
    for version in (1, 2):
        path = tmp_path / ("bad%d_h.nc" % version)
        write_classic(path, version, [("TIME", None)], [("TIME", ("TIME",), "i8", np.arange(3, dtype=np.int64), {})])
        assert L.tsu_input_time_levels(str(path).encode(), C.byref(T)) == 2
        assert b"malformed" in L.tsu_last_error()


def test_longer_velocity_series_use_their_first_levels(tmp_path):
    """The level count comes from the h file; u, v and the times (read from
    the u file) contribute their first T levels (interpolations.cxx:405-447)."""
    b, e = coastal_case(tmp_path, seed=5, T=4)
    from scipy.io import netcdf_file
    ref = ingest.load_input_files(b, e, params(40, 30))
    for c, var in (("u", "UA"), ("v", "VA")):
        path = e.replace("h.nc", c + ".nc")
        with netcdf_file(path, "r", mmap=False) as f:
            lon, lat, t = (np.array(f.variables[k].data) for k in ("LON", "LAT", "TIME"))
            data = np.array(f.variables[var].data)
        extra = np.concatenate([data, 100.0 + data[:2]])
        with netcdf_file(path, "w") as f:
            f.createDimension("TIME", None)
            f.createDimension("LAT", len(lat))
            f.createDimension("LON", len(lon))
            for k, arr, dims in (("LON", lon, ("LON",)), ("LAT", lat, ("LAT",))):
                f.createVariable(k, "f4", dims)[:] = arr
            f.createVariable(var, "f4", ("TIME", "LAT", "LON"))[:] = extra.astype(np.float32)
            f.createVariable("TIME", "f4", ("TIME",))[:] = np.concatenate([t, [999.0, 1999.0]])
    longer = ingest.load_input_files(b, e, params(40, 30))
    compare(longer, OI.load_problem_files(b, e, "0", 0.0, 40, 30))
    for k in ("h", "H", "CC", "tCC"):
        np.testing.assert_array_equal(longer[k], ref[k], err_msg=k)


def test_fixed_bed_map_as_short_integers(tmp_path):
    b, e = coastal_case(tmp_path, seed=3)
    lon = 235.78 + 0.004 * np.arange(40)
    lat = 46.84 + 0.003 * np.arange(35)
    mask = np.ones((35, 40), np.int16)
    mask[10:20, 5:30] = 0
    write_bathy(tmp_path / "fix.nc", lon, lat, mask, var_dtype="i2", coord_names=("x", "y"), version=2)
    lib_out = ingest.load_input_files(b, e, params(50, 40), fixed_bed_file=str(tmp_path / "fix.nc"))
    orc = OI.load_problem_files(b, e, str(tmp_path / "fix.nc"), 0.0, 50, 40)
    compare(lib_out, orc)
    assert 0 < (lib_out["fixH"] == 0).sum() < lib_out["fixH"].size


def test_input_errors(tmp_path):
    """Return codes of include/tsuner.h: 2 (TSU_ERR_HOST) unreadable input,
    3 (TSU_ERR_ARG) buffers sized for another number of time levels."""
    L = _lib.lib()
    T = C.c_ulong(0)
    assert L.tsu_input_time_levels(str(tmp_path / "missing_h.nc").encode(), C.byref(T)) == 2
    assert b"missing_h.nc" in L.tsu_last_error()
    (tmp_path / "junk_h.nc").write_bytes(b"not a grid file at all")
    assert L.tsu_input_time_levels(str(tmp_path / "junk_h.nc").encode(), C.byref(T)) == 2
    assert b"neither" in L.tsu_last_error()
    b, e = coastal_case(tmp_path, T=4)
    assert L.tsu_input_time_levels(e.encode(), C.byref(T)) == 0 and T.value == 4
    n = 10 * 8
    buf = [np.zeros(k) for k in (n, 2 * n, n)]
    fix = np.zeros(n, np.int32)
    CC = np.zeros(3 * (2 * 10 + 2 * 8) * 4)
    tCC = np.zeros(3)
    sp = _lib.make_params(params(10, 8))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    code = L.tsu_load_input_files(b.encode(), e.encode(), b"0", 0.0, C.byref(sp), dp(buf[0]),
                                  buf[1].ctypes.data_as(C.c_void_p), dp(buf[2]),
                                  fix.ctypes.data_as(C.POINTER(C.c_int)), dp(CC), dp(tCC), 3)
    assert code == 3 and b"time levels" in L.tsu_last_error()
    os.remove(e.replace("h.nc", "v.nc"))
    with pytest.raises(RuntimeError, match="v.nc"):
        ingest.load_input_files(b, e, params(10, 8))


# ----------------------------------------------------- the example files ----

@pytest.mark.skipif(not os.path.isdir(EXAMPLE), reason="reference example files not present")
@pytest.mark.parametrize("npx,npy", [(100, 100), (64, 90)])
def test_reference_example_files(npx, npy):
    """The reference's data.dat inputs: NetCDF-3 bathymetry (2701^2, 0.33
    arcsec) and NetCDF-4/HDF5 NCTR series (10 levels of 721 x 961)."""
    b = os.path.join(EXAMPLE, "westport_bathy_example.nc")
    e = os.path.join(EXAMPLE, "new_westport_runupB_h.nc")
    lib_out = ingest.load_input_files(b, e, params(npx, npy))
    orc = OI.load_problem_files(b, e, "0", 0.0, npx, npy)
    compare(lib_out, orc)
    # domain: 0.25 x 0.25 degrees (+ half cells) at 46.82-47.07 N
    span = 0.25 + 0.25 / 2700
    assert abs(lib_out["params"]["ymax"] - OI.dy_transform() * span) < 1e-6
    lat_s = 46.82 - 0.5 * 0.25 / 2700 + 0.5 * span / npy
    assert abs(lib_out["params"]["xmax"] - OI.dy_transform() * math.cos(math.radians(lat_s)) * span) < 1e-3
    np.testing.assert_array_equal(lib_out["tCC"], 30.0 * np.arange(1, 11))
    assert lib_out["H"].min() > -117 and lib_out["H"].max() < 50


@pytest.mark.skipif(not os.path.isdir(EXAMPLE), reason="reference example files not present")
def test_westport_fixture_matches_the_library():
    """tests/golden/westport_100x100.npz (written by make_westport.py with the
    oracle) is what the library loads from the example files today."""
    from ingest_files import WESTPORT, westport_scenario
    z = np.load(WESTPORT)
    lib_out = ingest.load_input_files(os.path.join(EXAMPLE, "westport_bathy_example.nc"),
                                      os.path.join(EXAMPLE, "new_westport_runupB_h.nc"), params(100, 100))
    for k in ("h", "H", "fixH", "CC", "tCC"):
        np.testing.assert_array_equal(lib_out[k], z[k], err_msg=k)
    sc = westport_scenario()
    assert sc.params["dx"] == lib_out["params"]["dx"] and sc.params["FR"] == 2.5e-2 / 8
