"""The C-ABI library loads on a CPU-only host and exports every entry point
that include/tsuner.h declares; host-only entries work without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tsuner_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tsuner.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
    return sorted(set(re.findall(r"TSU_API\s+[\w\s\*]*?\b(\w+)\s*\(", text)))


def test_header_declares_the_reference_entries():
    names = declared_symbols()
    for ref in ("TSUNER", "checkCUDAsupport", "initializeGPU"):
        assert ref in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_prototypes_cover_the_header():
    assert set(declared_symbols()) <= set(_lib._PROTOS)


def test_struct_layout_matches_reference_simparams():
    # 6 doubles, 3 ints (+4 bytes padding), 14 doubles (TSUNER.h:10-30)
    assert C.sizeof(_lib.SimParams) == 6 * 8 + 3 * 4 + 4 + 14 * 8
    assert _lib.SimParams.CFL.offset == 64


def test_version_and_errors_without_gpu():
    L = _lib.lib()
    assert b"gfx950" in L.tsu_version()
    assert L.tsu_state_destroy(None) == 0


def test_write_output_roundtrip(tmp_path):
    from scipy.io import netcdf_file
    from tsuner_amd import scenarios as S
    sc = S.sediment_shelf(n=20, ny=12, dx=10.0, inflow=False)
    V = sc.V.copy()
    V[4] = 0.01 * V[0]
    V[1] = 0.3 * V[0]
    recs = np.stack([V, V * 1.5])
    times = np.array([0.0, 12.5])
    hmax = V[0] + 1.0
    bmax = V[3] + 2.0
    emax = V[0] - V[3]
    prefix = str(tmp_path / "out")
    L = _lib.lib()
    rc = L.tsu_write_output(prefix.encode(), _lib.make_params(sc.params), 2, _lib.ptr(times),
                            _lib.ptr(np.ascontiguousarray(recs)), _lib.ptr(hmax), _lib.ptr(bmax), _lib.ptr(emax))
    assert rc == 0
    with netcdf_file(prefix + ".nc", "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.dimensions["x"] == 20 and f.dimensions["y"] == 12
        np.testing.assert_array_equal(f.variables["time"][:], times)
        eps = np.finfo(float).eps
        for r in range(2):
            W = recs[r]
            np.testing.assert_array_equal(f.variables["eta"][r], W[0] - W[3])
            np.testing.assert_array_equal(f.variables["ux"][r], W[1] / (W[0] + eps))
            np.testing.assert_array_equal(f.variables["uy"][r], W[2] / (W[0] + eps))
            np.testing.assert_array_equal(f.variables["b"][r], -W[3])
            np.testing.assert_array_equal(f.variables["c"][r], W[4] / (W[0] + eps))
        np.testing.assert_array_equal(f.variables["hmax"][:], hmax)
        np.testing.assert_array_equal(f.variables["bmin"][:], -bmax)
        np.testing.assert_array_equal(f.variables["eta_max"][:], emax)
        x = f.variables["x"][:]
        np.testing.assert_allclose(x, (np.arange(20) + 0.5) * 10.0)
        assert f.variables["c"].long_name == b"Volumetric Sediment Concentration"
        assert f.Conventions == b"CF-1.0"


@pytest.mark.parametrize("bad", ["order", "bc", "precision"])
def test_invalid_arguments_rejected_before_device_use(bad):
    from tsuner_amd import scenarios as S
    sc = S.plain_swe(n=8, dx=10.0)
    s = dict(sc.scheme)
    prec = 0
    if bad == "order":
        s["order"] = 3
    elif bad == "bc":
        s["bc"] = [S.BC_PERIODIC, S.BC_WALL, S.BC_WALL, S.BC_WALL]
    else:
        prec = 7
    h = C.c_void_p()
    rc = _lib.lib().tsu_state_create(C.byref(h), _lib.make_params(sc.params), _lib.make_scheme(s), prec, 0)
    assert rc in (3, 5)
    assert not h.value


def test_network_kernel_kinds_validated_before_device_use():
    """Unknown kinds (a weight-gradient kind given to the convolution, an
    unknown flag) and misaligned operands are rejected before any device
    work; the 64-wide weight-gradient kinds report their workspace."""
    L = _lib.lib()
    p = C.c_void_p(16)
    for kind in (5, 5 | 16, 32, 1 | 32, -1):
        assert L.tsu_nn_conv(kind, p, p, p, p, p, 1, 1, 1, None) == 3, kind
    assert L.tsu_nn_conv(1, C.c_void_p(8), p, p, None, p, 1, 1, 1, None) == 3      # misaligned input
    assert L.tsu_nn_conv(3 | 16, p, p, None, None, p, 1, 1, 1, None) == 3          # data gradient needs act
    for kind, my, kx in ((5, 32, 8), (6, 32, 32), (7, 8, 32), (5 | 16, 64, 8), (6 | 16, 64, 64), (7 | 16, 8, 64)):
        assert L.tsu_nn_wgrad_workspace(kind) == 9 * 256 * my * kx + 256 * my, kind
    for kind in (0, 4, 8, 6 | 32):
        assert L.tsu_nn_wgrad_workspace(kind) == -1, kind


def test_edge_kat_arguments_rejected_before_device_use():
    """The edge known-answer entries check their arguments on the host:
    null buffers, a negative count, an odd count for the paired entry and a
    normal that is not an axis direction return TSU_ERR_ARG (3) with a
    message, and no device is touched."""
    from tsuner_amd import scenarios as S
    sc = S.sediment_shelf(n=8, dx=10.0, inflow=False)
    p, s = _lib.make_params(sc.params), _lib.make_scheme(sc.scheme)
    L = _lib.lib()
    faces = np.zeros((3, 5))
    normals = np.array([(1, 0), (0, 1), (0, -1)], dtype=np.int32)
    fixed = np.zeros(3, dtype=np.int32)
    out = np.zeros((3, 5))
    args = lambda n, nrm=normals: (p, s, n, _lib.ptr(faces), _lib.ptr(faces), _lib.ptr(nrm, C.c_int),
                                   _lib.ptr(fixed, C.c_int), _lib.ptr(out), _lib.ptr(out), 0)
    assert L.tsu_kat_edge_flux_pairs(*args(3)) == 3
    assert b"even" in L.tsu_last_error()
    for entry in (L.tsu_kat_edge_flux, L.tsu_kat_edge_flux_pairs):
        assert entry(*args(-2)) == 3
        assert entry(p, s, 2, None, None, None, None, None, None, 0) == 3
        bad = np.array([(1, 1), (0, 1)], dtype=np.int32)
        assert entry(*args(2, bad)) == 3
        assert b"axis-aligned" in L.tsu_last_error()
        assert entry(*args(0)) == 0
