"""data.dat parsing (Problem_loader.cxx:264-342 semantics) through the C ABI."""
import ctypes as C
import math
import os

from tsuner_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "data_small.dat")


def parse(path):
    p = _lib.SimParams()
    prefix = C.create_string_buffer(256)
    files = C.create_string_buffer(1024)
    slr = C.c_double()
    rc = _lib.lib().tsu_parse_data_dat(path.encode(), C.byref(p), prefix, 256, files, 1024, C.byref(slr))
    return rc, p, prefix.value.decode(), files.value.decode().split("\n"), slr.value


def test_parse_values_and_derived_fields():
    rc, p, prefix, files, slr = parse(FIXTURE)
    assert rc == 0
    assert (p.xmin, p.xmax, p.ymin, p.ymax) == (-5.0, 5.0, -4.0, 4.0)
    assert (p.total_time, p.save_time, p.npx, p.npy, p.nvol) == (0.25, 0.1, 80, 64, 80 * 64)
    assert (p.CFL, p.epsilon_h, p.R, p.DS, p.porosity, p.OC, p.Dsg) == (0.9, 1e-4, 2.65, 2e-4, 0.4, 0.0047, 100e-6)
    assert p.FR == 2.5e-2 / 8.0                        # Chezy-type coefficient / 8
    assert p.HQB == 5.0 * 2e-4
    assert p.Q == 2e-4 * math.sqrt(9.81 * (2.65 - 1.0) * 2e-4) / (1.0 - 0.4)
    assert p.Opar == p.FR / ((2.65 - 1.0) * 9.81 * 2e-4)
    assert p.dx == 10.0 / 80 and p.dy == 8.0 / 64
    assert prefix == "small_run"
    assert files == ["none_bathy.nc", "none_eta.nc", "0", "0"]
    assert slr == 0.5


def test_parse_errors(tmp_path):
    bad = tmp_path / "bad.dat"
    bad.write_text("# only a few values\n# xmin\n1\n# xmax\n2\n")
    rc, *_ = parse(str(bad))
    assert rc == 3
    rc, *_ = parse(str(tmp_path / "missing.dat"))
    assert rc == 2
