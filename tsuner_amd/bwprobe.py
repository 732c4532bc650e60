"""Measures achievable HBM bandwidth with libtsuner's plain copy kernel:
python -m tsuner_amd.bwprobe [bytes] [reps]."""
import ctypes as C
import sys

from . import _lib


def probe(nbytes: int = 4 << 30, reps: int = 20, device: int = 0) -> float:
    out = C.c_double()
    _lib.check(_lib.lib().tsu_bandwidth_probe(device, int(nbytes), int(reps), C.byref(out)), "tsu_bandwidth_probe")
    return out.value


if __name__ == "__main__":
    nb = int(float(sys.argv[1])) if len(sys.argv) > 1 else 4 << 30
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 20
    print("copy %d bytes x %d: %.1f GB/s (read + write)" % (nb, reps, probe(nb, reps)), flush=True)
