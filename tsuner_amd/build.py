"""Builds libtsuner.so for gfx950 with hipcc (in-tree, no JIT cache).

Flags that matter for results:
  -ffp-contract=off    no FMA contraction (the reference builds its device
                       code with -fmad=false, GPU/makefile:10);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                       no packed FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtsuner.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["tsu_kernels.hip", "tsu_api.cpp", "tsu_run.cpp", "tsu_netcdf.cpp", "tsu_ensemble.cpp", "tsu_datadat.cpp",
           "tsu_nn.hip", "tsu_nn_api.cpp", "tsu_gridfile.cpp", "tsu_ingest.cpp"]
CLI = os.path.join(HERE, "tsuner")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fvisibility=hidden",
          "-Wall", "-Wno-unused-function", "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include")]
DEVICE = ["-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-mcumode"]


def _compile(src: str, defines=(), build_dir: str = BUILD) -> str:
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-D" + d for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build_variant(out: str, defines) -> str:
    """The library alone, compiled with extra -D defines (e.g. the stage
    kernels' TSU_WAVES_FP64 / TSU_WAVES_FP32 occupancy targets) into `out`;
    load it with TSUNER_AMD_LIB=out."""
    bdir = os.path.join(BUILD, "variant_" + "_".join(d.replace("=", "") for d in defines))
    os.makedirs(bdir, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda s: _compile(s, defines, bdir), SOURCES))
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out] + objs + ["-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    # the `tsuner data.dat [device]` program, linked against the library
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(os.path.dirname(HERE), "include"),
           os.path.join(CSRC, "tsu_cli.cpp"), "-o", CLI + ".tmp", "-L" + HERE, "-ltsuner",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("cli build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(CLI + ".tmp", CLI)
    if verbose:
        print("built", LIB, "and", CLI)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
