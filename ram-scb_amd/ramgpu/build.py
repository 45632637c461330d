"""Build the gfx950 library (libramgpu.so) in-tree with hipcc.

    python -m ramgpu.build        (from ram-scb_amd/)
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent             # ram-scb_amd/
REPO = ROOT.parent
CSRC = ROOT / "csrc"
INCLUDE = REPO / "include"
LIB = PKG / "libramgpu.so"

HIP_SOURCES = ["kernels.hip", "api.hip", "batch.hip", "adjoint.hip", "adjoint_api.hip", "moments.hip",
               "coulomb.hip", "pitch.hip", "bounce.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off", "-Wall",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or Path(c).exists()):
            return c
    raise FileNotFoundError("hipcc not found")


def build_library(verbose: bool = False) -> Path:
    srcs = [str(CSRC / s) for s in HIP_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", *srcs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{err}")
    if verbose and err.strip():
        print(err)
    os.replace(tmp, LIB)
    return LIB


FORTRAN = ROOT / "fortran"
FDRIVER = FORTRAN / "ramgpu_driver"


def build_fortran(verbose: bool = False) -> Path:
    """The ISO_C_BINDING module and the example Fortran host program, linked
    against libramgpu.so (rpath to its directory)."""
    fc = os.environ.get("AMDFLANG", "/opt/rocm/bin/amdflang")
    obj = FORTRAN / "_obj"
    obj.mkdir(exist_ok=True)
    cmd = [fc, "-O2", f"-module-dir", str(obj), str(FORTRAN / "ramgpu_mod.f90"),
           str(FORTRAN / "ramgpu_driver.f90"), "-o", str(FDRIVER), f"-L{PKG}", "-lramgpu",
           f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../ramgpu"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=str(obj))
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed ({r.returncode}):\n{r.stderr}")
    return FDRIVER


LEVEL1_SRC = REPO / "tests" / "fortran"
LEVEL1_DRIVER = LEVEL1_SRC / "_build" / "level1_driver"


def build_level1(verbose: bool = False) -> Path:
    """The Level-1 drop-in (fortran/ModRamDrift_gpu.f90) compiled against the
    test-side module stand-ins of tests/fortran/ and linked into the test's
    host program, with OpenMP and the reference's precision flags."""
    fc = os.environ.get("AMDFLANG", "/opt/rocm/bin/amdflang")
    obj = LEVEL1_DRIVER.parent
    obj.mkdir(exist_ok=True)
    srcs = [FORTRAN / "ramgpu_mod.f90", LEVEL1_SRC / "level1_modules.f90",
            FORTRAN / "ModRamDrift_gpu.f90", LEVEL1_SRC / "level1_driver.f90"]
    cmd = [fc, "-O2", "-fopenmp", "-fdefault-real-8", "-fdefault-double-8", "-module-dir", str(obj),
           *map(str, srcs), "-o", str(LEVEL1_DRIVER), f"-L{PKG}", "-lramgpu", f"-Wl,-rpath,{PKG}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=str(obj))
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed ({r.returncode}):\n{r.stderr}")
    return LEVEL1_DRIVER


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
    print(build_level1(verbose="-v" in sys.argv))
