"""The Fortran side of the boundary compiles: the ISO_C_BINDING module, the
example host program, and the drop-in ModRamDrift replacement against the
reference's own module interfaces (the .mod files of oracle/_ref)."""
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
FORTRAN = REPO / "ram-scb_amd" / "fortran"
REFMODS = REPO / "oracle" / "_ref" / "build"
AMDFLANG = "/opt/rocm/bin/amdflang"


def _compile(tmp_path, *srcs, incs=()):
    cmd = [AMDFLANG, "-c", "-fdefault-real-8", "-fdefault-double-8", "-module-dir", str(tmp_path)]
    cmd += [f"-I{d}" for d in incs] + [str(s) for s in srcs]
    return subprocess.run(cmd, capture_output=True, text=True, cwd=str(tmp_path))


@pytest.mark.skipif(not Path(AMDFLANG).exists(), reason="amdflang absent")
def test_interface_module_and_host_compile(tmp_path):
    r = _compile(tmp_path, FORTRAN / "ramgpu_mod.f90", FORTRAN / "ramgpu_driver.f90")
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not (REFMODS / "modramvariables.mod").exists(),
                    reason="reference modules not built (make -C oracle ref)")
def test_dropin_compiles_against_reference_modules(tmp_path):
    r = _compile(tmp_path, FORTRAN / "ramgpu_mod.f90", FORTRAN / "ModRamDrift_gpu.f90",
                 incs=(REFMODS,))
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "modramdrift.mod").exists()
