"""The C ABI library: loads without a GPU, exports every entry point that
include/ramgpu.h declares, and fails loudly (no CPU fallback) when no device
is present."""
import ctypes as C
import re
from pathlib import Path

import pytest

from conftest import gpu_available
from ramgpu import capi

HEADER = Path(__file__).resolve().parent.parent / "include" / "ramgpu.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(ramgpu_[A-Za-z_0-9]+)\s*\(", text)))


def test_header_declares_the_operator_set():
    names = declared()
    for op in ("driftpara", "driftr", "driftp", "drifte", "driftmu", "driftend", "cepara",
               "charexchange", "atmol", "sumrc", "ram_run"):
        assert f"ramgpu_{op}" in names


def test_library_exports_every_declared_symbol():
    lib = capi.load_library()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared()) == set(capi.EXPORTS)


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_no_device_is_an_error_not_a_fallback():
    lib = capi.load_library()
    h = C.c_void_p()
    rc = lib.ramgpu_create(C.byref(h), 1, 8, 9, 8, 10, 0)
    assert rc == 3  # RAMGPU_EHIP
    assert b"device" in lib.ramgpu_last_error()
    with pytest.raises(capi.RamGpuError):
        capi.RamGpu(1, 8, 9, 8, 10)


def test_argument_errors_are_reported():
    lib = capi.load_library()
    h = C.c_void_p()
    assert lib.ramgpu_create(C.byref(h), 0, 8, 9, 8, 10, 0) == 1  # NS=0
    assert b"NS" in lib.ramgpu_last_error()
    assert lib.ramgpu_create(C.byref(h), 1, 3, 9, 8, 10, 0) == 1  # NR<4
    assert lib.ramgpu_create(None, 1, 8, 9, 8, 10, 0) == 1
    assert lib.ramgpu_driftr(None, 1, None) == 1
