"""The no-allocation schedule entry point (ramgpu_ram_run_schedule_noalloc)
against ramgpu_ram_run_schedule, through the C ABI: two calls in a row, the
second on a smaller problem with a shorter schedule, each bit-identical to the
allocating entry point on a twin context (same kernels, same staging
contents).  Also: a schedule too long for the fixed staging array is refused
before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available
from helpers import random_problem
from ramgpu import inputs, species as sp
from ramgpu.capi import RamGpu, RamGpuError, SpeciesOut

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="no GPU")]

STAGING_DOUBLES = 1 << 16


def _twin_run(p, kp, bs):
    a, b = RamGpu.from_problem(p), RamGpu.from_problem(p)
    try:
        a.set_kp(p.kp)  # the grid's L shells and local times for the device-side VT
        b.set_kp(p.kp)
        ra = a.ram_run_schedule(kp, bs)
        rb = b.ram_run_schedule(kp, bs, noalloc=True)
        fa, fb = a.get_f2(), b.get_f2()
    finally:
        a.close()
        b.close()
    assert np.array_equal(fa, fb), f"F2 differs in {np.count_nonzero(fa != fb)} cells"
    assert ra == rb
    return fb


def test_noalloc_schedule_matches_two_calls_in_a_row():
    rng = np.random.default_rng(7)
    # first call: 4 species on the multi-species grid, 6 steps with boundary scales
    p1 = inputs.make_problem(20, 25, 35, 72, inputs.SPECIES4, inputs.COMPOSITION4, kp=2.0, seed=3)
    kp1 = np.linspace(2.0, 6.0, 6)
    bs1 = np.exp(0.3 * rng.standard_normal(6 * 4))
    print("noalloc: call 1, 4 species 20x25x35x72, 6 steps", flush=True)
    f1 = _twin_run(p1, kp1, bs1)
    assert np.isfinite(f1).all()
    # second call: a smaller problem and a shorter schedule, without scales;
    # the staging array still holds the first call's longer schedule
    p2 = random_problem(nr=8, nt=9, ne=7, npa=10, species=(sp.HYDROGEN, sp.OXYGEN), seed=2)
    kp2 = np.array([3.5, 1.0, 5.0])
    print("noalloc: call 2, 2 species 8x9x7x10, 3 steps", flush=True)
    _twin_run(p2, kp2, None)


def test_noalloc_refuses_a_schedule_beyond_the_array():
    p = random_problem(nr=8, nt=9, ne=7, npa=10, species=(sp.HYDROGEN, sp.OXYGEN), seed=4)
    g = RamGpu.from_problem(p)
    try:
        g.set_kp(p.kp)
        f0 = g.get_f2()
        n = STAGING_DOUBLES // (1 + g.ns) + 1
        kp = np.full(n, 3.0)
        bs = np.ones(n * g.ns)
        out = (SpeciesOut * g.ns)()
        rc = g.lib.ramgpu_ram_run_schedule_noalloc(g.h, n, kp.ctypes.data_as(C.POINTER(C.c_double)),
                                                  bs.ctypes.data_as(C.POINTER(C.c_double)), out)
        assert rc == 1  # RAMGPU_EINVAL
        assert b"fixed array" in g.lib.ramgpu_last_error()
        assert np.array_equal(g.get_f2(), f0)
        with pytest.raises(RamGpuError):
            g.ram_run_schedule(kp, bs, noalloc=True)
        # a schedule that fits runs
        g.ram_run_schedule(np.full(3, 3.0), None, noalloc=True)
    finally:
        g.close()
