"""GPU parity: the HIP path through the C ABI against the oracle.

Tolerance (the stated bar for this FP64 path): the device result must match
the oracle's to <= 1e-12 relative, norm-wise and cell-wise over cells above
1e-12 of the maximum.  The drift operators are in fact bit-identical
(checked exactly in test_drift_operators_bit_exact); the loss factors use
a table-driven exponential and the device's log/log10/pow, which may differ
from the host's in the last ulp or two.  SUMRC sums are compared to 1e-12 relative of SETRC (the oracle sums
sequentially like the reference; the device sums in a tree).
"""
import numpy as np
import pytest

from conftest import gpu_available
from helpers import max_rel_err, random_problem, rel_err
from oracle.oracle import Oracle
from ramgpu import inputs
from ramgpu.capi import RamGpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="no GPU")]

RTOL = 1e-12


def _assert_close(a, b, what):
    e1, e2 = rel_err(a, b), max_rel_err(a, b)
    print(f"{what}: norm rel {e1:.3e}  max rel {e2:.3e}", flush=True)
    assert e1 <= RTOL and e2 <= RTOL, f"{what}: norm {e1:.3e} max {e2:.3e}"


OPS = ["driftr", "driftp", "drifte", "driftmu", "charexchange", "atmol"]


@pytest.mark.parametrize("seed", [0, 1])
def test_operators_random_geometry(seed):
    p = random_problem(seed=seed, spiky=(seed == 1))
    o = Oracle(p)
    g = RamGpu.from_problem(p)
    for s in (1, 2):
        o.call("cepara", s)
        o.call("driftpara", s)
        g.cepara(s)
        g.driftpara(s)
        for op in OPS + ["driftmu", "drifte", "driftp", "driftr"]:
            o.call(op, s)
            dt = getattr(g, op)(s)
            _assert_close(g.get_f2(), o.f2, f"{op}(S={s}) seed={seed}")
            if op.startswith("drift"):
                ref = o.dtdrift()[s - 1, ["driftr", "driftp", "drifte", "driftmu"].index(op)]
                assert abs(dt - ref) <= RTOL * abs(ref), (op, dt, ref)
    g.close()


def test_sumrc_matches():
    p = random_problem(seed=4)
    o = Oracle(p)
    g = RamGpu.from_problem(p)
    for s in (1, 2):
        o.call("sumrc", s)
        setrc, elorc = g.sumrc(s)
        ref = o.arr("setrc")[s - 1]
        assert abs(setrc - ref) <= RTOL * abs(ref), (setrc, ref)
        assert abs(elorc - o.arr("elorc")[s - 1]) <= RTOL * abs(ref)
    g.close()


def _run_compare(p, nsteps, what):
    o = Oracle(p)
    g = RamGpu.from_problem(p)
    for _ in range(nsteps):
        o.ram_run(4)
    out = g.ram_run(nsteps)
    _assert_close(g.get_f2(), o.f2, what)
    dt = o.dtdrift()
    for s in range(p.grid.ns):
        ref = dict(setrc=o.arr("setrc")[s], lsdr=o.arr("lsdr")[s], lscha=o.arr("lscha")[s],
                   lsatm=o.arr("lsatm")[s], lswae=o.arr("lswae")[s])
        scale = abs(ref["setrc"])
        for k, v in ref.items():
            assert abs(out[s][k] - v) <= RTOL * scale, (what, s, k, out[s][k], v)
        for n, k in enumerate(("dtdrift_r", "dtdrift_p", "dtdrift_e", "dtdrift_mu")):
            assert abs(out[s][k] - dt[s, n]) <= RTOL * abs(dt[s, n]), (what, s, k)
    g.close()


def test_ram_run_plain_h_one_step():
    _run_compare(inputs.plain_h(seed=11), 1, "plain H+ 20x25x35x72, 1 step")


@pytest.mark.parametrize("seed", [2, 3])
def test_ram_run_random_geometry(seed):
    _run_compare(random_problem(seed=seed, spiky=True), 3, f"random geometry seed={seed}, 3 steps")


def test_ram_run_multi_species():
    p = inputs.multi_species(seed=5)
    _run_compare(p, 2, "e-,H+,He+,O+ 20x25x35x72, 2 steps")


def test_ram_run_with_wavelo():
    # the reference's default electron physics: WAVELO twice in the loss chain
    # (src/ModRamRun.f90:91-99,151-159), LSWAE its energy
    p = inputs.make_problem(20, 25, 35, 72, inputs.SPECIES4, inputs.COMPOSITION4, kp=3.0, seed=13, wave=True)
    assert p.wave.wpi.tolist() == [1, 0, 0, 0]
    _run_compare(p, 2, "e-(WAVELO),H+,He+,O+ 20x25x35x72, 2 steps")


def test_wavelo_operator_and_kp_switch():
    # WAVELO alone, then a Kp ramp across the wave-amplitude switch at Kp 4
    p = random_problem(seed=14, species=(inputs.SPECIES4[0], inputs.SPECIES4[1]))
    p.wave = inputs.wave_inputs(p.grid, 3.0, kpmax12=4.5)
    o = Oracle(p)
    g = RamGpu.from_problem(p)
    o.call("wavelo", 1)
    g.wavelo(1)
    _assert_close(g.get_f2(), o.f2, "WAVELO(S=1)")
    g.set_kp(3.0)
    kps = np.array([3.0, 3.9, 4.0, 5.5])
    for kp in kps:
        o.set_vt(inputs.volland_stern(p.grid, kp))
        o.set_wave_kp(kp)
        o.ram_run(2)
    out = g.ram_run_schedule(kps)
    _assert_close(g.get_f2(), o.f2, "Kp 3 -> 5.5 schedule with WAVELO")
    for s in range(p.grid.ns):
        for k in ("setrc", "lsdr", "lscha", "lsatm", "lswae"):
            ref = o.arr(k)[s]
            assert abs(out[s][k] - ref) <= RTOL * abs(o.arr("setrc")[s]), (s, k, out[s][k], ref)
    g.close()


def test_ram_run_uneven_large_grid():
    p = inputs.make_problem(nr=40, nt=48, ne=48, npa=90, species=inputs.SPECIES4,
                            composition=inputs.COMPOSITION4, kp=4.0, seed=9)
    _run_compare(p, 1, "4 species 40x48x48x90, 1 step")


def test_ram_run_high_resolution_planes():
    # 64x90 planes (the high-resolution ensemble's): 1024-thread plane blocks
    p = inputs.make_problem(nr=8, nt=10, ne=64, npa=90, species=(inputs.SPECIES4[1], inputs.SPECIES4[3]),
                            composition=(0.7, 0.25), kp=5.0, seed=12)
    _run_compare(p, 2, "H+,O+ 8x10x64x90, 2 steps")


@pytest.mark.parametrize("shape", [(256, 16), (768, 8), (1024, 8)])
def test_plane_shape_changes_no_drift_result(shape):
    # every launch shape of the plane kernels gives the same F2, bit for bit
    # (the per-cell arithmetic does not depend on the thread mapping); the
    # SUMRC sums add in the mapping's order and stay within the bar
    p = inputs.multi_species(seed=8)
    g0 = RamGpu.from_problem(p)
    out0 = g0.ram_run(2)
    g1 = RamGpu.from_problem(p)
    g1.set_plane_shape(*shape)
    assert g1.plane_shape() == shape
    out1 = g1.ram_run(2)
    assert np.array_equal(g0.get_f2(), g1.get_f2()), f"plane shape {shape} changed F2"
    for a, b in zip(out0, out1):
        for k in ("setrc", "lsdr", "lscha", "lsatm"):
            assert abs(a[k] - b[k]) <= RTOL * abs(a["setrc"]), (shape, k, a[k], b[k])
        for k in ("dtdrift_r", "dtdrift_p", "dtdrift_e", "dtdrift_mu"):
            assert a[k] == b[k], (shape, k)
    g1.set_plane_shape(0, 0)
    assert g1.plane_shape() == g0.plane_shape()
    g0.close()
    g1.close()


def test_ram_run_schedule_kp_ramp():
    p = inputs.multi_species(seed=6)
    sched = inputs.multi_species_schedule(4, seed=6)
    o = Oracle(p)
    g = RamGpu.from_problem(p)
    g.set_kp(sched[0][0])
    kps = np.array([k for k, _ in sched])
    bs = np.repeat(np.array([b for _, b in sched]), p.grid.ns)
    for kp, b in sched:
        o.set_vt(inputs.volland_stern(p.grid, kp))
        o.set_wave_kp(kp)  # the electrons' WAVELO follows the step's Kp
        o.set_fgeos(p.fgeos * b)
        o.ram_run(4)
    g.ram_run_schedule(kps, bs)
    _assert_close(g.get_f2(), o.f2, "Kp ramp schedule, 4 steps")
    g.close()


@pytest.mark.parametrize("seed", [0, 1, 5])
def test_drift_operators_bit_exact(seed):
    """Drift operators alone involve no transcendental beyond sqrt: the
    device follows the reference's rounding sequence, so F2 and DtDrift* are
    bit-identical to the oracle (which is bit-identical to the reference)."""
    p = random_problem(seed=seed, spiky=True)
    o = Oracle(p)
    g = RamGpu.from_problem(p)
    for s in (1, 2):
        o.call("driftpara", s)
        g.driftpara(s)
        for op in ["driftr", "driftp", "drifte", "driftmu", "driftmu", "drifte", "driftp", "driftr"]:
            o.call(op, s)
            dt = getattr(g, op)(s)
            a, b = g.get_f2(), o.f2
            nd = int(np.count_nonzero(a != b))
            print(f"{op}(S={s}) seed={seed}: {nd} cells differ", flush=True)
            assert nd == 0, f"{op}(S={s}): {nd} cells differ, max rel {max_rel_err(a, b):.3e}"
            ref = o.dtdrift()[s - 1, ["driftr", "driftp", "drifte", "driftmu"].index(op)]
            assert dt == ref, (op, dt, ref)
    g.close()


def test_sweep_lds_limits_at_the_edge():
    # 4x4 planes: a 256-line block spans 65 (J,L) or (I,L) pairs and 17 J or
    # I values, so the face terms the sweeps stage in LDS grow fastest with
    # NR (radial) and NT (azimuthal).  One past each limit is refused; at the
    # limit (about 159 KB of face terms) the step matches the oracle.
    import ctypes as C
    from ramgpu import capi
    lib = capi.load_library()
    h = C.c_void_p()
    assert lib.ramgpu_create(C.byref(h), 1, 8, 63, 4, 4, 0) == 1
    assert b"NT=63" in lib.ramgpu_last_error()
    assert lib.ramgpu_create(C.byref(h), 1, 150, 9, 4, 4, 0) == 1
    assert b"NR=150" in lib.ramgpu_last_error()
    _run_compare(random_problem(nr=8, nt=62, ne=4, npa=4, seed=21, spiky=True), 2, "8x62x4x4, 2 steps")
    _run_compare(random_problem(nr=149, nt=9, ne=4, npa=4, seed=22, spiky=True), 2, "149x9x4x4, 2 steps")
