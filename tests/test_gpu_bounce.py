"""GPU parity of computehI (src/ModRamScb.f90:144-643) through the C ABI.

The device integrates the same closed forms as the oracle, in the same
order along each line, and repeats its corrections, Steffen interpolation,
smoothing and time derivatives operation for operation, so FNHS, FNIS,
BOUNHS, BOUNIS, BNES and the derivatives equal the oracle's bit for bit.
HDNS carries the Rairden density, whose 10**x is the device's pow: within
1e-13 relative.  The step that follows uses the geometry the kernels wrote,
bit for bit the same as the same arrays uploaded by ramgpu_set_geometry.
"""
import types

import numpy as np
import pytest

from conftest import gpu_available
from oracle import bounce as ob
from ramgpu import fieldlines as fl
from ramgpu import inputs
from ramgpu.capi import RamGpu, RamGpuError

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="no GPU")]

EXACT = ("fnhs", "fnis", "bounhs", "bounis", "bnes", "didt", "dibndt", "dbdt")


def _lines(p, seed):
    lines = fl.dipole_lines(p.grid.lz, p.grid.nt)
    # uneven lines: B not monotone along them, the equator not always the minimum
    lines.b = lines.b * np.random.default_rng(seed).uniform(0.97, 1.03, lines.b.shape)
    return lines


def _oracle(p, lines, prev, dthi, smooth, outside, scale_at):
    g = p.grid
    return ob.computehI(lines.x, lines.y, lines.z, lines.b, lines.chi, lines.ieq - 1, g.mu, g.pa, g.pabn, g.lz,
                        outside, scale_at, lines.bnormal, dthi, smooth, prev)


def _compare(got, want, what):
    for k in EXACT:
        d = np.argwhere(got[k] != want[k])
        assert d.size == 0, f"{what} {k}: {len(d)} cells differ, first at {d[0]}: {got[k][tuple(d[0])]} {want[k][tuple(d[0])]}"
    rel = np.max(np.abs(got["hdns"] / want["hdns"] - 1))
    assert rel <= 1e-13, (what, rel)
    assert not got["eir"][0].any() and not got["eip"][0].any()
    print(f"{what}: bit-exact, HDNS within {rel:.1e}", flush=True)


@pytest.mark.parametrize("dims,smooth,beyond", [((10, 10, 8, 20), False, True), ((10, 10, 8, 20), True, False),
                                                ((20, 25, 35, 72), True, True)])
def test_computehI_matches_oracle(dims, smooth, beyond):
    nr, nt, ne, npa = dims
    p = inputs.make_problem(nr, nt, ne, npa, inputs.SPECIES4, inputs.COMPOSITION4, kp=4.0, seed=3)
    g = RamGpu.from_problem(p)
    lines = _lines(p, 5)
    outside = np.zeros((nr, nt), dtype=np.int32)
    scale_at = np.zeros(nt, dtype=np.int32)
    if beyond:
        scale_at[2], scale_at[nt - 2] = nr - 3, nr - 1
        outside[nr - 2:, 2] = 1
        outside[nr - 1, nt - 2] = 1
    prev = g.get_geometry()
    g.computehI(lines, p.grid.pa, p.grid.pabn, p.grid.lz, outside, scale_at, smooth)
    got = g.get_geometry()
    want = _oracle(p, lines, prev, 0.0, smooth, outside, scale_at)
    _compare(got, want, f"{dims} smooth={smooth} first call")
    assert np.array_equal(got["outside_mgnp"], outside)
    # 300 s later, on changed lines: the time derivatives over DthI
    g.set_time(300.0)
    lines.b = lines.b * 1.01
    g.computehI(lines, p.grid.pa, p.grid.pabn, p.grid.lz, outside, scale_at, smooth)
    got2 = g.get_geometry()
    want2 = _oracle(p, lines, got, 300.0, smooth, outside, scale_at)
    _compare(got2, want2, f"{dims} smooth={smooth} second call")
    assert np.abs(got2["didt"]).max() > 0
    assert g.told() == 300.0
    # TOld as a restart restores it: DthI = 600 - 500
    g.set_time(600.0)
    g.set_told(500.0)
    lines.b = lines.b * 0.995
    g.computehI(lines, p.grid.pa, p.grid.pabn, p.grid.lz, outside, scale_at, smooth)
    _compare(g.get_geometry(), _oracle(p, lines, got2, 100.0, smooth, outside, scale_at),
             f"{dims} smooth={smooth} after set_told")
    g.close()


def test_step_uses_the_geometry_computehI_wrote():
    p = inputs.make_problem(10, 10, 8, 20, inputs.SPECIES4, inputs.COMPOSITION4, kp=4.0, seed=4)
    a = RamGpu.from_problem(p)
    a.computehI(_lines(p, 6), p.grid.pa, p.grid.pabn, p.grid.lz, smooth=True)
    geo = a.get_geometry()
    b = RamGpu.from_problem(p)
    b.set_geometry(types.SimpleNamespace(**geo))
    for n in range(2):
        a.ram_run(1)
        b.ram_run(1)
        d = np.argwhere(a.get_f2() != b.get_f2())
        assert d.size == 0, f"step {n}: {len(d)} cells differ"
    a.close()
    b.close()


def test_computehI_refuses_what_the_reference_cannot_do():
    p = inputs.make_problem(8, 7, 6, 10, inputs.SPECIES4, inputs.COMPOSITION4, kp=4.0, seed=1)
    g = RamGpu.from_problem(p)
    lines = _lines(p, 2)
    with pytest.raises(RamGpuError, match="smoothing"):
        g.computehI(lines, p.grid.pa, p.grid.pabn, p.grid.lz, smooth=True)   # 9x9 kernel on 8x7
    bad = np.zeros(p.grid.nt, dtype=np.int32)
    bad[1] = 2
    with pytest.raises(RamGpuError, match="ScaleAt"):
        g.computehI(lines, p.grid.pa, p.grid.pabn, p.grid.lz, scale_at=bad, smooth=False)
    g.computehI(lines, p.grid.pa, p.grid.pabn, p.grid.lz, smooth=False)
    g.close()
