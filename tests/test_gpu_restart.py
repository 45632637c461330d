"""A run resumed from a restart file (the reference's layout,
ramgpu/restart.py) continues bit for bit as the run that wrote it."""
import numpy as np
import pytest

from conftest import gpu_available
from ramgpu import inputs, restart
from ramgpu.capi import RamGpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="no GPU")]


def test_resume_from_restart_is_bit_exact(tmp_path):
    p = inputs.make_problem(20, 25, 35, 72, inputs.SPECIES4, inputs.COMPOSITION4, kp=3.0, seed=12, wave=True,
                            coulomb=True)
    a = RamGpu.from_problem(p)
    out = a.ram_run(3)
    dts_next = min(min(o[k] for k in ("dtdrift_r", "dtdrift_p", "dtdrift_e", "dtdrift_mu")) for o in out)
    restart.save(tmp_path, a, p, niter=3, dts_next=dts_next)
    b = RamGpu.from_problem(p)
    r = restart.load_into(tmp_path, b, p)
    assert b.time() == a.time() == 30.0 and r["niter"] == 3 and r["DtNext"] == dts_next
    a.ram_run(2)
    b.ram_run(2)
    fa, fb = a.get_f2(), b.get_f2()
    diff = np.argwhere(fa != fb)
    assert diff.size == 0, f"{len(diff)} cells differ after resuming, first at {diff[0]}"
    print("resumed run equals the continuous one over 2 steps", flush=True)
    a.close()
    b.close()
