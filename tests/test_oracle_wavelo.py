"""WAVELO and its lifetime tables (src/ModRamWPI.f90:18-182,580-636), which
cannot be built here (ModRamWPI needs the absent GSL): the host tables of
ramgpu/wpi.py are pinned by the published lifetime table itself (read from
the reference's source text when it is present), by interpolation identities
and by closed forms; the oracle's WAVELO by the closed-form decay factor on
every branch (inside and outside the plasmapause, the energy corrections,
the 1 MeV switch, the wave amplitude switch at Kp 4) and its range."""
import math
import re
from pathlib import Path

import numpy as np
import pytest

from oracle.oracle import Oracle
from ramgpu import inputs, species as sp, wpi
from ramgpu.grid import PI, RE

REF_WPI = Path("/root/reference/src/ModRamWPI.f90")


@pytest.mark.skipif(not REF_WPI.exists(), reason="reference source absent")
def test_lifetime_table_is_the_published_one():
    text = REF_WPI.read_text()
    got = {(int(a), int(b)): float(v) for a, b, v in re.findall(r"rlife\((\d),(\d)\)=([0-9.]+)", text)}
    assert len(got) == 40
    for j in range(5):
        for i in range(8):
            assert wpi.R_LIFE[j][i] == got[(j + 1, i + 1)]
    assert "DATA rEa/0.2,0.5,1.0,1.5,2.0/" in text and "DATA rL/5.0,4.5,4.0,3.5,3.0,2.5,2.0,1.65/" in text


def test_wavepara1_interpolates_through_the_table_nodes():
    # at a table L the L-interpolation returns the column; at a table E the
    # E-interpolation returns that lifetime
    for i, l in enumerate(wpi.R_L):
        if l == 5.0:
            # no branch of the reference covers L = 5 itself: the previous
            # iteration's values stand
            prev = [1.0, 2.0, 3.0, 4.0, 5.0]
            assert wpi._clife_at(l, prev) == prev
            l = 5.0 + 1e-13
        cl = wpi._clife_at(l)
        for j in range(5):
            assert math.isclose(cl[j], wpi.R_LIFE[j][i], rel_tol=1e-9)
            assert math.isclose(wpi._xlife(wpi.R_EA[j] - (1e-12 if j == 4 else 0.0), cl),
                                cl[j], rel_tol=1e-9)
    # continuity across every interior node, in L and in E
    for l in wpi.R_L[1:-1]:
        a, b = wpi._clife_at(l - 1e-9), wpi._clife_at(l + 1e-9)
        assert all(math.isclose(x, y, rel_tol=1e-6) for x, y in zip(a, b))
    cl = wpi._clife_at(3.3)
    for e in wpi.R_EA[1:-1]:
        assert math.isclose(wpi._xlife(e - 1e-9, cl), wpi._xlife(e + 1e-9, cl), rel_tol=1e-6)


def test_wavepara2_closed_forms():
    g = inputs.make_problem(nr=8, nt=6, ne=12, npa=10, species=(sp.ELECTRON,), seed=1).grid
    w2, w3 = wpi.wavepara2(g, 1)
    for i in range(1, g.nr):
        cone = math.degrees(math.asin(math.sqrt(((RE + 2e5) / g.rlz[i]) ** 3 /
                                                math.sqrt(4 - 3 * (RE + 2e5) / g.rlz[i]))))
        for k in range(1, g.ne):
            e = g.ekev[k] / 1000
            rate = min(0.08 * e ** -1.32, 0.4 * 10 ** (2 * g.lz[i] - 6 + 0.4 * math.log10(29 * e)))
            assert math.isclose(w2[i, k], 86400.0 / rate, rel_tol=1e-13)
            ts = 64 * g.lz[i] * RE / 35 / 0.75 / math.sin(math.radians(cone)) ** 2 / g.v[0, k]
            assert math.isclose(w3[i, k], ts, rel_tol=1e-12)
    assert np.all(w2[0] == 0) and np.all(w2[:, 0] == 0)


@pytest.mark.parametrize("kp", [3.0, 5.0])
def test_wavelo_is_the_lifetime_decay(kp):
    p = inputs.make_problem(nr=9, nt=7, ne=12, npa=10, species=(sp.ELECTRON, sp.HYDROGEN),
                            composition=(1.0, 1.0), kp=kp, seed=3, wave=True)
    g = p.grid
    # energies across every branch: below 1.1 keV, 1.1-5 keV, above, and above 1 MeV
    g.ekev[-3:] = (1200.0, 2000.0, 3500.0)
    p.wave = inputs.wave_inputs(g, kp, kpmax12=kp + 0.5)
    w = p.wave
    inside = g.lz[None, :g.nr] <= w.rlpp[:, None]
    assert inside.any() and (~inside).any()
    o = Oracle(p)
    before = o.f2.copy()
    o.call("wavelo", 1)
    ratio = o.f2[0] / before[0]
    bw = 100.0 if kp >= 4 else 30.0
    for i in range(1, g.nr):
        for j in range(g.nt):
            for k in range(1, g.ne):
                e = g.ekev[k]
                if g.lz[i] <= w.rlpp[j]:
                    tau = w.walos1[0, i, k] * (10.0 / bw) ** 2
                elif e <= 1000.0:
                    tau = w.walos2[0, i, k] + w.walos3[0, i, k]
                    if e <= 1.1:
                        tau *= 37.5813 * math.exp(-1.81255 * e)
                    elif e <= 5.0:
                        tau *= 7.5 - 1.15 * e
                else:
                    tau = 5 * 86400 / kp
                np.testing.assert_allclose(ratio[i, j, k, 1:], math.exp(-p.dts / tau), rtol=1e-13)
    assert np.all(ratio[0] == 1) and np.all(ratio[:, :, 0] == 1) and np.all(ratio[:, :, :, 0] == 1)
    assert np.array_equal(o.f2[1], before[1])


def test_ram_run_with_wavelo_accumulates_lswae():
    """WAVELO sits in ram_run's loss chain twice for the electrons
    (src/ModRamRun.f90:91-99,151-159): the step with it equals the step
    without it composed with the decay on both sides of the charge exchange
    and atmospheric losses, and LSWAE holds the energy it removed."""
    p = inputs.make_problem(nr=7, nt=6, ne=9, npa=8, species=(sp.ELECTRON, sp.HYDROGEN),
                            composition=(1.0, 1.0), kp=3.0, seed=5, wave=True)
    o = Oracle(p)
    o.ram_run(1)
    assert o.arr("lswae")[0] > 0 and o.arr("lswae")[1] == 0
    tot = sum(o.arr(n)[0] for n in ("lsdr", "lscha", "lsatm", "lswae"))
    assert math.isclose(tot, 0.0 - o.arr("setrc")[0], rel_tol=1e-12)
    q = inputs.make_problem(nr=7, nt=6, ne=9, npa=8, species=(sp.ELECTRON, sp.HYDROGEN),
                            composition=(1.0, 1.0), kp=3.0, seed=5, wave=False)
    r = Oracle(q)
    r.ram_run(1)
    assert np.array_equal(o.f2[1], r.f2[1])          # H+ takes no WAVELO
    assert not np.allclose(o.f2[0], r.f2[0], rtol=1e-6)
