"""The oracle's implicit pitch-angle diffusion: WPADIF (src/ModRamWPI.f90:
643-714) and FLCscatter (src/ModRamLoss.f90:513-575).

The reference's modules holding them need GSL and NetCDF, absent here, so
the restatement is pinned by what the mathematics fixes:

* the line solve equals numpy's dense solve of the tridiagonal system
  -GN g(L-1) + (1+AN+GN) g(L) - AN g(L+1) = F2(L)/FACMU(L), L = 2..NPA-1,
  with the boundary conditions the recurrences impose, g(1) = g(2) (RK(1)=0,
  RL(1)=-1) and g(NPA) = g(NPA-1), and F2 = g*FACMU on L = 1..NPA;
* the scheme is conservative: with the weights FACMU*WMU its couplings are
  symmetric (FACMU(L)*WMU(L)*AN(L) = FACMU(L+1)*WMU(L+1)*GN(L+1) =
  DTs*D(L)/DMU(L)), so sum over L=2..NPA-1 of WMU(L)*F2(L) of every line is
  unchanged -- a wrong DMU, WMU or FACMU index breaks it;
* F2 = c*FNHS*MU (uniform phase-space density along the line) is a fixed
  point;
* FLCscatter is the same solve with D = FLC_coef, and does nothing before
  TimeRamElapsed reaches Dt_bc;
* ram_run places WPADIF at the WPI positions in place of WAVELO, at the
  EMIC positions, and FLCscatter around the atmospheric losses without a
  SUMRC (src/ModRamRun.f90:91-150).
"""
import numpy as np

from oracle.oracle import Oracle
from ramgpu import inputs


def _problem(seed=0, nr=6, nt=5, ne=7, npa=20):
    p = inputs.make_problem(nr=nr, nt=nt, ne=ne, npa=npa, species=inputs.SPECIES4,
                            composition=inputs.COMPOSITION4, kp=4.0, seed=seed)
    # the pitch-angle grid of this NPA is uniform; uneven widths let an index
    # mistake in DMU or WMU show
    rng = np.random.default_rng(seed + 50)
    p.grid.dmu = p.grid.dmu * rng.uniform(0.7, 1.3, npa)
    p.grid.wmu = p.grid.wmu * rng.uniform(0.7, 1.3, npa)
    d1 = inputs.pitch_diffusion_coef(p.grid, p.geom, seed=seed + 1)
    d2 = inputs.pitch_diffusion_coef(p.grid, p.geom, scale=3e-6, seed=seed + 2)
    return p, d1, d2


def test_wpadif_is_the_tridiagonal_solve():
    p, d1, d2 = _problem(1)
    g = p.grid
    o = Oracle(p)
    s = 2
    f0 = o.f2.copy()
    o.wpadif(s, d1, d2)
    out = o.f2
    npa, dts = g.npa, p.dts
    rng = np.random.default_rng(5)
    for _ in range(12):
        i, j, k = int(rng.integers(1, g.nr)), int(rng.integers(0, g.nt)), int(rng.integers(1, g.ne))
        fac = p.geom.fnhs[i, j, :] * g.mu
        dd = d1[i, j, k, :] + d2[i, j, k, :]
        n = npa - 2                        # unknowns g(2..NPA-1)
        a = np.zeros((n, n))
        rhs = f0[s - 1, i, j, k, 1:npa - 1] / fac[1:npa - 1]
        for r in range(n):
            l = r + 1                      # 0-based row of L = r+2
            an = dd[l] / g.dmu[l] * dts / fac[l] / g.wmu[l]
            gn = dd[l - 1] / g.dmu[l - 1] * dts / fac[l] / g.wmu[l]
            a[r, r] = 1.0 + an + gn
            if r > 0:
                a[r, r - 1] = -gn
            else:
                a[r, r] -= gn              # g(1) = g(2)
            if r < n - 1:
                a[r, r + 1] = -an
            else:
                a[r, r] -= an              # g(NPA) = g(NPA-1)
        gs = np.linalg.solve(a, rhs)
        want = np.concatenate([[gs[0]], gs, [gs[-1]]]) * fac
        got = out[s - 1, i, j, k, :]
        assert np.allclose(got, want, rtol=1e-12, atol=0), (i, j, k, np.max(np.abs(got / want - 1)))
    # other species untouched, I=1 and K=1 untouched
    assert np.array_equal(out[0], f0[0]) and np.array_equal(out[s - 1, 0], f0[s - 1, 0])
    assert np.array_equal(out[s - 1, :, :, 0], f0[s - 1, :, :, 0])
    o.close()


def test_wpadif_conserves_and_keeps_uniform_density():
    p, d1, d2 = _problem(2, npa=24)
    g = p.grid
    o = Oracle(p)
    f0 = o.f2.copy()
    for s in range(1, g.ns + 1):
        o.wpadif(s, d1, d2)
    lines = (slice(None), slice(1, None), slice(None), slice(1, None))
    w = g.wmu[1:-1]
    a = np.einsum("sijkl,l->sijk", f0[..., 1:-1], w)[lines]
    b = np.einsum("sijkl,l->sijk", o.f2[..., 1:-1], w)[lines]
    assert np.max(np.abs(a - b) / a) < 1e-13
    assert np.max(np.abs(o.f2[lines][..., 1:] / f0[lines][..., 1:] - 1)) > 1e-2  # it diffused
    # uniform phase-space density along a line is a fixed point
    fac = p.geom.fnhs[None, :g.nr, :, None, :] * g.mu[None, None, None, None, :]
    c = np.random.default_rng(3).uniform(1.0, 2.0, (g.ns, g.nr, g.nt, g.ne, 1))
    o.set_f2(np.asfortranarray(c * fac))
    before = o.f2.copy()
    o.wpadif(1, d1, d2)
    x, y = o.f2[lines][..., 1:-1], before[lines][..., 1:-1]
    assert np.max(np.abs(x / y - 1)) < 1e-13
    o.close()


def test_flcscatter_waits_for_dt_bc_and_is_the_same_solve():
    p, d1, _ = _problem(3)
    o = Oracle(p)
    o.set_coulomb(time_elapsed=100.0, on=False)
    f0 = o.f2.copy()
    o.flcscatter(3, d1, dt_bc=300.0)
    assert np.array_equal(o.f2, f0)
    o.flcscatter(3, d1, dt_bc=100.0)
    r = Oracle(p)
    r.wpadif(3, d1)
    assert np.array_equal(o.f2, r.f2) and not np.array_equal(o.f2, f0)
    o.close()
    r.close()


def test_ram_run_places_wpadif_emic_and_flc():
    p, d1, d2 = _problem(4)
    p.wave = inputs.wave_inputs(p.grid, 4.0)   # electrons have species%WPI
    flc = inputs.pitch_diffusion_coef(p.grid, p.geom, scale=2e-6, seed=9)
    o = Oracle(p)
    o.set_coulomb(time_elapsed=600.0, on=False)
    o.set_wpadif(1, d1, d2, at_wpi=True, at_emic=False)     # e-: WPADIF in place of WAVELO
    o.set_wpadif(2, d2, d1, at_wpi=False, at_emic=True)     # H+: EMIC
    o.set_flc(3, flc, dt_bc=300.0)                          # He+: FLC
    o.ram_run(1)
    m = Oracle(p)
    m.set_coulomb(time_elapsed=600.0, on=False)
    ls = {n: np.zeros(p.grid.ns) for n in ("lsdr", "lswae", "lscha", "lsatm")}

    def sumrc(s, acc):
        m.call("sumrc", s)
        ls[acc][s - 1] += m.arr("elorc")[s - 1]

    for s in range(1, p.grid.ns + 1):
        cex = p.grid.cex_kind[s - 1] != 0
        for op in ("cepara", "driftpara", "driftr", "driftp", "drifte", "driftmu"):
            m.call(op, s)
        sumrc(s, "lsdr")
        half = []
        if s == 1:
            half.append(("wpadif", (d1, d2), "lswae"))
        elif p.wave.wpi[s - 1]:
            half.append(("wavelo", None, "lswae"))
        if s == 2:
            half.append(("wpadif", (d2, d1), "lswae"))
        if cex:
            half.append(("charexchange", None, "lscha"))
        if s == 3:
            half.append(("flc", flc, None))
        for op, arg, acc in half:
            (m.wpadif(s, *arg) if op == "wpadif" else m.flcscatter(s, arg, 300.0) if op == "flc"
             else m.call(op, s))
            if acc:
                sumrc(s, acc)
        for _ in range(2):
            m.call("atmol", s)
            sumrc(s, "lsatm")
        for op, arg, acc in reversed(half):
            (m.wpadif(s, *arg) if op == "wpadif" else m.flcscatter(s, arg, 300.0) if op == "flc"
             else m.call(op, s))
            if acc:
                sumrc(s, acc)
        for op in ("driftmu", "drifte", "driftp", "driftr"):
            m.call(op, s)
        sumrc(s, "lsdr")
    f = m.f2.copy()
    f[:, :, -1] = f[:, :, 0]
    f[:, p.geom.outside_mgnp == 1] = 1e-31
    assert np.array_equal(o.f2, f)
    for n, v in ls.items():
        assert np.array_equal(o.arr(n), v), n
    assert ls["lswae"][1] != 0  # H+ EMIC losses were booked
    o.close()
    m.close()
