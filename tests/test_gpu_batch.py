"""GPU ensembles: members of different grid sizes stepped together in one
launch per kernel (graph-captured step) must give, bit for bit, what each
member gives stepped alone; one member is also checked against the oracle.
"""
import numpy as np
import pytest

from conftest import gpu_available
from helpers import max_rel_err, random_problem, rel_err
from oracle.oracle import Oracle
from ramgpu import inputs, species as sp
from ramgpu.capi import RamBatch, RamGpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="no GPU")]


def _members():
    """Uneven members: different grids, species sets and fields; the
    electrons of the second take WAVELO."""
    m2 = random_problem(nr=10, nt=12, ne=9, npa=14, species=(sp.ELECTRON, sp.HYDROGEN, sp.HELIUM), seed=22)
    m2.wave = inputs.wave_inputs(m2.grid, m2.kp, kpmax12=4.0)
    return [
        random_problem(nr=8, nt=9, ne=7, npa=10, species=(sp.HYDROGEN, sp.OXYGEN), seed=21),
        m2,
        inputs.make_problem(nr=6, nt=7, ne=6, npa=8, species=(sp.OXYGEN,), kp=2.0, seed=23),
    ]


def _assert_same(a, b, what):
    assert a.shape == b.shape, what
    diff = np.argwhere(a != b)
    assert diff.size == 0, f"{what}: {len(diff)} cells differ, first at {diff[0]}"


def _outs_same(oa, ob, what):
    for s, (x, y) in enumerate(zip(oa, ob)):
        for k in x:
            assert x[k] == y[k], (what, s, k, x[k], y[k])


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_equals_members_alone(graph):
    probs = _members()
    alone = [RamGpu.from_problem(p) for p in probs]
    outs_alone = [g.ram_run(3) for g in alone]
    mem = [RamGpu.from_problem(p) for p in probs]
    b = RamBatch(mem)
    b.use_graph(graph)
    outs = b.ram_run(3)
    for m, (ga, gm) in enumerate(zip(alone, mem)):
        _assert_same(gm.get_f2(), ga.get_f2(), f"member {m} F2 (graph={graph})")
        _outs_same(outs[m], outs_alone[m], f"member {m}")
    print(f"ensemble of {len(mem)} uneven members == members alone (graph={graph})", flush=True)
    b.close()
    for g in alone + mem:
        g.close()


def test_ensemble_schedule_equals_members_alone():
    probs = _members()
    nsteps = 4
    rng = np.random.default_rng(5)
    kp = rng.uniform(1.0, 7.0, (nsteps, len(probs)))
    sum_ns = sum(p.grid.ns for p in probs)
    bs = np.exp(0.5 * rng.standard_normal((nsteps, sum_ns)))
    alone, outs_alone, off = [], [], 0
    for m, p in enumerate(probs):
        g = RamGpu.from_problem(p)
        g.set_kp(3.0)
        ns = p.grid.ns
        outs_alone.append(g.ram_run_schedule(kp[:, m], bs[:, off:off + ns].ravel()))
        off += ns
        alone.append(g)
    mem = []
    for p in probs:
        g = RamGpu.from_problem(p)
        g.set_kp(3.0)
        mem.append(g)
    b = RamBatch(mem)
    outs = b.ram_run_schedule(kp, bs)
    for m, (ga, gm) in enumerate(zip(alone, mem)):
        _assert_same(gm.get_f2(), ga.get_f2(), f"member {m} F2 (schedule)")
        _outs_same(outs[m], outs_alone[m], f"member {m} (schedule)")
    # a second scheduled run replays the same graph with the counter reset
    outs2 = b.ram_run_schedule(kp, bs)
    off = 0
    for m, (p, ga) in enumerate(zip(probs, alone)):
        ns = p.grid.ns
        ref = ga.ram_run_schedule(kp[:, m], bs[:, off:off + ns].ravel())
        off += ns
        _assert_same(mem[m].get_f2(), ga.get_f2(), f"member {m} F2 (second schedule)")
        _outs_same(outs2[m], ref, f"member {m} (second schedule)")
    b.close()
    for g in alone + mem:
        g.close()


def test_autotune_keeps_state_and_results():
    probs = _members()
    mem = [RamGpu.from_problem(p) for p in probs]
    before = [g.get_f2() for g in mem]
    b = RamBatch(mem)
    ms = b.autotune(reps=2)
    shapes = b.shapes()
    print(f"autotune best ms {ms} shapes {shapes}", flush=True)
    for m, g in enumerate(mem):
        _assert_same(g.get_f2(), before[m], f"member {m} F2 after autotune")
    outs = b.ram_run(2)
    alone = [RamGpu.from_problem(p) for p in probs]
    for m, ga in enumerate(alone):
        ref = ga.ram_run(2)
        _assert_same(mem[m].get_f2(), ga.get_f2(), f"member {m} F2 (tuned shapes)")
        _outs_same(outs[m], ref, f"member {m} (tuned)")
    # every fixed persistent grid gives the same result too
    for cap in (64, 256):
        for m, (g, p) in enumerate(zip(mem, probs)):
            g.put_f2(p.f2)
        b.set_shapes(cap, cap, cap)
        b.ram_run(2)
        for m, ga in enumerate(alone):
            _assert_same(mem[m].get_f2(), ga.get_f2(), f"member {m} F2 (cap {cap})")
    b.close()
    for g in alone + mem:
        g.close()


def test_ensemble_member_matches_oracle():
    probs = _members()
    mem = [RamGpu.from_problem(p) for p in probs]
    b = RamBatch(mem)
    b.ram_run(2)
    for m, p in enumerate(probs):
        o = Oracle(p)
        o.ram_run(4)
        o.ram_run(4)
        e1, e2 = rel_err(mem[m].get_f2(), o.f2), max_rel_err(mem[m].get_f2(), o.f2)
        print(f"member {m}: norm rel {e1:.3e} max rel {e2:.3e}", flush=True)
        assert e1 <= 1e-12 and e2 <= 1e-12
    b.close()
    for g in mem:
        g.close()
