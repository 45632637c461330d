"""Test-only problem generators and comparison helpers.

``random_problem`` perturbs the dipole inputs so that every branch of the
drift operators is reached: radial drifts that change sign along a line
(inflow and outflow edges, the stale outer ghost of DRIFTR), azimuthal and
energy drifts of both signs, nonzero induced fields (EIR, EIP, dBdt, dIdt,
dIbndt), cells outside the magnetopause, and negative intermediate values
that the operators clamp.
"""
from __future__ import annotations

import numpy as np

from ramgpu import inputs, species as sp
from ramgpu.grid import build_grid


def random_problem(nr=8, nt=9, ne=7, npa=10, species=(sp.HYDROGEN, sp.OXYGEN),
                   composition=None, seed=0, kp=4.0, mgnp_frac=0.1, noise=0.3,
                   field_scale=1.0, spiky=False) -> inputs.Problem:
    rng = np.random.default_rng(seed)
    composition = composition or tuple(1.0 for _ in species)
    grid = build_grid(nr, nt, ne, npa, species)
    geom = inputs.dipole_geometry(grid)
    sh3 = geom.fnhs.shape
    sh2 = geom.bnes.shape

    def pert(a, amp):
        return np.asfortranarray(a * (1.0 + amp * rng.uniform(-1, 1, a.shape)))

    geom.fnhs = pert(geom.fnhs, 0.05)
    geom.fnis = pert(geom.fnis, 0.05)
    geom.bounhs = pert(geom.bounhs, 0.05)
    geom.bounis = pert(geom.bounis, 0.05)
    geom.hdns = pert(geom.hdns, 0.2)
    geom.bnes = pert(geom.bnes, 0.05)
    # induced fields comparable to the convection terms
    geom.eir = np.asfortranarray(field_scale * 2e-4 * rng.standard_normal(sh2))
    geom.eip = np.asfortranarray(field_scale * 2e-3 * rng.standard_normal(sh2))
    geom.dbdt = np.asfortranarray(field_scale * 1e-9 * rng.standard_normal(sh2))
    geom.didt = np.asfortranarray(field_scale * 1e-4 * rng.standard_normal(sh3))
    geom.dibndt = np.asfortranarray(field_scale * 1e-4 * rng.standard_normal(sh3))
    mg = (rng.uniform(0, 1, (nr, nt)) < mgnp_frac).astype(np.int32)
    mg[:, nt - 1] = mg[:, 0]
    geom.outside_mgnp = np.asfortranarray(mg)
    vt = inputs.volland_stern(grid, kp)
    vt = np.asfortranarray(vt * (1.0 + 0.5 * rng.uniform(-1, 1, vt.shape)))
    f2 = inputs.initial_f2(grid, geom, composition, rng, noise_sigma=noise)
    if spiky:  # isolated spikes and holes drive the limiter and the clamps
        m = rng.uniform(0, 1, f2.shape)
        f2 = np.where(m < 0.05, f2 * 50.0, f2)
        f2 = np.where(m > 0.95, 0.0, f2)
        f2 = np.asfortranarray(f2)
    fgeos = inputs.boundary_fgeos(grid, composition, scale=float(rng.uniform(0.5, 2.0)))
    fgeos = np.asfortranarray(fgeos * rng.uniform(0.5, 1.5, fgeos.shape))
    return inputs.Problem(grid, geom, vt, fgeos, f2, kp, tuple(composition))


def rel_err(a, b):
    """Norm-wise relative difference ||a-b|| / ||b||."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def max_rel_err(a, b, floor=1e-12):
    """Max relative difference over cells with |b| > floor*max|b|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m = np.abs(b).max()
    mask = np.abs(b) > floor * m
    if not mask.any():
        return 0.0
    return float((np.abs(a - b)[mask] / np.abs(b)[mask]).max())
