"""Properties at the benchmark's full size (GPU): the 128x64x64 Harris sheet
on a 20x10x10 box with its four species at 5x5x5 particles per cell,
262,144,000 particles generated on the device, through two engine phases,
a tile sort and a third phase, in FP64 and in the FP32 form; and the stress
shape (eight species, 314,572,800 particles, FP32) through eleven phases
with a tile sort before every fifth.

No oracle finishes at this size; the checks are what the mathematics fixes
whatever the size, on every particle and node:

* every particle comes back at its load index exactly once: the
  destination arrays start as NaN, and q (which the push leaves alone, and
  which varies with y for the sheet species) returns bit for bit;
* partition of unity (charge): sum(rhon) / invVOL = sum(q);
* the current: sum(Jx) / invVOL = sum(q u), with the pushed u;
* the second moment: sum(pxx) / invVOL = sum(q u^2);
* linear precision along x, which fixes the pairing of weights and nodes:
  sum_n(x_n rhon_n) / invVOL = sum(q x), node i at x = (i - 1) dx
  (src/Grid.cpp:54-56), before any ghost-node fold.

Tolerances, relative to the sum of magnitudes: 1e-12 in FP64; 2e-6 in the
FP32 form (FP32 weights and register sums, FP32 tile partials).
"""
from __future__ import annotations

import numpy as np
import pytest

import sputnipic_amd as S

pytestmark = pytest.mark.gpu

GRID = (128, 64, 64)
BOX = (20.0, 10.0, 10.0)
PPC = (5, 5, 5)


def _close(got: float, want: float, scale: float, tol: float, what: str) -> None:
    assert abs(got - want) <= tol * scale, f"{what}: {got} vs {want} (scale {scale})"


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_full_size_phases_and_tile_sort_keep_the_invariants(prec):
    _run(prec, S.harris_species(PPC), phases=3, sort_every=0, sort_after=2)


def test_stress_shape_keeps_the_invariants():
    """The stress shape: the four Harris species plus four x-beams (drift
    +-0.1, thermal speed 0.02) at 1x5x5 per cell, 314,572,800 particles in
    FP32, a tile sort before every fifth phase, eleven phases: the beams
    cross tile faces every few phases, so the tile sorts move many particles
    between segments."""
    _run("f32", S.harris_species(PPC) + S.beam_species(), phases=11, sort_every=5,
         sort_after=None)


def _run(prec, species, phases, sort_every, sort_after):
    S.set_async(False)
    S.set_strict_fp(False)
    grd = S.Grid(*GRID, *BOX, prec)
    gem = S.GemParams()
    field = S.gem_fields(grd, gem)
    ns = len(species)
    eng = S.Engine(grd, ns, set_moments=True, sort_every=sort_every)
    counts, q0 = [], []
    for s, sp in enumerate(species):
        p = S.gem_particles(grd, gem, sp, s, seed=0)
        counts.append(p.npmax)
        q0.append(p.d["q"].download())
        eng.load(s, p)
        for d in p.d.values():
            d.free()
    assert sum(counts) == {4: 262_144_000, 8: 314_572_800}[ns]
    moms = [S.InterpDensSpecies(grd) for _ in range(ns)]
    param = S.Parameters(dt=0.25, c=1.0, NiterMover=3)
    for ph in range(phases):
        if ph == sort_after:
            eng.sort()
            assert eng.stats()["overflow"] == 0
        eng.phase(field, moms, param)
    if sort_every:
        assert eng.stats()["sorts"] == ns * (1 + (phases - 1) // sort_every)

    tol = 1e-12 if prec == "f64" else 2e-6
    iv = float(grd.invVOL)
    dx = BOX[0] / GRID[0]
    xn = (np.arange(grd.nxn, dtype=np.float64) - 1.0) * dx
    dt = np.float64 if prec == "f64" else np.float32
    for s in range(ns):
        out = S.DeviceParticles(counts[s], species[s].qom, prec)
        for d in out.d.values():
            d.upload(np.full(counts[s], np.nan, dtype=dt))
        eng.store(s, out)
        q = out.d["q"].download()
        assert np.array_equal(q, q0[s]), f"s{s}: q not returned at every load index"
        x = out.d["x"].download().astype(np.float64)
        u = out.d["u"].download().astype(np.float64)
        for k in ("y", "z", "v", "w"):
            assert np.all(np.isfinite(out.d[k].download())), f"s{s} {k}: slot not written"
        assert np.all(np.isfinite(x)) and np.all(np.isfinite(u))
        for d in out.d.values():
            d.free()
        qd = q.astype(np.float64)
        m = moms[s].download()
        _close(float(m["rhon"].sum()) / iv, float(qd.sum()), float(np.abs(qd).sum()), tol,
               f"s{s} charge")
        qu = qd * u
        _close(float(m["Jx"].sum()) / iv, float(qu.sum()), float(np.abs(qu).sum()), tol,
               f"s{s} Jx")
        quu = qu * u
        _close(float(m["pxx"].sum()) / iv, float(quu.sum()), float(np.abs(quu).sum()), tol,
               f"s{s} pxx")
        rx = m["rhon"].reshape(grd.nxn, grd.nyn, grd.nzn).sum(axis=(1, 2))
        qx = qd * x
        _close(float((xn * rx).sum()) / iv, float(qx.sum()), float(np.abs(qx).sum()), tol,
               f"s{s} linear precision")
    eng.close()
