"""Two particles per thread and batch (spic_engine_opts.particles_per_thread
= 2, FP32 fixed point) on the GPU.

A particle's push does not depend on how a tile's particles are batched, so
an engine with two particles per thread gives the particles of the engine
with one, bit for bit in load order, after phases with re-sorts between
them.  The moments differ only in the order of the FP32 register sums:
within 1e-5 of each moment's largest magnitude (the FP32 form's bound).  A
crowded cell fills the doubled buckets and sends the rest through the list
ring (and, once the ring is full, the direct deposit), with the same result.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S
from helpers import assert_close

pytestmark = pytest.mark.gpu

CFG = dict(nxc=16, nyc=8, nzc=8, npcelx=[2] * 4, npcely=[2] * 4, npcelz=[2] * 4)


@pytest.fixture(autouse=True)
def _modes():
    S.set_async(False)
    S.set_strict_fp(False)
    yield


def _run(grd, fields, arrays, qoms, param, ppt, tile, phases):
    field = S.EMField(grd, fields)
    parts = [S.Particles(p, qom, "f32") for p, qom in zip(arrays, qoms)]
    moms = [S.InterpDensSpecies(grd) for _ in parts]
    eng = S.Engine(grd, len(parts), tile=tile, set_moments=True, particles_per_thread=ppt)
    try:
        for s, p in enumerate(parts):
            eng.load(s, p)
        for k in range(phases):
            if k:
                eng.sort()
            eng.phase(field, moms, param)
        stats = eng.stats()
        for s, p in enumerate(parts):
            eng.store(s, p)
    finally:
        eng.close()
    assert stats["positions"] == S.SPIC_POS_FIXED
    return [p.download() for p in parts], [m.download() for m in moms]


def _compare(a, b):
    (pa, ma), (pb, mb) = a, b
    for s in range(len(pa)):
        for k in O.PARTICLE_NAMES:
            assert np.array_equal(pa[s][k], pb[s][k]), (s, k)
        for n in O.MOMENT_NAMES:
            assert_close(mb[s][n], ma[s][n], 1e-5, f"s{s} {n}")


@pytest.mark.parametrize("tile", [(4, 4, 4), (8, 4, 4)])
def test_two_particles_per_thread_move_the_same_particles(tile):
    st = O.State(O.GemConfig(**CFG), "f32")
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f32")
    fields = {n: st.fields[n].astype(np.float32) for n in O.FIELD_NAMES}
    param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                         PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
    args = (grd, fields, st.parts, c.qom, param)
    _compare(_run(*args, ppt=1, tile=tile, phases=3), _run(*args, ppt=2, tile=tile, phases=3))


def test_two_particles_per_thread_crowded_cell():
    """6000 particles in one cell of the first tile, moving within the tile
    (no fields): per batch of 512 the cell's bucket takes 14, the list ring
    128 and the rest deposit directly."""
    cfg = O.GemConfig(**CFG)
    dx = cfg.Lx / cfg.nxc
    rng = np.random.default_rng(11)
    n = 6000
    p = {"x": (1.0 + rng.uniform(0.1, 0.9, n)) * dx, "y": (1.0 + rng.uniform(0.1, 0.9, n)) * dx,
         "z": (1.0 + rng.uniform(0.1, 0.9, n)) * dx,
         "u": rng.normal(0.0, 0.2 * dx / cfg.dt, n), "v": rng.normal(0.0, 0.2 * dx / cfg.dt, n),
         "w": rng.normal(0.0, 0.2 * dx / cfg.dt, n), "q": np.full(n, -1e-3)}
    grd = S.Grid(cfg.nxc, cfg.nyc, cfg.nzc, cfg.Lx, cfg.Ly, cfg.Lz, "f32")
    zero = np.zeros(grd.nn, np.float32)
    fields = {f: zero for f in O.FIELD_NAMES}
    param = S.Parameters(dt=cfg.dt, c=cfg.c, NiterMover=cfg.NiterMover)
    args = (grd, fields, [p], [-64.0], param)
    _compare(_run(*args, ppt=1, tile=(4, 4, 4), phases=2),
             _run(*args, ppt=2, tile=(4, 4, 4), phases=2))
