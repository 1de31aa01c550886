"""HIP-graph capture of the particle phase (spic_graph_*): two cycles
(zero the moments, engine phase) captured once and replayed, against the
same cycles run eagerly.

Each replay is checked on its own, with eager work and a synchronization
between replays.  Tolerances: particle state bit for bit (strict
arithmetic); node moments within 1e-13 of each array's max magnitude (the
deposit's summation order is not fixed from run to run).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S
from helpers import assert_bitwise, assert_close

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _modes():
    S.set_async(False)
    S.set_strict_fp(True)
    yield
    S.set_async(False)
    S.set_strict_fp(False)


def _setup(st):
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f64")
    field = S.EMField(grd, {n: st.fields[n] for n in O.FIELD_NAMES})
    parts = [S.Particles(p, c.qom[s], "f64", c.NiterMover) for s, p in enumerate(st.parts)]
    ids = [S.InterpDensSpecies(grd) for _ in range(c.ns)]
    param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                         PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
    eng = S.Engine(grd, c.ns)
    for s, p in enumerate(parts):
        eng.load(s, p)
    return grd, field, parts, ids, param, eng


def test_graph_replays_match_eager_cycles():
    st = O.State(O.GemConfig(), "f64")
    c = st.cfg
    # eager: four cycles, keeping the state after cycles 2 and 4
    grd, field, parts, ids, param, eng = _setup(st)
    want = []
    for cyc in range(4):
        S.setZeroSpeciesDensities(ids, grd, c.ns)
        eng.phase(field, ids, param)
        if cyc in (1, 3):
            for s, p in enumerate(parts):
                eng.store(s, p)
            want.append(([p.download() for p in parts], [m.download() for m in ids]))
    eng.close()

    # captured: two cycles in one graph, replayed twice
    grd, field, parts, ids, param, eng = _setup(st)
    stream = S.Stream()
    with S.Graph(stream) as g:
        for _ in range(2):
            S.setZeroSpeciesDensities(ids, grd, c.ns, stream)
            eng.phase(field, ids, param, stream)
    for rep in range(2):
        g.launch()
        stream.synchronize()
        S.set_async(False)
        S.check_domain(stream)
        for s, p in enumerate(parts):
            eng.store(s, p, stream)
        got_p = [p.download() for p in parts]
        got_m = [m.download() for m in ids]
        want_p, want_m = want[rep]
        for s in range(c.ns):
            for k in O.PARTICLE_NAMES:
                assert_bitwise(got_p[s][k], want_p[s][k], f"replay {rep} s{s} {k}")
            for n in O.MOMENT_NAMES:
                assert_close(got_m[s][n], want_m[s][n], 1e-13, f"replay {rep} s{s} {n}")
        S.set_async(True)
    S.set_async(False)
    g.close()
    eng.close()
