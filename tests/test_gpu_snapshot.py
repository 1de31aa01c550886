"""GPU snapshots (spic_engine_snapshot_be32, spic_snapshot_be32 through the
C ABI): the big-endian float32 datasets of HDF5_Write_Particles
(src/RW_IO.cpp:376-460), in load order, equal the oracle's particles after
the same phases rounded to float and byte-swapped, bit for bit (strict
arithmetic keeps the particles bit-identical).  The copy runs on a side
stream: a phase enqueued after the snapshot does not change it."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S
from helpers import PRECS
from sputnipic_amd.snapshot import DATASETS, Snapshot

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _modes():
    S.set_async(False)
    S.set_strict_fp(True)
    yield
    S.set_async(False)
    S.set_strict_fp(False)


def _setup(prec):
    st = O.State(O.GemConfig(), prec)
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, prec)
    field = S.EMField(grd, {n: st.fields[n] for n in O.FIELD_NAMES})
    parts = [S.Particles(p, c.qom[s], prec, c.NiterMover) for s, p in enumerate(st.parts)]
    ids = [S.InterpDensSpecies(grd) for _ in range(c.ns)]
    param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                         PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
    return st, grd, field, parts, ids, param


def _check(got, st, what):
    for s, arrs in enumerate(got):
        for d in DATASETS:
            want = np.asarray(st.parts[s][d], dtype=np.float64).astype(">f4")
            assert np.array_equal(arrs[d].view(">u4"), want.view(">u4")), f"{what} s{s} {d}"


@pytest.mark.parametrize("prec", PRECS)
def test_engine_snapshot_matches_oracle(prec):
    st, grd, field, parts, ids, param = _setup(prec)
    c = st.cfg
    eng = S.Engine(grd, c.ns, set_moments=True)
    for s, p in enumerate(parts):
        eng.load(s, p)
    stream = S.Stream()
    for _ in range(2):
        eng.phase(field, ids, param, stream)
    snap = Snapshot([len(p["x"]) for p in st.parts])
    S.set_async(True)
    snap.capture_engine(eng, stream)
    eng.phase(field, ids, param, stream)  # runs while the copy is in flight
    got = snap.result()
    stream.synchronize()
    S.set_async(False)
    S.check_domain(stream)
    for _ in range(2):
        st.cycle()
    _check(got, st, "engine")
    snap.close()
    eng.close()


@pytest.mark.parametrize("prec", PRECS)
def test_caller_order_snapshot_matches_oracle(prec):
    st, grd, field, parts, ids, param = _setup(prec)
    for s, p in enumerate(parts):
        S.mover_PC(p, field, grd, param)
    for _ in range(1):
        for s in range(st.cfg.ns):
            st.mover_pc(s)
    snap = Snapshot([p.nop for p in parts])
    snap.capture_species(grd, parts)
    _check(snap.result(), st, "caller order")
    snap.close()
