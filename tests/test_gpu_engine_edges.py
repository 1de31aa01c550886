"""GPU edge cases of the tile-resident engine's storage and moment contract.

* A device allocation that fails during a load leaves the species without
  storage: the load reports SPIC_ENOMEM, the engine's device byte count
  returns to what it held before the species was first loaded plus the
  sort workspace that load left, and the next load (smaller than the failed
  one, so that a partial set of buffers would otherwise be reused)
  allocates every buffer again; its phase then matches the oracle bit for
  bit (strict arithmetic).  The failure is injected with
  spic_test_fail_engine_alloc at each allocation of the load in turn.
* With set_moments, a phase writes the moments of every loaded species,
  including one reloaded without particles: its moments come back zero
  (they start as garbage), for a single engine and for a batch.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S
from sputnipic_amd._capi import SPIC_ENOMEM, SpicError, lib

pytestmark = pytest.mark.gpu

CFG = dict(nxc=16, nyc=8, nzc=8, npcelx=[2] * 4, npcely=[2] * 4, npcelz=[2] * 4)


@pytest.fixture(autouse=True)
def _modes():
    S.set_async(False)
    S.set_strict_fp(True)
    yield
    lib().spic_test_fail_engine_alloc(-1)
    S.set_strict_fp(False)


def _setup(prec="f64"):
    st = O.State(O.GemConfig(**CFG), prec)
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, prec)
    field = S.EMField(grd, st.fields)
    param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                         PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
    return st, grd, field, param


def _garbage(ids, grd, seed):
    rng = np.random.default_rng(seed)
    for m in ids:
        m.upload({n: rng.normal(0, 1e3, grd.nn) for n in O.MOMENT_NAMES})


# eight allocations per load that (re)allocates: x..q block, ids, counters
# and segment table of each of the two buffers; then the sort workspace
@pytest.mark.parametrize("after", [0, 1, 3, 4, 5, 7, 8])
def test_failed_load_then_smaller_load_allocates_every_buffer(after):
    st, grd, field, param = _setup()
    c = st.cfg
    eng = S.Engine(grd, 1, set_moments=True)
    try:
        base = eng.stats()["device_bytes"]
        small = {k: v[: len(v) // 2].copy() for k, v in st.parts[0].items()}
        big = {k: np.concatenate([v, v]) for k, v in st.parts[0].items()}
        eng.load(0, S.Particles(small, c.qom[0], "f64"))
        lib().spic_test_fail_engine_alloc(after)
        with pytest.raises(SpicError) as exc:
            eng.load(0, S.Particles(big, c.qom[0], "f64"))
        assert exc.value.code == SPIC_ENOMEM
        lib().spic_test_fail_engine_alloc(-1)
        if after < 8:
            # the sort workspace of the first load stays; nothing else does
            got = eng.stats()
            assert got["device_bytes"] == base + got["workspace_bytes"], \
                "a failed load kept device memory"
            assert got["workspace_bytes"] > 0
        # (after == 8: the sort workspace failed; the storage, complete, is kept)
        # a load smaller than the failed one: without the fix it would reuse
        # the partly allocated storage
        part = S.Particles(small, c.qom[0], "f64")
        eng.load(0, part)
        ids = [S.InterpDensSpecies(grd)]
        _garbage(ids, grd, 3)
        eng.phase(field, ids, param)
        eng.store(0, part)
        got_p, got_m = part.download(), ids[0].download()
    finally:
        eng.close()
    st.parts[0] = small
    st.mover_pc(0)
    st.zero_moments(0)
    st.interp_p2g(0)
    for k in O.PARTICLE_NAMES:
        assert np.array_equal(got_p[k], st.parts[0][k]), k
    for n in O.MOMENT_NAMES:
        want = st.moments[0][n]
        err = np.max(np.abs(got_m[n] - want)) / np.max(np.abs(want))
        assert err <= 1e-13, (n, err)


def _empty_like(parts):
    return {k: np.zeros(0, v.dtype) for k, v in parts.items()}


def test_set_moments_zeroes_a_species_reloaded_empty():
    st, grd, field, param = _setup()
    c = st.cfg
    eng = S.Engine(grd, 2, set_moments=True)
    try:
        eng.load(0, S.Particles(st.parts[0], c.qom[0], "f64"))
        eng.load(1, S.Particles(st.parts[1], c.qom[1], "f64"))
        ids = [S.InterpDensSpecies(grd) for _ in range(2)]
        eng.phase(field, ids, param)
        assert np.any(ids[1].download()["rhon"] != 0)
        eng.load(1, S.Particles(_empty_like(st.parts[1]), c.qom[1], "f64"))
        _garbage(ids, grd, 5)
        eng.phase(field, ids, param)
        m1 = ids[1].download()
        for n in O.MOMENT_NAMES:
            assert np.all(m1[n] == 0.0), f"{n} of the empty species not zeroed"
        # species 0 phased twice, its moments those of its second phase
        m0 = ids[0].download()
    finally:
        eng.close()
    for _ in range(2):
        st.mover_pc(0)
    st.zero_moments(0)
    st.interp_p2g(0)
    for n in O.MOMENT_NAMES:
        want = st.moments[0][n]
        err = np.max(np.abs(m0[n] - want)) / np.max(np.abs(want))
        assert err <= 1e-13, (n, err)


@pytest.mark.parametrize("all_empty", [False, True])
def test_batch_set_moments_zeroes_empty_species(all_empty):
    st, grd, field, param = _setup()
    c = st.cfg
    engs, moms = [], []
    try:
        for i in range(2):
            eng = S.Engine(grd, 2)
            p0 = st.parts[0] if (i == 0 and not all_empty) else _empty_like(st.parts[0])
            eng.load(0, S.Particles(p0, c.qom[0], "f64"))
            eng.load(1, S.Particles(_empty_like(st.parts[1]), c.qom[1], "f64"))
            engs.append(eng)
            ids = [S.InterpDensSpecies(grd) for _ in range(2)]
            _garbage(ids, grd, 7 + i)
            moms.append(ids)
        batch = S.Batch([(e, field, m, param) for e, m in zip(engs, moms)], set_moments=True)
        try:
            batch.phase()
        finally:
            batch.close()
        for i in range(2):
            for s in range(2):
                m = moms[i][s].download()
                if i == 0 and s == 0 and not all_empty:
                    assert np.any(m["rhon"] != 0)
                    continue
                for n in O.MOMENT_NAMES:
                    assert np.all(m[n] == 0.0), f"member {i} species {s} {n} not zeroed"
    finally:
        for e in engs:
            e.close()
