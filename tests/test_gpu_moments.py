"""GPU parity of the steps after the deposit: the ghost-node fold
(spic_apply_bc_ids, applyBCids src/BC.cpp:362-619) and the species sum
(spic_sum_species, sumOverSpecies src/InterpDensNet.cpp:120-144), through
the C ABI.

Tolerances: both work on FP64 moment arrays with the reference's additions
in the reference's order, so FP64 results are bit-identical to the oracle's
(and the reference's).  Against the FP32 reference fixtures, whose fold
rounds every addition to float, the FP64 device result agrees to 1e-6 of
each array's largest magnitude.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S
from helpers import CASES, assert_bitwise, assert_close, load_case, oracle_from_fixture

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _modes():
    S.set_async(False)
    S.set_strict_fp(False)
    yield
    S.set_strict_fp(False)


def _state(per=(1, 0, 1), seed=0, shape=(6, 5, 4)):
    cfg = O.GemConfig(nxc=shape[0], nyc=shape[1], nzc=shape[2], npcelx=[1] * 4,
                      npcely=[1] * 4, npcelz=[1] * 4, PERIODICX=bool(per[0]),
                      PERIODICY=bool(per[1]), PERIODICZ=bool(per[2]))
    st = O.State(cfg, "f64", build_ic=False)
    rng = np.random.default_rng(seed)
    for s in range(cfg.ns):
        for n in O.MOMENT_NAMES:
            st.moments[s][n][:] = rng.normal(0, 1, st.nn)
    return st


def _device(st):
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f64")
    ids = [S.InterpDensSpecies(grd) for _ in range(c.ns)]
    for s in range(c.ns):
        ids[s].upload(st.moments[s])
    param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                         PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
    return grd, ids, param


@pytest.mark.parametrize("per", [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)])
def test_fold_matches_oracle_bitwise(per):
    st = _state(per, seed=7 + 4 * per[0] + 2 * per[1] + per[2])
    grd, ids, param = _device(st)
    S.applyBCids(ids[0], grd, param)        # one species, the reference's call
    S.applyBCids_all(ids[1:], grd, param)   # the others in one call
    for s in range(st.cfg.ns):
        st.apply_bc_ids(s)
        got = ids[s].download()
        for n in O.MOMENT_NAMES:
            assert_bitwise(got[n], st.moments[s][n], f"{per} s{s} {n}")


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name", CASES)
def test_fold_matches_reference_fixture(name, prec):
    """The reference's own applyBCids output after cycle 1 (stage bc1)."""
    data = load_case(name, prec)
    st = oracle_from_fixture(name, prec, data)
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f64")
    param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                         PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
    ids = [S.InterpDensSpecies(grd) for _ in range(c.ns)]
    for s in range(c.ns):
        ids[s].upload({n: data[f"c1_s{s}_{n}"].astype(np.float64) for n in O.MOMENT_NAMES})
    S.applyBCids_all(ids, grd, param)
    for s in range(c.ns):
        got = ids[s].download()
        for n in O.MOMENT_NAMES:
            want = data[f"bc1_s{s}_{n}"]
            if prec == "f64":
                assert_bitwise(got[n], want, f"bc1 s{s} {n}")
            else:
                assert_close(got[n], want, 1e-6, f"bc1 s{s} {n}")


def test_species_sum_matches_oracle_bitwise():
    st = _state(seed=21, shape=(10, 6, 7))
    grd, ids, param = _device(st)
    net = S.InterpDensNet(grd)
    start = {n: np.linspace(-1, 1, st.nn) * (k + 1) for k, n in enumerate(O.MOMENT_NAMES)}
    net.upload(start)
    S.sumOverSpecies(net, ids, grd, st.cfg.ns)
    want = st.sum_species(start)
    got = net.download()
    for n in O.MOMENT_NAMES:
        assert_bitwise(got[n], want[n], n)


def test_species_sum_edge_counts():
    """One species (ns = 1) passes through; ns outside 1..8 is rejected."""
    st = _state(seed=5)
    grd, ids, param = _device(st)
    net = S.InterpDensNet(grd)
    S.sumOverSpecies(net, ids, grd, 1)
    got = net.download()
    for n in O.MOMENT_NAMES:
        assert_bitwise(got[n], st.moments[0][n], n)
    with pytest.raises(S.SpicError) as ei:
        S.sumOverSpecies(net, ids, grd, 0)
    assert ei.value.code == S.SPIC_EINVAL


def test_cycle_with_fold_and_sum_matches_oracle():
    """The reference's whole particle step on the small Harris case
    (src/sputniPIC_CPU.cpp:232-259): the engine phase (mover_PC and
    interpP2G of every species), then applyBCids per species, then
    sumOverSpecies, against the oracle, in strict arithmetic: the moments
    differ from the oracle's only by the deposit's summation order.  The net
    charge density nearly cancels (electrons against ions), so the error is
    measured against the largest species magnitude of each moment: 1e-13."""
    S.set_strict_fp(True)
    cfg = O.GemConfig()
    st = O.State(cfg, "f64")
    grd = S.Grid(cfg.nxc, cfg.nyc, cfg.nzc, cfg.Lx, cfg.Ly, cfg.Lz, "f64")
    field = S.EMField(grd, {n: st.fields[n] for n in O.FIELD_NAMES})
    parts = [S.Particles(p, cfg.qom[s], "f64", cfg.NiterMover) for s, p in enumerate(st.parts)]
    ids = [S.InterpDensSpecies(grd) for _ in range(cfg.ns)]
    param = S.Parameters(dt=cfg.dt, c=cfg.c, NiterMover=cfg.NiterMover,
                         PERIODICX=cfg.PERIODICX, PERIODICY=cfg.PERIODICY,
                         PERIODICZ=cfg.PERIODICZ)
    eng = S.Engine(grd, cfg.ns)
    for s, p in enumerate(parts):
        eng.load(s, p)
    S.setZeroSpeciesDensities(ids, grd, cfg.ns)
    eng.phase(field, ids, param)
    S.applyBCids_all(ids, grd, param)
    net = S.InterpDensNet(grd)
    S.sumOverSpecies(net, ids, grd, cfg.ns)
    eng.close()
    st.cycle()
    for s in range(cfg.ns):
        st.apply_bc_ids(s)
    want = st.sum_species()
    got = net.download()
    for n in O.MOMENT_NAMES:
        scale = max(float(np.abs(st.moments[s][n]).max()) for s in range(cfg.ns))
        err = float(np.abs(got[n] - want[n]).max()) / scale
        assert err <= 1e-13, f"{n}: error {err:.2e} of the largest species magnitude"
