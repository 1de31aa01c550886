"""GPU parity of batches (spic_batch_* through the C ABI): independent
Harris-sheet instances with different grids, species counts, particle
counts and perturbation amplitudes, advanced together with one launch per
kernel and phase.

Each instance's results are the oracle's for that instance run alone.
Tolerances, as for the engine on its own: in strict arithmetic the particle
state is bit-identical after every phase, whatever the sort schedule; node
moments differ only in summation order: 1e-13 of each array's max magnitude
in FP64, 1e-5 in FP32 (the FP32-form bound).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S
from helpers import PRECS, assert_bitwise, assert_close, load_case, oracle_from_fixture

pytestmark = pytest.mark.gpu

# grids with partial tiles, 1..4 species, 1..12 particles per cell
INSTANCES = [
    dict(nxc=8, nyc=6, nzc=4, ppc=(2, 2, 2), ns=2, pertX=0.01),
    dict(nxc=12, nyc=8, nzc=8, ppc=(1, 2, 2), ns=4, pertX=0.05),
    dict(nxc=16, nyc=4, nzc=10, ppc=(3, 2, 2), ns=3, pertX=0.4),
    dict(nxc=4, nyc=4, nzc=4, ppc=(1, 1, 1), ns=1, pertX=0.005),
]


@pytest.fixture(autouse=True)
def _modes():
    S.set_async(False)
    S.set_strict_fp(True)
    yield
    S.set_async(False)
    S.set_strict_fp(False)


def _oracle(spec, prec):
    p = spec["ppc"]
    cfg = O.GemConfig(nxc=spec["nxc"], nyc=spec["nyc"], nzc=spec["nzc"], ns=spec["ns"],
                      npcelx=[p[0]] * 4, npcely=[p[1]] * 4, npcelz=[p[2]] * 4,
                      pertX=spec["pertX"])
    return O.State(cfg, prec)


class Instance:
    def __init__(self, st, prec, tile=(4, 4, 4), lds_partials=0, sort_mode=0):
        c = st.cfg
        dt = np.float64 if prec == "f64" else np.float32
        self.grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, prec)
        self.field = S.EMField(self.grd, {n: st.fields[n].astype(dt) for n in O.FIELD_NAMES})
        self.parts = [S.Particles(p, c.qom[s], prec, c.NiterMover)
                      for s, p in enumerate(st.parts)]
        self.ids = [S.InterpDensSpecies(self.grd) for _ in range(c.ns)]
        self.param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover,
                                  PERIODICX=c.PERIODICX, PERIODICY=c.PERIODICY,
                                  PERIODICZ=c.PERIODICZ)
        self.eng = S.Engine(self.grd, c.ns, tile=tile, lds_partials=lds_partials,
                            sort_mode=sort_mode)
        for s, p in enumerate(self.parts):
            self.eng.load(s, p)
        self.ns = c.ns

    def member(self):
        return (self.eng, self.field, self.ids, self.param)

    def zero(self, stream=None):
        S.setZeroSpeciesDensities(self.ids, self.grd, self.ns, stream)

    def results(self, stream=None):
        for s, p in enumerate(self.parts):
            self.eng.store(s, p, stream)
        return [p.download() for p in self.parts], [m.download() for m in self.ids]


def _check(inst_results, states, prec, what):
    tol = 1e-13 if prec == "f64" else 1e-5
    for i, ((got_p, got_m), st) in enumerate(zip(inst_results, states)):
        for s in range(st.cfg.ns):
            for k in O.PARTICLE_NAMES:
                assert_bitwise(got_p[s][k], st.parts[s][k], f"{what} inst {i} s{s} {k}")
            for n in O.MOMENT_NAMES:
                assert_close(got_m[s][n], st.moments[s][n], tol, f"{what} inst {i} s{s} {n}")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("tile,sort_every", [((4, 4, 4), 0), ((4, 4, 4), 2), ((8, 4, 4), 1)])
def test_batch_matches_oracle_per_instance(prec, tile, sort_every):
    states = [_oracle(spec, prec) for spec in INSTANCES]
    insts = [Instance(st, prec, tile) for st in states]
    batch = S.Batch([x.member() for x in insts], sort_every=sort_every)
    phases = 3
    for _ in range(phases):
        for x in insts:
            x.zero()
        batch.phase()
    for st in states:
        for _ in range(phases):
            st.cycle()
    _check([x.results() for x in insts], states, prec, f"{tile} sort {sort_every}")
    for x in insts:
        stats = x.eng.stats()
        assert stats["phases"] == phases
        if sort_every:
            assert stats["sorts"] == x.ns * (1 + (phases - 1) // sort_every)
    batch.close()
    for x in insts:
        x.eng.close()


@pytest.mark.parametrize("prec", PRECS)
def test_batch_radix_sort_matches_oracle(prec):
    """Members whose re-sorts are radix sorts, re-sorted every phase for
    six phases (the other batch tests re-sort by tile sort, the default)."""
    states = [_oracle(spec, prec) for spec in INSTANCES]
    insts = [Instance(st, prec, sort_mode=S.SPIC_SORT_RADIX) for st in states]
    batch = S.Batch([x.member() for x in insts], sort_every=1, set_moments=True)
    for _ in range(6):
        batch.phase()
    for st in states:
        for _ in range(6):
            st.cycle()
    _check([x.results() for x in insts], states, prec, "sort mode 1")
    batch.close()
    for x in insts:
        x.eng.close()


@pytest.mark.parametrize("prec", PRECS)
def test_batch_set_moments_overwrites(prec):
    """With set_moments the phase writes the moments: started from garbage
    and never zeroed, every phase's moments are that phase's deposit."""
    states = [_oracle(spec, prec) for spec in INSTANCES]
    insts = [Instance(st, prec) for st in states]
    rng = np.random.default_rng(5)
    for x in insts:
        for m in x.ids:
            m.upload({n: rng.normal(0, 1e3, x.grd.nn) for n in O.MOMENT_NAMES})
    batch = S.Batch([x.member() for x in insts], sort_every=2, set_moments=True)
    for _ in range(3):
        batch.phase()
    for st in states:
        for _ in range(3):
            st.cycle()
    _check([x.results() for x in insts], states, prec, "set_moments")
    batch.close()
    for x in insts:
        x.eng.close()


@pytest.mark.parametrize("prec", PRECS)
def test_batch_set_moments_keeps_global_deposits(prec):
    """An instance whose particles spill into the overflow region (gem_hot,
    no fields, species 0's last cell layer of tile 0 drifting +0.9 cell per
    step into tile 1) beside a plain one: the overflow pass deposits with
    global atomics, which the written moments must include."""
    data = load_case("gem_hot", prec)
    hot = oracle_from_fixture("gem_hot", prec, data)
    for n in O.FIELD_NAMES:
        hot.fields[n][:] = 0.0
    c = hot.cfg
    dx = c.Lx / c.nxc
    for s_, p in enumerate(hot.parts):
        for k in ("u", "v", "w"):
            p[k][:] = 0.0
        if s_ == 0:
            last = (p["x"] >= 3 * dx) & (p["x"] < 4 * dx)
            p["u"][last] = 0.9 * dx / c.dt
    states = [hot, _oracle(INSTANCES[1], prec)]
    insts = [Instance(st, prec) for st in states]
    batch = S.Batch([x.member() for x in insts], set_moments=True)
    for _ in range(3):
        batch.phase()
    for st in states:
        for _ in range(3):
            st.cycle()
    assert insts[0].eng.stats()["global_deposits"] > 0
    _check([x.results() for x in insts], states, prec, "set_moments overflow")
    batch.close()
    for x in insts:
        x.eng.close()


@pytest.mark.parametrize("prec", PRECS)
def test_batch_members_loaded_again_rebuild_the_tables(prec):
    """Three phases (re-sorted every second phase), then every member's
    species are loaded again -- member 0 from a denser initial condition,
    the others from their own current particles -- and the batch rebuilds
    its phase and sort tables at the next phase; three more phases match
    the oracle."""
    states = [_oracle(spec, prec) for spec in INSTANCES]
    insts = [Instance(st, prec) for st in states]
    batch = S.Batch([x.member() for x in insts], sort_every=2, set_moments=True)
    for _ in range(3):
        batch.phase()
    for st in states:
        for _ in range(3):
            st.cycle()
    _check([x.results() for x in insts], states, prec, "before the loads")
    dense = dict(INSTANCES[0], ppc=(3, 3, 3))
    states[0] = _oracle(dense, prec)
    c = states[0].cfg
    insts[0].parts = [S.Particles(p, c.qom[s_], prec, c.NiterMover)
                      for s_, p in enumerate(states[0].parts)]
    for x in insts:
        for s_, p in enumerate(x.parts):
            x.eng.load(s_, p)
    for _ in range(3):
        batch.phase()
    for st in states:
        for _ in range(3):
            st.cycle()
    _check([x.results() for x in insts], states, prec, "after the loads")
    batch.close()
    for x in insts:
        x.eng.close()


def test_batch_graph_replays_match_oracle():
    """Two phases (zero, batch phase) captured in one graph and replayed
    twice, each replay checked against the oracle after it."""
    states = [_oracle(spec, "f64") for spec in INSTANCES]
    insts = [Instance(st, "f64") for st in states]
    batch = S.Batch([x.member() for x in insts], sort_every=3)
    stream = S.Stream()
    with S.Graph(stream) as g:
        for _ in range(2):
            for x in insts:
                x.zero(stream)
            batch.phase(stream)
    for rep in range(2):
        g.launch()
        stream.synchronize()
        S.set_async(False)
        S.check_domain(stream)
        for st in states:
            for _ in range(2):
                st.cycle()
        _check([x.results(stream) for x in insts], states, "f64", f"replay {rep}")
        S.set_async(True)
    S.set_async(False)
    g.close()
    batch.close()
    for x in insts:
        x.eng.close()


def test_batch_rejects_bad_members():
    st64 = _oracle(INSTANCES[0], "f64")
    st32 = _oracle(INSTANCES[3], "f32")
    a, b = Instance(st64, "f64"), Instance(st32, "f32")
    with pytest.raises(S.SpicError) as ei:
        S.Batch([a.member(), b.member()])
    assert ei.value.code == S.SPIC_EINVAL
    with pytest.raises(S.SpicError) as ei:
        S.Batch([a.member(), a.member()])
    assert ei.value.code == S.SPIC_EINVAL
    c = Instance(_oracle(INSTANCES[1], "f64"), "f64", tile=(8, 4, 4))
    with pytest.raises(S.SpicError) as ei:
        S.Batch([a.member(), c.member()])
    assert ei.value.code == S.SPIC_EINVAL
    e = Instance(_oracle(INSTANCES[1], "f64"), "f64", sort_mode=S.SPIC_SORT_RADIX)
    with pytest.raises(S.SpicError) as ei:
        S.Batch([a.member(), e.member()])
    assert ei.value.code == S.SPIC_EINVAL
    # a member advanced on its own falls out of step with the others
    d = Instance(_oracle(INSTANCES[2], "f64"), "f64")
    batch = S.Batch([a.member(), d.member()])
    d.zero()
    d.eng.phase(d.field, d.ids, d.param)
    with pytest.raises(S.SpicError) as ei:
        batch.phase()
    assert ei.value.code == S.SPIC_EINVAL
    batch.close()
    for x in (a, b, c, d, e):
        x.eng.close()


def test_autotuner_times_every_candidate_and_picks_the_fastest():
    """tune_batch builds the instances once per tile shape, times every
    sort schedule, returns the fastest, and leaves the caller's state
    alone (it advances only its own instances)."""
    from sputnipic_amd.tune import tune_batch
    states = [_oracle(spec, "f32") for spec in INSTANCES]
    built = []

    def make(tile):
        insts = [Instance(st, "f32", tile) for st in states]
        for x in insts:
            x.close = x.eng.close
        built.append(tile)
        return insts

    res = tune_batch(make, tiles=((4, 4, 4), (8, 4, 4)), sort_choices=(0, 2), phases=2,
                     warmup=1, set_moments=True)
    assert built == [(4, 4, 4), (8, 4, 4)]
    assert len(res.trials) == 4
    best = min(res.trials, key=lambda t: t.ms_per_phase)
    assert (res.tile, res.sort_every, res.ms_per_phase) == (best.tile, best.sort_every,
                                                            best.ms_per_phase)
    assert all(t.ms_per_phase > 0 for t in res.trials)
    d = res.as_dict()
    assert d["tile"] in ("4x4x4", "8x4x4") and len(d["trials"]) == 4


def test_engine_autotuner_times_every_candidate_and_picks_the_fastest():
    """tune_engine builds a fresh engine per (tile, sort interval), times
    every one, closes each, and returns the fastest; the engines it times
    advance the same particles as an engine outside it (checked on the last
    candidate's phases against the oracle)."""
    from sputnipic_amd.tune import tune_engine
    st = _oracle(INSTANCES[0], "f64")
    c = st.cfg
    built, closed, phases = [], [], {}

    class Runner:
        def __init__(self, tile, se):
            self.key = (tile, se)
            built.append(self.key)
            self.grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f64")
            self.field = S.EMField(self.grd, st.fields)
            self.parts = [S.Particles(p, c.qom[s], "f64", c.NiterMover)
                          for s, p in enumerate(st.parts)]
            self.ids = [S.InterpDensSpecies(self.grd) for _ in range(c.ns)]
            self.param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover,
                                      PERIODICX=c.PERIODICX, PERIODICY=c.PERIODICY,
                                      PERIODICZ=c.PERIODICZ)
            self.eng = S.Engine(self.grd, c.ns, tile=tile, sort_every=se, set_moments=True)
            for s, p in enumerate(self.parts):
                self.eng.load(s, p)

        def phase(self, stream):
            self.eng.phase(self.field, self.ids, self.param, stream)
            phases[self.key] = phases.get(self.key, 0) + 1

        def close(self):
            for s, p in enumerate(self.parts):
                self.eng.store(s, p)
            self.out = [p.download() for p in self.parts]
            self.eng.close()
            closed.append(self.key)
            last[0] = self

    last = [None]
    res = tune_engine(lambda tile, se: Runner(tile, se), tiles=((4, 4, 4), (8, 4, 4)),
                      sort_choices=(0, 2), phases=3, warmup=1)
    assert built == closed == [((4, 4, 4), 0), ((4, 4, 4), 2), ((8, 4, 4), 0),
                               ((8, 4, 4), 2)]
    assert all(n == 4 for n in phases.values())
    assert len(res.trials) == 4 and all(t.ms_per_phase > 0 for t in res.trials)
    best = min(res.trials, key=lambda t: t.ms_per_phase)
    assert (res.tile, res.sort_every) == (best.tile, best.sort_every)
    assert res.as_dict()["tile"] in ("4x4x4", "8x4x4")
    # the last candidate (8x4x4, a re-sort every 2 phases) ran 4 phases:
    # its particles are the oracle's after 4 cycles, within production
    # arithmetic's bounds
    for _ in range(4):
        st.cycle()
    for s in range(c.ns):
        for k in ("u", "v", "w"):
            got, want = last[0].out[s][k], st.parts[s][k]
            assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), (s, k)


def test_engine_autotuner_times_particles_per_thread():
    """With ppt_choices the tuner builds each (tile, sort interval) with one
    and with two particles per thread (an FP32 fixed-point engine), passes
    the choice to the builder, records it per trial and returns the fastest
    trial's; both choices advance the same particles, bit for bit."""
    from sputnipic_amd.tune import tune_engine
    S.set_strict_fp(False)  # fixed-point positions: production arithmetic
    st = _oracle(INSTANCES[0], "f32")
    c = st.cfg
    built, outs = [], {}

    class Runner:
        def __init__(self, tile, se, ppt):
            self.key = (tile, se, ppt)
            built.append(self.key)
            self.grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f32")
            self.field = S.EMField(self.grd, {n: st.fields[n].astype(np.float32)
                                              for n in O.FIELD_NAMES})
            self.parts = [S.Particles(p, c.qom[s], "f32", c.NiterMover)
                          for s, p in enumerate(st.parts)]
            self.ids = [S.InterpDensSpecies(self.grd) for _ in range(c.ns)]
            self.param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover,
                                      PERIODICX=c.PERIODICX, PERIODICY=c.PERIODICY,
                                      PERIODICZ=c.PERIODICZ)
            self.eng = S.Engine(self.grd, c.ns, tile=tile, sort_every=se, set_moments=True,
                                particles_per_thread=ppt)
            for s, p in enumerate(self.parts):
                self.eng.load(s, p)

        def phase(self, stream):
            self.eng.phase(self.field, self.ids, self.param, stream)

        def close(self):
            for s, p in enumerate(self.parts):
                self.eng.store(s, p)
            outs[self.key] = [p.download() for p in self.parts]
            self.eng.close()

    res = tune_engine(lambda tile, se, ppt: Runner(tile, se, ppt), tiles=((4, 4, 4),),
                      sort_choices=(0,), phases=2, warmup=1, ppt_choices=(1, 2))
    assert built == [((4, 4, 4), 0, 1), ((4, 4, 4), 0, 2)]
    assert [t.ppt for t in res.trials] == [1, 2]
    best = min(res.trials, key=lambda t: t.ms_per_phase)
    assert res.ppt == best.ppt and res.as_dict()["particles_per_thread"] == best.ppt
    a, b = outs[((4, 4, 4), 0, 1)], outs[((4, 4, 4), 0, 2)]
    for s in range(c.ns):
        for k in O.PARTICLE_NAMES:
            assert np.array_equal(a[s][k], b[s][k]), (s, k)
