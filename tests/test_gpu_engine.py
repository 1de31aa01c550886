"""GPU parity of the tile-resident engine (spic_engine_* through the C ABI):
one launch per particle phase over particles binned by tile, LDS-staged
fields, per-cell LDS buckets and register sums for the deposit, particles
re-binned into their destination tile's segment as they are written.

Tolerances: in strict arithmetic the particle state after any number of
phases is bit-identical to the oracle (and so to the reference), whatever
the tile shape, sort schedule or overflow traffic, because binning only
reorders particles and the arithmetic per particle is the reference's; node
moments then differ only in summation order: 1e-13 of each array's max
magnitude in FP64, 1e-5 in FP32 (the FP32-form bound).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S
from helpers import (CASES, PRECS, assert_bitwise, assert_close, load_case,
                     oracle_from_fixture)

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _modes():
    S.set_async(False)
    S.set_strict_fp(False)
    yield
    S.set_strict_fp(False)


_CACHE = {}


def oracle_small(prec):
    if prec not in _CACHE:
        _CACHE[prec] = O.State(O.GemConfig(), prec)
    fresh = O.State(_CACHE[prec].cfg, prec, build_ic=False)
    fresh.copy_from(_CACHE[prec])
    return fresh


def run_engine(st: O.State, prec: str, phases: int, tile=(4, 4, 4), sort_every=0,
               slack_pct=0, set_moments=False, sort_mode=0, positions=0):
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, prec)
    dt = np.float64 if prec == "f64" else np.float32
    field = S.EMField(grd, {n: st.fields[n].astype(dt) for n in O.FIELD_NAMES})
    parts = [S.Particles(p, c.qom[s], prec, c.NiterMover) for s, p in enumerate(st.parts)]
    ids = [S.InterpDensSpecies(grd) for _ in range(c.ns)]
    param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                         PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
    eng = S.Engine(grd, c.ns, tile=tile, sort_every=sort_every, slack_pct=slack_pct,
                   set_moments=set_moments, sort_mode=sort_mode, positions=positions)
    for s, p in enumerate(parts):
        eng.load(s, p)
    if set_moments:
        # the phase writes the moments: start from garbage, never zero
        rng = np.random.default_rng(11)
        for m in ids:
            m.upload({n: rng.normal(0, 1e3, grd.nn) for n in O.MOMENT_NAMES})
    for _ in range(phases):
        if not set_moments:
            S.setZeroSpeciesDensities(ids, grd, c.ns)
        eng.phase(field, ids, param)
    for s, p in enumerate(parts):
        eng.store(s, p)
    out_p = [p.download() for p in parts]
    out_m = [m.download() for m in ids]
    stats = eng.stats()
    eng.close()
    return out_p, out_m, stats


def oracle_phases(st: O.State, phases: int):
    for _ in range(phases):
        st.cycle()


@pytest.mark.parametrize("prec", PRECS)
def test_engine_strict_matches_oracle_bitwise(prec):
    st = oracle_small(prec)
    S.set_strict_fp(True)
    got_p, got_m, stats = run_engine(st, prec, phases=3)
    oracle_phases(st, 3)
    assert stats["phases"] == 3
    for s in range(st.cfg.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[s][k], st.parts[s][k], f"s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], 1e-13 if prec == "f64" else 1e-5,
                         f"s{s} {n}")


@pytest.mark.parametrize("prec", PRECS)
def test_engine_set_moments_matches_oracle(prec):
    st = oracle_small(prec)
    S.set_strict_fp(True)
    got_p, got_m, _ = run_engine(st, prec, phases=2, sort_every=1, set_moments=True)
    oracle_phases(st, 2)
    for s in range(st.cfg.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[s][k], st.parts[s][k], f"s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], 1e-13 if prec == "f64" else 1e-5,
                         f"s{s} {n}")


@pytest.mark.parametrize("tile,sort_every", [((8, 4, 4), 1), ((4, 4, 4), 2), ((4, 4, 4), 1)])
def test_engine_tile_shapes_and_sorting_do_not_change_trajectories(tile, sort_every):
    st = oracle_small("f64")
    S.set_strict_fp(True)
    got_p, got_m, stats = run_engine(st, "f64", phases=3, tile=tile, sort_every=sort_every)
    oracle_phases(st, 3)
    assert stats["tile_x"] == tile[0]
    assert stats["sorts"] >= st.cfg.ns * (1 + (2 // sort_every))
    for s in range(st.cfg.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[s][k], st.parts[s][k], f"s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], 1e-13, f"s{s} {n}")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("tile", [(4, 4, 4), (8, 4, 4)])
@pytest.mark.parametrize("sort_mode", [S.SPIC_SORT_TILE, S.SPIC_SORT_RADIX])
def test_engine_sort_modes_keep_trajectories(prec, tile, sort_mode):
    """A re-sort every phase over seven phases, tile or radix sorts:
    particles, ids and counters pass through every re-sort bit-exactly."""
    st = oracle_small(prec)
    S.set_strict_fp(True)
    got_p, got_m, stats = run_engine(st, prec, phases=7, tile=tile, sort_every=1,
                                     sort_mode=sort_mode)
    oracle_phases(st, 7)
    assert stats["sorts"] == st.cfg.ns * 7
    for s in range(st.cfg.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[s][k], st.parts[s][k], f"s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], 1e-13 if prec == "f64" else 1e-5,
                         f"s{s} {n}")


@pytest.mark.parametrize("prec", PRECS)
def test_engine_tile_sort_of_a_crowded_tile(prec):
    """One 4x4x4-cell tile holding 729 particles per cell per species
    (46,656): more than the tile sort stages in LDS (it writes each particle
    at its slot directly) and cells ranked beyond the slot tables' 512
    columns (slots from the loop over the cells); re-sorted every phase,
    the trajectories stay exact."""
    cfg = O.GemConfig(nxc=4, nyc=4, nzc=4, ns=2, npcelx=[9] * 4, npcely=[9] * 4,
                      npcelz=[9] * 4)
    st = O.State(cfg, prec)
    assert len(st.parts[0]["x"]) == 64 * 729
    S.set_strict_fp(True)
    got_p, got_m, stats = run_engine(st, prec, phases=3, sort_every=1,
                                     sort_mode=S.SPIC_SORT_TILE)
    oracle_phases(st, 3)
    assert stats["sorts"] == cfg.ns * 3
    for s in range(cfg.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[s][k], st.parts[s][k], f"s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], 1e-13 if prec == "f64" else 1e-5,
                         f"s{s} {n}")


@pytest.mark.parametrize("sort_mode", [S.SPIC_SORT_TILE, S.SPIC_SORT_RADIX])
def test_engine_sort_places_crossing_particles_in_their_tiles(sort_mode):
    """gem_hot without fields, at rest, except species 0's last cell layer
    of tile 0, which drifts +0.9 cell per step into tile 1: more particles
    than tile 1's segment has slack for, so the first phase puts some into
    the overflow region.  A sort, in either mode, moves every particle into
    its tile's segment (an empty overflow region) and the trajectories stay
    exact."""
    data = load_case("gem_hot", "f64")
    st = oracle_from_fixture("gem_hot", "f64", data)
    for n in O.FIELD_NAMES:
        st.fields[n][:] = 0.0
    c = st.cfg
    dx = c.Lx / c.nxc
    for s_, p in enumerate(st.parts):
        for k in ("u", "v", "w"):
            p[k][:] = 0.0
        if s_ == 0:
            last = (p["x"] >= 3 * dx) & (p["x"] < 4 * dx)
            p["u"][last] = 0.9 * dx / c.dt
    S.set_strict_fp(True)
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f64")
    field = S.EMField(grd, {n: st.fields[n] for n in O.FIELD_NAMES})
    parts = [S.Particles(p, c.qom[s], "f64", c.NiterMover) for s, p in enumerate(st.parts)]
    ids = [S.InterpDensSpecies(grd) for _ in range(c.ns)]
    param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                         PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
    eng = S.Engine(grd, c.ns, sort_mode=sort_mode)
    for s, p in enumerate(parts):
        eng.load(s, p)
    spilled = 0
    for _ in range(4):
        S.setZeroSpeciesDensities(ids, grd, c.ns)
        eng.phase(field, ids, param)
        spilled = max(spilled, eng.stats()["overflow"])
        eng.sort()
        assert eng.stats()["overflow"] == 0
    assert spilled > 0
    for s, p in enumerate(parts):
        eng.store(s, p)
    got_p = [p.download() for p in parts]
    got_m = [m.download() for m in ids]
    eng.close()
    oracle_phases(st, 4)
    for s in range(c.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[s][k], st.parts[s][k], f"s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], 1e-13, f"s{s} {n}")


def test_engine_rejects_unknown_sort_mode():
    st = oracle_small("f64")
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f64")
    with pytest.raises(S.SpicError) as ei:
        S.Engine(grd, c.ns, sort_mode=2)
    assert ei.value.code == S.SPIC_EINVAL


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", CASES)
def test_engine_fixture_cycles(name, prec):
    """Reference fixtures: two phases, then compare with cycle 2 of the
    reference run (mover_PC + interpP2G twice from the same IC)."""
    data = load_case(name, prec)
    st = oracle_from_fixture(name, prec, data)
    S.set_strict_fp(True)
    got_p, got_m, _ = run_engine(st, prec, phases=2)
    for s in range(st.cfg.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[s][k], data[f"c2_s{s}_{k}"], f"c2 s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], data[f"c2_s{s}_{n}"], 1e-13 if prec == "f64" else 1e-5,
                         f"c2 s{s} {n}")


def test_engine_fast_particles_cross_tiles_every_phase():
    """gem_hot particles (E set to zero so speeds stay below one cell per
    step) cross cells and tiles every phase; each phase re-bins them and
    the trajectories stay exact."""
    data = load_case("gem_hot", "f64")
    st = oracle_from_fixture("gem_hot", "f64", data)
    for n in ("Ex", "Ey", "Ez"):
        st.fields[n][:] = 0.0
    x0 = [p["x"].copy() for p in st.parts]
    S.set_strict_fp(True)
    got_p, got_m, _ = run_engine(st, "f64", phases=8, sort_every=0)
    oracle_phases(st, 8)
    dx = st.cfg.Lx / st.cfg.nxc
    moved = sum(int(np.sum(np.abs(p["x"] - x) > 3 * dx)) for p, x in zip(st.parts, x0))
    assert moved > 100, moved
    for s in range(st.cfg.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[s][k], st.parts[s][k], f"s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], 1e-13, f"s{s} {n}")


@pytest.mark.parametrize("strict", [True, False])
@pytest.mark.parametrize("prec", PRECS)
def test_engine_iterates_outside_the_staged_block(prec, strict):
    """The ion species' particles in cell layer 3 of the gem_hot grid (8x4x4
    cells, 4x4x4 tiles, so the last layer of tile 0) drift +4.4 cells per
    step along x with the fixture's fields: the predictor-corrector iterates
    (half a step, about 2.2 cells) land in cell layers 5-6, beyond the
    staged block and its one-cell halo, so the push gathers the stencil
    from global memory; the final positions reach tile 1 or wrap through
    the periodic x face into tile 0.  Ions only: their gyration per step is
    0.18 rad at the fixture's largest B, so the drift stays along x and the
    iterates stay within the grid (the electrons' 11.6 rad would turn it
    into y, towards the reflecting walls).  Strict arithmetic: trajectories
    bit for bit; production arithmetic: within the bounds of the other
    production tests."""
    data = load_case("gem_hot", prec)
    st = oracle_from_fixture("gem_hot", prec, data)
    c = st.cfg
    dx = c.Lx / c.nxc
    ions = [s for s in range(c.ns) if c.qom[s] > 0]
    assert len(ions) == 2
    for s in ions:
        p = st.parts[s]
        layer3 = (p["x"] >= 3 * dx) & (p["x"] < 4 * dx)
        assert int(np.sum(layer3)) == 32
        p["u"][layer3] = 4.4 * dx / c.dt
    S.set_strict_fp(strict)
    got_p, got_m, stats = run_engine(st, prec, phases=1)
    oracle_phases(st, 1)
    assert stats["global_gathers"] >= 2 * 32, stats
    for s in range(c.ns):
        if strict:
            for k in O.PARTICLE_NAMES:
                assert_bitwise(got_p[s][k], st.parts[s][k], f"s{s} {k}")
        else:
            tol = 1e-12 if prec == "f64" else 1e-5
            for k in ("u", "v", "w"):
                assert_close(got_p[s][k], st.parts[s][k], tol, f"s{s} {k}")
            for k in ("x", "y", "z"):
                assert np.max(np.abs(got_p[s][k] - st.parts[s][k])) < (
                    1e-11 if prec == "f64" else 1e-5), f"s{s} {k}"
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n],
                         (1e-13 if strict else 1e-12) if prec == "f64" else 1e-5,
                         f"s{s} {n}")


@pytest.mark.parametrize("set_moments", [False, True])
@pytest.mark.parametrize("prec", PRECS)
def test_engine_overflow_region(prec, set_moments):
    """A full segment spills into the overflow region, which the next phase
    pushes with the global-memory kernel.  gem_hot grid (8x4x4 cells, two
    4x4x4 tiles along x) with no fields: species 0's particles in the last
    cell layer of tile 0 drift +0.9 cell per step into tile 1, all others
    rest, so tile 1 gains 32 particles against 16 free slots.  With
    set_moments the phase writes the moments, global-atomic deposits
    included."""
    data = load_case("gem_hot", prec)
    st = oracle_from_fixture("gem_hot", prec, data)
    for n in O.FIELD_NAMES:
        st.fields[n][:] = 0.0
    c = st.cfg
    dx = c.Lx / c.nxc
    for s, p in enumerate(st.parts):
        for k in ("u", "v", "w"):
            p[k][:] = 0.0
        if s == 0:
            last = (p["x"] >= 3 * dx) & (p["x"] < 4 * dx)
            p["u"][last] = 0.9 * dx / c.dt
            assert int(np.sum(last)) == 32
    S.set_strict_fp(True)
    got_p, got_m, stats = run_engine(st, prec, phases=3, set_moments=set_moments)
    oracle_phases(st, 3)
    assert stats["overflow"] > 0, stats
    assert stats["global_deposits"] > 0, stats
    for s in range(c.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[s][k], st.parts[s][k], f"s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], 1e-13 if prec == "f64" else 1e-5,
                         f"s{s} {n}")


@pytest.mark.parametrize("prec,positions", [("f64", 0), ("f32", S.SPIC_POS_FLOAT),
                                            ("f32", S.SPIC_POS_FIXED)])
def test_engine_load_store_round_trip(prec, positions):
    """Load then store without a phase returns every particle at its load
    index: bit for bit with float positions (and for velocities and
    charges always); fixed-point positions come back within half their
    unit (2^-(k+1) cells) rounded to float."""
    st = oracle_small(prec)
    got_p, _, stats = run_engine(st, prec, phases=0, positions=positions)
    assert stats["sorts"] == st.cfg.ns and stats["overflow"] == 0
    assert stats["positions"] == (positions or S.SPIC_POS_FLOAT)
    c = st.cfg
    for s in range(c.ns):
        for k in O.PARTICLE_NAMES:
            if positions == S.SPIC_POS_FIXED and k in ("x", "y", "z"):
                L, n = {"x": (c.Lx, c.nxc), "y": (c.Ly, c.nyc), "z": (c.Lz, c.nzc)}[k]
                kbits = max(b for b in range(25) if (n + 64) << b < 2 ** 31)
                want = st.parts[s][k].astype(np.float64)
                err = np.abs(got_p[s][k].astype(np.float64) - want)
                bound = 0.5 * (L / n) * 2.0 ** -kbits + np.spacing(np.abs(want).astype(
                    np.float32)).astype(np.float64)
                assert np.all(err <= bound), f"s{s} {k}: {np.max(err - bound)}"
            else:
                assert_bitwise(got_p[s][k], st.parts[s][k], f"s{s} {k}")


def test_engine_fp32_fixed_point_keeps_fp64_input_precision():
    """An FP32 engine with fixed-point positions loaded from FP64 arrays
    (spic_engine_load_as) and stored into FP64 arrays returns positions
    within half the fixed-point unit of the FP64 input -- about 1e-8 of a
    box length here, 100 times finer than float lengths near the far faces
    -- and velocities and charges rounded to float."""
    st = oracle_small("f64")
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f32")
    eng = S.Engine(grd, c.ns, positions=S.SPIC_POS_FIXED)
    try:
        for s in range(c.ns):
            p64 = S.Particles(st.parts[s], c.qom[s], "f64")
            eng.load(s, p64)
            out = S.DeviceParticles(len(st.parts[s]["x"]), c.qom[s], "f64")
            eng.store(s, out)
            got = out.download()
            for k, L, n in (("x", c.Lx, c.nxc), ("y", c.Ly, c.nyc), ("z", c.Lz, c.nzc)):
                kbits = max(b for b in range(25) if (n + 64) << b < 2 ** 31)
                err = np.max(np.abs(got[k] - st.parts[s][k]))
                assert err <= 0.5 * (L / n) * 2.0 ** -kbits * (1 + 1e-9), (s, k, err)
            for k in ("u", "v", "w", "q"):
                assert_bitwise(got[k], st.parts[s][k].astype(np.float32).astype(np.float64),
                               f"s{s} {k}")
    finally:
        eng.close()


def test_engine_fixed_point_refused_in_strict_arithmetic():
    st = oracle_small("f32")
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f32")
    eng = S.Engine(grd, 1)  # created in production arithmetic: fixed point
    try:
        eng.load(0, S.Particles(st.parts[0], c.qom[0], "f32"))
        assert eng.stats()["positions"] == S.SPIC_POS_FIXED
        field = S.EMField(grd, {n: st.fields[n] for n in O.FIELD_NAMES})
        S.set_strict_fp(True)
        with pytest.raises(S.SpicError) as ei:
            eng.phase(field, [S.InterpDensSpecies(grd)], S.Parameters())
        assert ei.value.code == S.SPIC_EINVAL
    finally:
        S.set_strict_fp(False)
        eng.close()
    with pytest.raises(S.SpicError) as ei:
        S.Engine(S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f64"), 1,
                 positions=S.SPIC_POS_FIXED)
    assert ei.value.code == S.SPIC_EUNSUPPORTED


@pytest.mark.parametrize("prec", PRECS)
def test_engine_fast_against_f64_oracle(prec):
    st = oracle_small("f64")
    got_p, got_m, _ = run_engine(st, prec, phases=2, sort_every=1)
    oracle_phases(st, 2)
    tol = 1e-12 if prec == "f64" else 1e-5
    for s in range(st.cfg.ns):
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], tol, f"s{s} {n}")
        for k in ("u", "v", "w"):
            assert_close(got_p[s][k], st.parts[s][k], tol, f"s{s} {k}")
        for k in ("x", "y", "z"):
            assert np.max(np.abs(got_p[s][k] - st.parts[s][k])) < (1e-11 if prec == "f64"
                                                                   else 1e-5), k


@pytest.mark.parametrize("prec", PRECS)
def test_engine_deposit_is_interp_p2g(prec):
    """spic_engine_deposit: interpP2G of the loaded particles without a push,
    against the oracle's interpP2G of the same arrays; the particles come
    back unchanged, the call is not a phase, and a phase after it is the
    oracle's mover_PC + interpP2G."""
    st = oracle_small(prec)
    c = st.cfg
    S.set_strict_fp(True)
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, prec)
    dt = np.float64 if prec == "f64" else np.float32
    field = S.EMField(grd, {n: st.fields[n].astype(dt) for n in O.FIELD_NAMES})
    parts = [S.Particles(p, c.qom[s], prec, c.NiterMover) for s, p in enumerate(st.parts)]
    ids = [S.InterpDensSpecies(grd) for _ in range(c.ns)]
    param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                         PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
    eng = S.Engine(grd, c.ns)
    for s, p in enumerate(parts):
        eng.load(s, p)
    before = [{k: v.copy() for k, v in p.items()} for p in st.parts]
    S.setZeroSpeciesDensities(ids, grd, c.ns)
    eng.deposit(ids)
    assert eng.stats()["phases"] == 0
    tol = 1e-13 if prec == "f64" else 1e-5
    for s in range(c.ns):
        st.zero_moments(s)
        st.interp_p2g(s)
        got = ids[s].download()
        for n in O.MOMENT_NAMES:
            assert_close(got[n], st.moments[s][n], tol, f"deposit s{s} {n}")
    S.setZeroSpeciesDensities(ids, grd, c.ns)
    eng.phase(field, ids, param)
    for s, p in enumerate(parts):
        eng.store(s, p)
    for s in range(c.ns):
        for k in O.PARTICLE_NAMES:
            st.parts[s][k][:] = before[s][k]
    oracle_phases(st, 1)
    got_p = [p.download() for p in parts]
    got_m = [m.download() for m in ids]
    for s in range(c.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[s][k], st.parts[s][k], f"phase s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], tol, f"phase s{s} {n}")
    eng.close()


@pytest.mark.parametrize("prec", PRECS)
def test_engine_species_storage_grows_and_is_reused(prec):
    """One species slot loaded three times -- a small species, an eight
    times larger one (the storage grows), the small one again (the storage
    is kept, with the smaller component stride) -- each followed by a phase
    that matches the oracle."""
    S.set_strict_fp(True)
    small = O.GemConfig(npcelx=[1] * 4, npcely=[1] * 4, npcelz=[1] * 4)
    big = O.GemConfig(npcelx=[2] * 4, npcely=[2] * 4, npcelz=[2] * 4)
    eng, grd, used = None, None, []
    tol = 1e-13 if prec == "f64" else 1e-5
    for cfg in (small, big, small):
        st = O.State(cfg, prec)
        c = st.cfg
        if eng is None:
            grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, prec)
            eng = S.Engine(grd, 1)
        dt = np.float64 if prec == "f64" else np.float32
        field = S.EMField(grd, {n: st.fields[n].astype(dt) for n in O.FIELD_NAMES})
        part = S.Particles(st.parts[0], c.qom[0], prec, c.NiterMover)
        ids = [S.InterpDensSpecies(grd)]
        param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                             PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
        eng.load(0, part)
        used.append(eng.stats()["device_bytes"])
        S.setZeroSpeciesDensities(ids, grd, 1)
        eng.phase(field, ids, param)
        eng.store(0, part)
        st.mover_pc(0)
        st.zero_moments(0)
        st.interp_p2g(0)
        got_p, got_m = part.download(), ids[0].download()
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[k], st.parts[0][k], f"n={len(st.parts[0]['x'])} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[n], st.moments[0][n], tol, f"n={len(st.parts[0]['x'])} {n}")
    assert used[0] < used[1] == used[2]
    eng.close()


def test_engine_conserves_charge():
    st = oracle_small("f64")
    got_p, got_m, _ = run_engine(st, "f64", phases=1)
    iv = O.State(st.cfg, "f64", build_ic=False).grid.invVOL
    for s in range(st.cfg.ns):
        q = got_p[s]["q"]
        assert abs(np.sum(got_m[s]["rhon"]) / iv - np.sum(q)) <= 1e-12 * np.sum(np.abs(q))


def test_engine_reports_out_of_domain():
    st = oracle_small("f64")
    st.parts[0]["x"][7] = -5.0
    with pytest.raises(S.SpicError) as ei:
        run_engine(st, "f64", phases=1)
    assert ei.value.code == S.SPIC_EDOMAIN


@pytest.mark.parametrize("prec", PRECS)
def test_engine_input_order_does_not_matter(prec):
    """Particles in random order (spic_permute_species: index i to
    (a i + b) mod n) give the engine the same trajectories, particle for
    particle, as the oracle run on the same permuted arrays."""
    import ctypes as C
    from sputnipic_amd._capi import check, lib

    st = oracle_small(prec)
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, prec)
    g_c = grd.c_struct()
    a, b = 7919, 101
    for s in range(c.ns):
        n = len(st.parts[s]["x"])
        src = S.Particles(st.parts[s], c.qom[s], prec, c.NiterMover)
        dst = S.DeviceParticles(n, c.qom[s], prec)
        check(lib().spic_permute_species(C.byref(g_c), C.byref(src.c_struct()),
                                         C.byref(dst.c_struct()), a, b, None), "permute")
        got = dst.download()
        j = (a * np.arange(n, dtype=np.uint64) + b) % np.uint64(n)
        assert np.unique(j).size == n
        for k in O.PARTICLE_NAMES:
            want = np.empty_like(st.parts[s][k])
            want[j] = st.parts[s][k]
            assert_bitwise(got[k], want, f"permuted s{s} {k}")
            st.parts[s][k] = want
    S.set_strict_fp(True)
    got_p, got_m, _ = run_engine(st, prec, phases=2)
    oracle_phases(st, 2)
    for s in range(c.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(got_p[s][k], st.parts[s][k], f"s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], 1e-13 if prec == "f64" else 1e-5,
                         f"s{s} {n}")
