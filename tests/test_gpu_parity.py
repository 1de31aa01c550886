"""GPU parity of the caller-order path (spic_mover_pc, spic_interp_p2g,
spic_mover_interp, spic_particle_sort through the C ABI) against the
oracle and the reference's golden fixtures.

Tolerances (stated per check):
* strict arithmetic (spic_set_strict_fp(1)): particle state bit-identical to
  the reference in FP64 and FP32; node moments differ only by summation
  order (FP64 device accumulation): 1e-13 of the array's max magnitude
  against an FP64 reference, 1e-5 against the FP32 reference, which
  accumulates in float.
* production arithmetic (FMA contraction): FP64 particles within 1e-12 and
  moments within 1e-12 of the FP64 oracle; FP32 moments within 1e-5 of the
  FP64 oracle (the FP32-form bound).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S
from helpers import (CASES, PRECS, assert_bitwise, assert_close, case_config, load_case,
                     max_rel_err, oracle_from_fixture)

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _modes():
    S.set_async(False)
    S.set_strict_fp(False)
    yield
    S.set_strict_fp(False)


class GpuState:
    """Device copies of an oracle state's grid, fields and species."""

    def __init__(self, st: O.State, prec: str):
        c = st.cfg
        self.cfg, self.prec = c, prec
        self.grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, prec)
        dt = np.float64 if prec == "f64" else np.float32
        assert_bitwise(self.grd.xn_host, st.XN.reshape(self.grd.shape)[:, 0, 0].astype(dt),
                       "device grid xn")
        self.field = S.EMField(self.grd, {n: st.fields[n].astype(dt) for n in O.FIELD_NAMES})
        self.parts = [S.Particles(p, c.qom[s], prec, c.NiterMover, c.n_sub_cycles)
                      for s, p in enumerate(st.parts)]
        self.ids = [S.InterpDensSpecies(self.grd) for _ in range(c.ns)]
        self.param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover,
                                  n_sub_cycles=c.n_sub_cycles, PERIODICX=c.PERIODICX,
                                  PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ,
                                  sort_cps=c.sort_cps)

    def zero(self):
        S.setZeroSpeciesDensities(self.ids, self.grd, self.cfg.ns)


# --------------------------------------------------------------------------
# Golden fixtures from the reference's own code
# --------------------------------------------------------------------------
@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", CASES)
def test_fixture_cycles_strict(name, prec):
    data = load_case(name, prec)
    st = oracle_from_fixture(name, prec, data)
    gs = GpuState(st, prec)
    S.set_strict_fp(True)
    mom_tol = 1e-13 if prec == "f64" else 1e-5
    ns = gs.cfg.ns
    for s in range(ns):
        assert S.mover_PC(gs.parts[s], gs.field, gs.grd, gs.param) == 0
    gs.zero()
    for s in range(ns):
        S.interpP2G(gs.parts[s], gs.ids[s], gs.grd)
        p = gs.parts[s].download()
        for k in O.PARTICLE_NAMES:
            assert_bitwise(p[k], data[f"c1_s{s}_{k}"], f"c1 s{s} {k}")
        m = gs.ids[s].download()
        for n in O.MOMENT_NAMES:
            assert_close(m[n], data[f"c1_s{s}_{n}"], mom_tol, f"c1 s{s} {n}")
    for s in range(ns):
        S.mover_PC(gs.parts[s], gs.field, gs.grd, gs.param)
    gs.zero()
    for s in range(ns):
        S.interpP2G(gs.parts[s], gs.ids[s], gs.grd)
        p = gs.parts[s].download()
        for k in O.PARTICLE_NAMES:
            assert_bitwise(p[k], data[f"c2_s{s}_{k}"], f"c2 s{s} {k}")
        m = gs.ids[s].download()
        for n in O.MOMENT_NAMES:
            assert_close(m[n], data[f"c2_s{s}_{n}"], mom_tol, f"c2 s{s} {n}")
    gs.zero()
    for s in range(ns):
        assert S.mover_interp(gs.parts[s], gs.field, gs.ids[s], gs.grd, gs.param) == 0
        p = gs.parts[s].download()
        for k in O.PARTICLE_NAMES:
            assert_bitwise(p[k], data[f"c3_s{s}_{k}"], f"c3 s{s} {k}")
        m = gs.ids[s].download()
        for n in O.MOMENT_NAMES:
            assert_close(m[n], data[f"c3_s{s}_{n}"], mom_tol, f"c3 s{s} {n}")
    for s in range(ns):
        S.particle_sort(gs.param, gs.parts[s], gs.grd)
        p = gs.parts[s].download()
        for k in O.PARTICLE_NAMES:
            assert_bitwise(p[k], data[f"sort_s{s}_{k}"], f"sort s{s} {k}")


# --------------------------------------------------------------------------
# The small Harris-sheet case (32x16x16 cells, 884,736 particles) vs the
# oracle
# --------------------------------------------------------------------------
_CACHE = {}


def oracle_small(prec):
    if prec not in _CACHE:
        _CACHE[prec] = O.State(O.GemConfig(), prec)
    st = _CACHE[prec]
    fresh = O.State(st.cfg, prec, build_ic=False)
    fresh.copy_from(st)
    return fresh


@pytest.mark.parametrize("prec", PRECS)
def test_small_case_strict_cycles_bitwise(prec):
    """Three cycles of mover_PC + zero + interpP2G, strict arithmetic."""
    st = oracle_small(prec)
    gs = GpuState(st, prec)
    S.set_strict_fp(True)
    mom_tol = 1e-13 if prec == "f64" else 1e-5
    for _cycle in range(3):
        st.cycle()
        for s in range(st.cfg.ns):
            S.mover_PC(gs.parts[s], gs.field, gs.grd, gs.param)
        gs.zero()
        for s in range(st.cfg.ns):
            S.interpP2G(gs.parts[s], gs.ids[s], gs.grd)
    for s in range(st.cfg.ns):
        p = gs.parts[s].download()
        for k in O.PARTICLE_NAMES:
            assert_bitwise(p[k], st.parts[s][k], f"s{s} {k}")
        m = gs.ids[s].download()
        for n in O.MOMENT_NAMES:
            assert_close(m[n], st.moments[s][n], mom_tol, f"s{s} {n}")


def test_small_case_fast_f64_within_tolerance():
    st = oracle_small("f64")
    gs = GpuState(st, "f64")
    st.cycle()
    for s in range(st.cfg.ns):
        S.mover_PC(gs.parts[s], gs.field, gs.grd, gs.param)
    gs.zero()
    for s in range(st.cfg.ns):
        S.interpP2G(gs.parts[s], gs.ids[s], gs.grd)
        p = gs.parts[s].download()
        for k in ("x", "y", "z", "u", "v", "w"):
            assert_close(p[k], st.parts[s][k], 1e-12, f"s{s} {k}")
        m = gs.ids[s].download()
        for n in O.MOMENT_NAMES:
            assert_close(m[n], st.moments[s][n], 1e-12, f"s{s} {n}")


def test_small_case_fast_f32_against_f64_oracle():
    """FP32 form vs the FP64 oracle, both from the FP64 initial state:
    node moments within 1e-5 of each array's max magnitude."""
    st = oracle_small("f64")
    gs = GpuState(st, "f32")
    st.cycle()
    for s in range(st.cfg.ns):
        S.mover_PC(gs.parts[s], gs.field, gs.grd, gs.param)
    gs.zero()
    for s in range(st.cfg.ns):
        S.interpP2G(gs.parts[s], gs.ids[s], gs.grd)
        m = gs.ids[s].download()
        for n in O.MOMENT_NAMES:
            assert_close(m[n], st.moments[s][n], 1e-5, f"s{s} {n}")
        p = gs.parts[s].download()
        for k in ("x", "y", "z"):
            assert np.max(np.abs(p[k] - st.parts[s][k])) < 1e-5, k
        for k in ("u", "v", "w"):
            assert_close(p[k], st.parts[s][k], 1e-5, k)


@pytest.mark.parametrize("prec", PRECS)
def test_fused_mover_interp_matches_oracle(prec):
    st = oracle_small(prec)
    gs = GpuState(st, prec)
    S.set_strict_fp(True)
    gs.zero()
    for s in range(st.cfg.ns):
        st.zero_moments(s)
        st.mover_interp(s)
        S.mover_interp(gs.parts[s], gs.field, gs.ids[s], gs.grd, gs.param)
        p = gs.parts[s].download()
        for k in O.PARTICLE_NAMES:
            assert_bitwise(p[k], st.parts[s][k], f"s{s} {k}")
        m = gs.ids[s].download()
        for n in O.MOMENT_NAMES:
            assert_close(m[n], st.moments[s][n], 1e-13 if prec == "f64" else 1e-5, n)


@pytest.mark.parametrize("sort_cps", [1, 2, 4])
@pytest.mark.parametrize("prec", PRECS)
def test_particle_sort_reference_order(prec, sort_cps):
    st = oracle_small(prec)
    st.cycle()  # move first so the input is not already in lattice order
    gs = GpuState(st, prec)
    for s in range(st.cfg.ns):
        st.particle_sort(s, sort_cps)
        gs.param.sort_cps = sort_cps
        S.particle_sort(gs.param, gs.parts[s], gs.grd)
        p = gs.parts[s].download()
        for k in O.PARTICLE_NAMES:
            assert_bitwise(p[k], st.parts[s][k], f"sort s{s} {k}")


# --------------------------------------------------------------------------
# Invariants and edge cases
# --------------------------------------------------------------------------
def test_deposit_conserves_charge_on_device():
    st = oracle_small("f64")
    gs = GpuState(st, "f64")
    for s in range(st.cfg.ns):
        S.interpP2G(gs.parts[s], gs.ids[s], gs.grd)
        m = gs.ids[s].download()
        q = st.parts[s]["q"]
        total = np.sum(m["rhon"]) / gs.grd.invVOL
        assert abs(total - np.sum(q)) <= 1e-12 * np.sum(np.abs(q))
        jx = np.sum(m["Jx"]) / gs.grd.invVOL
        qu = q * st.parts[s]["u"]
        assert abs(jx - np.sum(qu)) <= 1e-12 * np.sum(np.abs(qu))


def test_interp_accumulates_and_respects_nop():
    st = oracle_small("f64")
    gs = GpuState(st, "f64")
    p = gs.parts[0]
    S.interpP2G(p, gs.ids[0], gs.grd)
    once = gs.ids[0].download()
    S.interpP2G(p, gs.ids[0], gs.grd)
    twice = gs.ids[0].download()
    for n in O.MOMENT_NAMES:
        assert_close(twice[n], 2 * once[n], 1e-14, f"+= {n}")
    # only the first nop particles move / deposit
    half = p.npmax // 2
    p.nop = half
    before = p.download()
    S.mover_PC(p, gs.field, gs.grd, gs.param)
    after = p.download()
    assert not np.array_equal(after["x"][:half], before["x"][:half])
    assert_bitwise(after["x"][half:], before["x"][half:], "untouched tail")


def test_empty_species():
    st = oracle_small("f64")
    gs = GpuState(st, "f64")
    p = gs.parts[0]
    p.nop = 0
    S.mover_PC(p, gs.field, gs.grd, gs.param)
    S.interpP2G(p, gs.ids[0], gs.grd)
    m = gs.ids[0].download()
    assert all(not np.any(m[n]) for n in O.MOMENT_NAMES)


def test_out_of_domain_particle_is_reported():
    st = oracle_small("f64")
    bad = {k: v[:64].copy() for k, v in st.parts[0].items()}
    bad["x"][5] = -3 * (st.cfg.Lx / st.cfg.nxc)  # two cells outside the grid
    gs = GpuState(st, "f64")
    p = S.Particles(bad, st.cfg.qom[0], "f64")
    with pytest.raises(S.SpicError) as ei:
        S.interpP2G(p, gs.ids[0], gs.grd)
    assert ei.value.code == S.SPIC_EDOMAIN
    with pytest.raises(S.SpicError) as ei:
        S.mover_PC(p, gs.field, gs.grd, gs.param)
    assert ei.value.code == S.SPIC_EDOMAIN


def test_particle_at_upper_face_touches_ghost_with_zero_weight():
    """x == Lx lands in the last cell + 1 and weights the ghost node by 0
    (src/Mover.cpp:399-409); indices stay inside the arrays."""
    st = oracle_small("f64")
    one = {k: v[:1].copy() for k, v in st.parts[0].items()}
    one["x"][0] = st.cfg.Lx
    gs = GpuState(st, "f64")
    p = S.Particles(one, st.cfg.qom[0], "f64")
    S.interpP2G(p, gs.ids[0], gs.grd)
    m = gs.ids[0].download()
    o = O.State(st.cfg, "f64", build_ic=False)
    o.parts = [one]
    o.interp_p2g(0)
    for n in O.MOMENT_NAMES:
        assert_close(m[n], o.moments[0][n], 1e-15, n)
    assert max_rel_err(m["rhon"], o.moments[0]["rhon"]) == 0.0
