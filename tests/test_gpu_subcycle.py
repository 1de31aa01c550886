"""Sub-cycling of the mover (GPU): species with n_sub_cycles = 2 and 3.

The reference's two movers differ: the CPU mover pushes once with
dt / n_sub_cycles (src/Mover.cpp:11-12, no loop); the GPU mover pushes
n_sub_cycles times with that step and applies the particle BC after each
push (src/gpu/Particles_gpu.cu:319-437).  The oracle restates the CPU mover,
so:

* SPIC_SUBCYCLE_SINGLE against one oracle mover call with the species'
  n_sub_cycles;
* SPIC_SUBCYCLE_LOOP against n_sub_cycles oracle mover calls in a row (the
  GPU mover's loop is exactly that: its BC's `else if` differs from the CPU
  mover's two tests only for a displacement longer than the box, which no
  particle here makes; its step (FPpart)dt / (float)n equals the oracle's
  (FPpart)dt / (double)n for these n, checked below);
* SPIC_SUBCYCLE_REFUSE (default): the mover and an engine phase refuse.

Strict arithmetic: particle state bit for bit, moments to 1e-13 (FP64) and
1e-5 (FP32, whose oracle accumulates in float) of each array's largest
magnitude, as the other strict parity tests; the caller-order mover_PC and
mover_interp, and the tile engine's phase.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S
from sputnipic_amd._capi import SPIC_EUNSUPPORTED, SpicError

pytestmark = pytest.mark.gpu

CFG = dict(nxc=16, nyc=8, nzc=8, npcelx=[2] * 4, npcely=[2] * 4, npcelz=[2] * 4)
MOM_TOL = {"f64": 1e-13, "f32": 1e-5}


@pytest.fixture(autouse=True)
def _modes():
    S.set_async(False)
    S.set_strict_fp(True)
    yield
    S.set_subcycling(S.SPIC_SUBCYCLE_REFUSE)
    S.set_strict_fp(False)


def _setup(prec, nsub):
    cfg = O.GemConfig(n_sub_cycles=nsub, **CFG)
    st = O.State(cfg, prec)
    grd = S.Grid(cfg.nxc, cfg.nyc, cfg.nzc, cfg.Lx, cfg.Ly, cfg.Lz, prec)
    field = S.EMField(grd, st.fields)
    param = S.Parameters(dt=cfg.dt, c=cfg.c, NiterMover=cfg.NiterMover, n_sub_cycles=nsub,
                         PERIODICX=cfg.PERIODICX, PERIODICY=cfg.PERIODICY,
                         PERIODICZ=cfg.PERIODICZ)
    return cfg, st, grd, field, param


def _oracle_cycle(st, s, pushes):
    for _ in range(pushes):
        st.mover_pc(s)
    st.zero_moments(s)
    st.interp_p2g(s)


def _check(st, s, got_p, got_m, prec):
    for k in O.PARTICLE_NAMES:
        assert np.array_equal(got_p[k], st.parts[s][k]), (s, k)
    for n in O.MOMENT_NAMES:
        want = st.moments[s][n]
        err = np.max(np.abs(got_m[n] - want)) / max(np.max(np.abs(want)), 1e-300)
        assert err <= MOM_TOL[prec], (s, n, err)


@pytest.mark.parametrize("nsub", [2, 3])
def test_loop_step_equals_the_oracle_step(nsub):
    """(float)dt / (float)n and (float)((double)(float)dt / n) agree, so n
    oracle calls reproduce the GPU mover's loop in FP32 too."""
    dt = np.float32(0.25)
    assert np.float32(dt / np.float32(nsub)) == np.float32(np.float64(dt) / nsub)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("mode", ["loop", "single"])
@pytest.mark.parametrize("nsub", [2, 3])
def test_caller_order_mover_subcycles(prec, mode, nsub):
    cfg, st, grd, field, param = _setup(prec, nsub)
    S.set_subcycling(S.SPIC_SUBCYCLE_LOOP if mode == "loop" else S.SPIC_SUBCYCLE_SINGLE)
    pushes = nsub if mode == "loop" else 1
    for s in range(cfg.ns):
        part = S.Particles(st.parts[s], cfg.qom[s], prec, cfg.NiterMover, nsub)
        ids = S.InterpDensSpecies(grd)
        if s % 2 == 0:
            S.mover_PC(part, field, grd, param)
            S.interpP2G(part, ids, grd)
        else:
            S.mover_interp(part, field, ids, grd, param)
        _oracle_cycle(st, s, pushes)
        _check(st, s, part.download(), ids.download(), prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("nsub", [2, 3])
def test_engine_phase_subcycles(prec, nsub):
    cfg, st, grd, field, param = _setup(prec, nsub)
    S.set_subcycling(S.SPIC_SUBCYCLE_LOOP)
    parts = [S.Particles(p, cfg.qom[s], prec, cfg.NiterMover, nsub)
             for s, p in enumerate(st.parts)]
    ids = [S.InterpDensSpecies(grd) for _ in range(cfg.ns)]
    eng = S.Engine(grd, cfg.ns, set_moments=True, positions=S.SPIC_POS_FLOAT)
    try:
        for s, p in enumerate(parts):
            eng.load(s, p)
        eng.phase(field, ids, param)
        for s, p in enumerate(parts):
            eng.store(s, p)
    finally:
        eng.close()
    for s in range(cfg.ns):
        _oracle_cycle(st, s, nsub)
        _check(st, s, parts[s].download(), ids[s].download(), prec)


def test_refused_by_default():
    cfg, st, grd, field, param = _setup("f64", 2)
    part = S.Particles(st.parts[0], cfg.qom[0], "f64", cfg.NiterMover, 2)
    with pytest.raises(SpicError) as exc:
        S.mover_PC(part, field, grd, param)
    assert exc.value.code == SPIC_EUNSUPPORTED
    # an engine loaded while a mode was chosen refuses its phase once the
    # mode is back to refusing
    S.set_subcycling(S.SPIC_SUBCYCLE_LOOP)
    eng = S.Engine(grd, 1, set_moments=True, positions=S.SPIC_POS_FLOAT)
    try:
        eng.load(0, part)
        S.set_subcycling(S.SPIC_SUBCYCLE_REFUSE)
        with pytest.raises(SpicError) as exc:
            eng.phase(field, [S.InterpDensSpecies(grd)], param)
        assert exc.value.code == SPIC_EUNSUPPORTED
    finally:
        eng.close()
