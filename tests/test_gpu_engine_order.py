"""Engines without load order (SPIC_ORDER_ENGINE) and with a bounded
overflow region (overflow_pct), on the GPU.

* Without ids the engine moves the same particles the same way: in strict
  arithmetic, after phases with tile and radix re-sorts in between, the
  stored particles are the oracle's as a set (rows compared after sorting,
  bit for bit) and the moments match as in the other strict tests (1e-13
  FP64, 1e-5 FP32 against the FP32 oracle); the snapshot words are the
  oracle's float32 values as a set.  In production FP32 (fixed point) the
  set equals that of an engine keeping the load order, bit for bit.
* A bounded overflow region holds what a normal phase needs (the Harris
  sheet runs as with the default region, with less device memory); a
  phase that sends more particles into a full segment than the region
  holds reports SPIC_ENOMEM, while the default region takes them all.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S
from helpers import assert_close
from sputnipic_amd._capi import SPIC_ENOMEM, SpicError
from sputnipic_amd.snapshot import Snapshot

pytestmark = pytest.mark.gpu

CFG = dict(nxc=16, nyc=8, nzc=8, npcelx=[2] * 4, npcely=[2] * 4, npcelz=[2] * 4)


@pytest.fixture(autouse=True)
def _modes():
    S.set_async(False)
    S.set_strict_fp(False)
    yield
    S.set_strict_fp(False)


def _rows(p, keys=O.PARTICLE_NAMES):
    """Particles as rows sorted lexicographically (a canonical set order)."""
    a = np.stack([np.asarray(p[k]) for k in keys], axis=1)
    return a[np.lexsort(a.T[::-1])]


def _engine_run(st, prec, phases, order, sort_mode=0, overflow_pct=0, positions=0):
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, prec)
    dt = np.float64 if prec == "f64" else np.float32
    field = S.EMField(grd, {n: st.fields[n].astype(dt) for n in O.FIELD_NAMES})
    parts = [S.Particles(p, c.qom[s], prec, c.NiterMover) for s, p in enumerate(st.parts)]
    ids = [S.InterpDensSpecies(grd) for _ in range(c.ns)]
    param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                         PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
    eng = S.Engine(grd, c.ns, set_moments=True, sort_mode=sort_mode, order=order,
                   overflow_pct=overflow_pct, positions=positions)
    try:
        for s, p in enumerate(parts):
            eng.load(s, p)
        for k in range(phases):
            if k:
                eng.sort()
            eng.phase(field, ids, param)
        snap = Snapshot([len(p["x"]) for p in st.parts])
        snap.capture_engine(eng)
        words = snap.result()
        for s, p in enumerate(parts):
            eng.store(s, p)
        stats = eng.stats()
    finally:
        eng.close()
    return [p.download() for p in parts], [m.download() for m in ids], words, stats


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("sort_mode", [S.SPIC_SORT_TILE, S.SPIC_SORT_RADIX])
def test_engine_order_particles_are_the_oracles(prec, sort_mode):
    st = O.State(O.GemConfig(**CFG), prec)
    S.set_strict_fp(True)
    got_p, got_m, words, stats = _engine_run(st, prec, 3, S.SPIC_ORDER_ENGINE, sort_mode)
    for _ in range(3):
        st.cycle()
    for s in range(st.cfg.ns):
        assert np.array_equal(_rows(got_p[s]), _rows(st.parts[s])), f"s{s}"
        for n in O.MOMENT_NAMES:
            assert_close(got_m[s][n], st.moments[s][n], 1e-13 if prec == "f64" else 1e-5,
                         f"s{s} {n}")
        want = {k: st.parts[s][k].astype(np.float32) for k in "xyzuvw"}
        got = {k: words[s][k].astype(np.float32) for k in "xyzuvw"}
        assert np.array_equal(_rows(got, "xyzuvw"), _rows(want, "xyzuvw")), f"s{s} snapshot"


def test_engine_order_fixed_point_same_set_as_load_order():
    st = O.State(O.GemConfig(**CFG), "f32")
    a_p, a_m, _, _ = _engine_run(st, "f32", 2, S.SPIC_ORDER_LOAD)
    b_p, b_m, _, b_stats = _engine_run(st, "f32", 2, S.SPIC_ORDER_ENGINE)
    assert b_stats["positions"] == S.SPIC_POS_FIXED
    for s in range(st.cfg.ns):
        assert np.array_equal(_rows(a_p[s]), _rows(b_p[s])), f"s{s}"
        for n in O.MOMENT_NAMES:
            assert_close(b_m[s][n], a_m[s][n], 1e-5, f"s{s} {n}")


def test_engine_without_ids_holds_less_memory():
    st = O.State(O.GemConfig(**CFG), "f64")
    _, _, _, a = _engine_run(st, "f64", 1, S.SPIC_ORDER_LOAD)
    _, _, _, b = _engine_run(st, "f64", 1, S.SPIC_ORDER_ENGINE, overflow_pct=5)
    assert b["device_bytes"] < a["device_bytes"]


def test_bounded_overflow_runs_the_harris_sheet():
    st = O.State(O.GemConfig(**CFG), "f64")
    S.set_strict_fp(True)
    got_p, got_m, _, stats = _engine_run(st, "f64", 3, S.SPIC_ORDER_LOAD, overflow_pct=1)
    for _ in range(3):
        st.cycle()
    for s in range(st.cfg.ns):
        for k in O.PARTICLE_NAMES:
            assert np.array_equal(got_p[s][k], st.parts[s][k]), f"s{s} {k}"


def _crowd(n):
    """n particles in one cell of the first tile, all moving four cells
    along x in one step (no fields): every one lands in the second tile,
    whose segment holds only its slack."""
    cfg = O.GemConfig(**CFG)
    dx = cfg.Lx / cfg.nxc
    rng = np.random.default_rng(5)
    p = {"x": rng.uniform(0.1, 0.9, n) * dx, "y": rng.uniform(0.1, 0.9, n) * dx,
         "z": rng.uniform(0.1, 0.9, n) * dx}
    p["u"] = np.full(n, 4 * dx / cfg.dt)
    p["v"] = np.zeros(n)
    p["w"] = np.zeros(n)
    p["q"] = np.full(n, -1e-3)
    return cfg, p


@pytest.mark.parametrize("pct,fits", [(1, False), (0, True)])
def test_bounded_overflow_full_reports_enomem(pct, fits):
    cfg, arrays = _crowd(5000)
    grd = S.Grid(cfg.nxc, cfg.nyc, cfg.nzc, cfg.Lx, cfg.Ly, cfg.Lz, "f64")
    zero = np.zeros(grd.nn)
    field = S.EMField(grd, {n: zero for n in O.FIELD_NAMES})
    param = S.Parameters(dt=cfg.dt, c=cfg.c, NiterMover=cfg.NiterMover)
    part = S.Particles(arrays, -64.0, "f64")
    ids = [S.InterpDensSpecies(grd)]
    eng = S.Engine(grd, 1, set_moments=True, overflow_pct=pct)
    try:
        eng.load(0, part)
        if fits:
            eng.phase(field, ids, param)
            eng.store(0, part)
            got = part.download()
            assert np.allclose(got["x"], arrays["x"] + 4 * cfg.Lx / cfg.nxc, rtol=0, atol=1e-12)
        else:
            with pytest.raises(SpicError) as exc:
                eng.phase(field, ids, param)
            assert exc.value.code == SPIC_ENOMEM
    finally:
        eng.close()
