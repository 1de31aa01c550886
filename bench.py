#!/usr/bin/env python3
"""Throughput of sputniPIC's particle hot path (mover_PC + interpP2G) on one
MI355X, FP64 form and FP32 form, with its HBM-roofline fraction.

Workload: the 3-D Harris sheet (GEM) of inputfiles/GEM_3D.inp at the
published run's shape -- 128x64x64 cells on a 20x10x10 box, dt 0.25,
NiterMover 3, four species (qom -64, 1, -64, 1) at 5x5x5 particles per cell
= 65,536,000 per species, 262,144,000 in total, all resident in HBM.  The
fields are frozen at the initial condition (no field solve between cycles,
which is not on this path).  A step is one cycle: push and deposit every
species into freshly zeroed node moments (the engine writes the moments,
which folds the reference driver's zeroing into its reduce).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particles/s through mover_PC+interpP2G, FP32 vs FP64 form; % of HBM roofline"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor peak
GRID = (128, 64, 64)
BOX = (20.0, 10.0, 10.0)
PPC = (5, 5, 5)
NS = 4
# overflow region of the lean engine timed beside the headline, in percent of
# each species' particles: the Harris sheet's phases overflow no segment
LEAN_OVERFLOW_PCT = 2


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(n_particles: int, nn: int, w: int, n_species: int) -> int:
    """Minimum HBM traffic of fused mover + deposit: each particle's
    x,y,z,u,v,w,q read once and x..w written once (13 words), the six
    fields read once, the ten FP64 node moments of each species written
    once."""
    return n_particles * 13 * w + 6 * nn * w + 10 * n_species * nn * 8


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_binned_phase_gem128x64x64_262M_traffic.json")


def measured_traffic(prec: str):
    """HBM bytes per phase launch from the committed FETCH_SIZE / WRITE_SIZE
    profile of this workload (scaled by the copy calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            return float(json.load(f)["bytes_per_launch"][prec])
    except (OSError, KeyError, ValueError):
        return None


class EngineRunner:
    """A fresh engine of the 262M-particle Harris sheet with a given tile
    shape, sort interval and particles per thread, for the autotuner
    (sputnipic_amd.tune)."""

    def __init__(self, prec: str, tile, sort_every: int, ppt: int = 1):
        import sputnipic_amd as S
        self.grd = S.Grid(*GRID, *BOX, prec)
        gem = S.GemParams()
        self.field = S.gem_fields(self.grd, gem)
        self.eng = S.Engine(self.grd, NS, tile=tile, sort_every=sort_every, set_moments=True,
                            particles_per_thread=ppt)
        for s, sp in enumerate(S.harris_species(PPC)):
            p = S.gem_particles(self.grd, gem, sp, s, seed=0)
            self.eng.load(s, p)
            for d in p.d.values():
                d.free()
        self.moms = [S.InterpDensSpecies(self.grd) for _ in range(NS)]
        self.param = S.Parameters(dt=0.25, c=1.0, NiterMover=3)

    def phase(self, stream) -> None:
        self.eng.phase(self.field, self.moms, self.param, stream)

    def close(self) -> None:
        self.eng.close()
        for m in self.moms:
            for d in m.d.values():
                d.free()
        for d in self.field.d.values():
            d.free()


def tune_form(prec: str) -> dict:
    """Tile shape, sort interval and (FP32 form) particles per thread of the
    Harris-sheet engine, chosen by the autotuner on the workload itself
    (each candidate a fresh engine, 1 warm-up and 10 timed phases)."""
    from sputnipic_amd.tune import tune_engine
    t0 = time.perf_counter()
    res = tune_engine(lambda tile, se, ppt=1: EngineRunner(prec, tile, se, ppt), phases=10,
                      warmup=1, ppt_choices=(1, 2) if prec == "f32" else None)
    log(f"{prec}: tuned in {time.perf_counter() - t0:.1f} s: " + ", ".join(
        f"{'x'.join(map(str, t.tile))}/sort {t.sort_every}/ppt {t.ppt}: {t.ms_per_phase:.3f} ms"
        for t in res.trials))
    return res.as_dict()


def run_form(prec: str, steps: int, warmup: int, engine: str, sort_every: int,
             tile, order: str = "load", overflow_pct: int = 0, ppt: int = 1) -> dict:
    import ctypes as C

    import sputnipic_amd as S
    from sputnipic_amd._capi import check, lib, spic_moments

    L = lib()
    w = 8 if prec == "f64" else 4
    grd = S.Grid(*GRID, *BOX, prec)
    gem = S.GemParams()
    field = S.gem_fields(grd, gem)
    species = S.harris_species(PPC)
    parts = [S.gem_particles(grd, gem, sp, s, seed=0) for s, sp in enumerate(species)]
    moms = [S.InterpDensSpecies(grd) for _ in range(NS)]
    n_total = sum(p.npmax for p in parts)
    counts = [p.npmax for p in parts]
    log(f"{prec}: {n_total} particles generated, grid {grd.shape}, engine {engine}")
    param = S.Parameters(dt=0.25, c=1.0, NiterMover=3)
    stream = S.Stream()
    g_c = grd.c_struct()
    f_c = field.c_struct()
    p_c = param.push()
    m_c = [m.c_struct() for m in moms]
    m_arr = (spic_moments * NS)(*m_c)
    eng = None
    if engine == "tiled":
        # the phase writes the moments, so a cycle needs no zeroing pass
        # order "engine": no per-particle load index (particles stay in the
        # engine's tile order, as the reference's arrays after particle_sort)
        # overflow_pct: the overflow region in percent of the species'
        # particles (0: all of them, a phase can never run out)
        eng = S.Engine(grd, NS, tile=tile, sort_every=0, set_moments=True,
                       order=S.SPIC_ORDER_ENGINE if order == "engine" else S.SPIC_ORDER_LOAD,
                       overflow_pct=overflow_pct, particles_per_thread=ppt)
        t0 = time.perf_counter()
        for s, p in enumerate(parts):
            eng.load(s, p, stream)
        stream.synchronize()
        log(f"{prec}: loaded + binned in {time.perf_counter() - t0:.2f} s")
        for p in parts:
            for d in p.d.values():
                d.free()
        parts = None
        kernel = (f"k_binned_phase<{prec}, {tile[0]}x{tile[1]}x{tile[2]}> (all species, one "
                  f"launch; timed with its overflow, reduce and count-reset launches; "
                  f"particle order: {order})")
    else:
        s_c = [p.c_struct() for p in parts]
        kernel = "k_mover_deposit (plain fused, one launch per species)"
    S.set_async(True)
    marks = []  # (kind, start event, end event) of the timed cycles
    pool = [S.Event() for _ in range(2 * (steps + 1) * (NS + 1))]
    state = {"phase": 0, "ev": 0}

    def pair():
        i = state["ev"]
        state["ev"] = i + 2
        return pool[i % len(pool)], pool[(i + 1) % len(pool)]

    def cycle(timed: bool):
        if eng is None:
            for s in range(NS):
                check(L.spic_zero_moments(C.byref(g_c), C.byref(m_c[s]), stream.ptr),
                      "zero")
        if eng is not None:
            due = sort_every > 0 and state["phase"] > 0 and state["phase"] % sort_every == 0
            if due:
                a, b = pair()
                a.record(stream)
                eng.sort(stream)
                b.record(stream)
                if timed:
                    marks.append(("sort", a, b))
            a, b = pair()
            a.record(stream)
            check(L.spic_engine_phase(eng.ptr, C.byref(f_c), m_arr, C.byref(p_c), stream.ptr),
                  "engine phase")
            b.record(stream)
            if timed:
                marks.append(("kernel", a, b))
            state["phase"] += 1
        else:
            for s in range(NS):
                a, b = pair()
                a.record(stream)
                check(L.spic_mover_interp(C.byref(g_c), C.byref(f_c), C.byref(s_c[s]),
                                          C.byref(m_c[s]), C.byref(p_c), stream.ptr),
                      "mover_interp")
                b.record(stream)
                if timed:
                    marks.append(("kernel", a, b))

    try:
        for w_ in range(warmup):
            cycle(False)
        stream.synchronize()
        check(L.spic_check_domain(stream.ptr), "domain check after warmup")
        state["ev"] = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            cycle(True)
        stream.synchronize()
        t1 = time.perf_counter()
        check(L.spic_check_domain(stream.ptr), "domain check")
    finally:
        S.set_async(False)
    wall = t1 - t0
    kernel_ms = [a.elapsed_ms(b) for k, a, b in marks if k == "kernel"]
    sort_ms = [a.elapsed_ms(b) for k, a, b in marks if k == "sort"]
    # charge conservation on the last cycle: sum(rhon)/invVOL == sum(q)
    q_err = 0.0
    q_tot = [0.0] * NS
    if eng is not None:
        # charges are constant: regenerate the per-species totals from the IC
        tmp = [S.gem_particles(grd, gem, sp, s, seed=0) for s, sp in enumerate(species)]
        for s in range(NS):
            q_tot[s] = (float(tmp[s].d["q"].download().astype("float64").sum()),
                        float(abs(tmp[s].d["q"].download().astype("float64")).sum()))
        del tmp
    else:
        for s in range(NS):
            q = parts[s].d["q"].download().astype("float64")
            q_tot[s] = (float(q.sum()), float(abs(q).sum()))
    for s in range(NS):
        rho = moms[s].d["rhon"].download()
        tot = float(rho.sum()) / grd.invVOL
        q_err = max(q_err, abs(tot - q_tot[s][0]) / q_tot[s][1])
    avg_kernel_ms = sum(kernel_ms) / len(kernel_ms)
    if eng is not None:
        per_launch_bytes = algorithmic_bytes(n_total, grd.nn, w, NS)
    else:
        per_launch_bytes = sum(algorithmic_bytes(n, grd.nn, w, 1) for n in counts) / NS
    achieved = per_launch_bytes / (avg_kernel_ms * 1e-3) / 1e9
    value = n_total * steps / wall
    log(f"{prec}: {value / 1e9:.3f} G particles/s, {1e3 * wall / steps:.3f} ms/cycle, "
        f"kernel {avg_kernel_ms:.3f} ms avg, {achieved:.1f} GB/s "
        f"({100 * achieved / HBM_PEAK_GBS:.1f}% of HBM peak), sorts {len(sort_ms)} "
        f"({(sum(sort_ms) / len(sort_ms)) if sort_ms else 0:.2f} ms each), "
        f"charge err {q_err:.1e}")
    out = {
        "value": value,
        "ms_per_step": 1e3 * wall / steps,
        "particles": n_total,
        "kernel": kernel,
        "avg_kernel_ms": avg_kernel_ms,
        "median_kernel_ms": sorted(kernel_ms)[len(kernel_ms) // 2],
        "kernel_ms_range": [min(kernel_ms), max(kernel_ms)],
        "bytes_per_launch": per_launch_bytes,
        "achieved_GBs": achieved,
        "roofline_frac": achieved / HBM_PEAK_GBS,
        "sorts_timed": len(sort_ms),
        "avg_sort_ms": (sum(sort_ms) / len(sort_ms)) if sort_ms else None,
        "sort_every": sort_every if eng is not None else None,
        "charge_rel_err": q_err,
    }
    if eng is not None:
        st = eng.stats()
        out["engine_device_bytes"] = st["device_bytes"]
        out["overflow_pct"] = overflow_pct
        out["engine_stats"] = {k: st[k] for k in ("phases", "sorts", "lds_gathers",
                                                  "global_gathers", "lds_deposits",
                                                  "global_deposits", "overflow")}
        log(f"{prec}: engine stats {out['engine_stats']}")
        eng.close()
    else:
        for p in parts:
            for d in p.d.values():
                d.free()
    return out


STRESS_CYCLES = 50
STRESS_SORT_EVERY = 5


def run_stress(warmup: int, sort_mode: int = 0) -> dict:
    """The combined stress shape: the 128x64x64 Harris sheet with its four
    species at 5x5x5 per cell plus four uniform beams (electrons and ions,
    x drift +-0.1, thermal speed 0.02) at 1x5x5 = 25 per cell, 314,572,800
    particles in FP32; the engine rebuilds its tile segments every 5 phases;
    50 cycles (engine phases that write the moments) are captured in one HIP graph
    and timed as one launch.  Re-sorts are tile sorts (sort_mode 0) or radix
    sorts (1)."""
    import ctypes as C

    import sputnipic_amd as S
    from sputnipic_amd._capi import check, lib

    L = lib()
    grd = S.Grid(*GRID, *BOX, "f32")
    gem = S.GemParams()
    field = S.gem_fields(grd, gem)
    species = S.harris_species(PPC) + S.beam_species()
    ns = len(species)
    eng = S.Engine(grd, ns, sort_every=STRESS_SORT_EVERY, set_moments=True,
                   sort_mode=sort_mode)
    stream = S.Stream()
    n_total = 0
    for s, sp in enumerate(species):
        p = S.gem_particles(grd, gem, sp, s, seed=0)
        n_total += p.npmax
        eng.load(s, p, stream)
        stream.synchronize()
        for d in p.d.values():
            d.free()
    log(f"stress: {n_total} particles in {ns} species, FP32, re-sorted every "
        f"{STRESS_SORT_EVERY} phases, sort mode {sort_mode}")
    moms = [S.InterpDensSpecies(grd) for _ in range(ns)]
    param = S.Parameters(dt=0.25, c=1.0, NiterMover=3)

    def cycle():
        eng.phase(field, moms, param, stream)

    S.set_async(True)
    try:
        for _ in range(warmup):
            cycle()
        stream.synchronize()
        check(L.spic_check_domain(stream.ptr), "domain check after warmup")
        S.set_async(True)
        t0 = time.perf_counter()
        with S.Graph(stream) as graph:
            for _ in range(STRESS_CYCLES):
                cycle()
        t_capture = time.perf_counter() - t0
        a, b = S.Event(), S.Event()
        stream.synchronize()
        t0 = time.perf_counter()
        a.record(stream)
        graph.launch()
        b.record(stream)
        stream.synchronize()
        wall = time.perf_counter() - t0
        check(L.spic_check_domain(stream.ptr), "domain check")
    finally:
        S.set_async(False)
    gpu_ms = a.elapsed_ms(b)
    st = eng.stats()
    graph.close()
    eng.close()
    value = n_total * STRESS_CYCLES / wall
    per_cycle_bytes = algorithmic_bytes(n_total, grd.nn, 4, ns)
    achieved = per_cycle_bytes * STRESS_CYCLES / (gpu_ms * 1e-3) / 1e9
    log(f"stress: {value / 1e9:.3f} G particles/s, {1e3 * wall / STRESS_CYCLES:.3f} "
        f"ms/cycle ({gpu_ms / STRESS_CYCLES:.3f} ms/cycle on the GPU clock), "
        f"{achieved:.1f} GB/s, capture {t_capture:.2f} s, stats {st}")
    return {
        "value": value,
        "ms_per_step": 1e3 * wall / STRESS_CYCLES,
        "particles": n_total,
        "gpu_ms_per_cycle": gpu_ms / STRESS_CYCLES,
        "achieved_GBs": achieved,
        "roofline_frac": achieved / HBM_PEAK_GBS,
        "bytes_per_cycle": per_cycle_bytes,
        "engine_stats": {k: st[k] for k in ("phases", "sorts", "global_deposits",
                                            "overflow")},
    }


BATCH_INSTANCES = 16
BATCH_SEED = 3
BATCH_SCALES = (2, 3, 4, 5, 6)  # grid (16 s) x (8 s) x (8 s): 32x16x16 .. 96x48x48
# (npcelx, npcely): the generator's lattice loop runs its z index to npcely,
# as the reference's does (src/IC.cpp:128), so a cell holds npcelx * npcely^2
# particles: 1 .. 64
BATCH_PPC = tuple((px, py) for px in range(1, 5) for py in range(1, 5))
BATCH_N_RANGE = (500_000, 60_000_000)


def batch_specs(seed: int = BATCH_SEED):
    """The 16 independent Harris-sheet instances of the batching form: per
    instance a grid scale, a lattice per species (1..64 particles per cell),
    drawn again until the instance holds 0.5M..60M particles, a
    perturbation amplitude in [0.005, 0.05] and its own generator seed."""
    import numpy as np
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(BATCH_INSTANCES):
        while True:
            sc = int(rng.choice(BATCH_SCALES))
            grid = (16 * sc, 8 * sc, 8 * sc)
            ppc = [BATCH_PPC[int(rng.integers(len(BATCH_PPC)))] for _ in range(NS)]
            cells = grid[0] * grid[1] * grid[2]
            n = sum(cells * px * py * py for px, py in ppc)
            if BATCH_N_RANGE[0] <= n <= BATCH_N_RANGE[1]:
                break
        specs.append({"grid": grid, "ppc": ppc, "pertX": float(rng.uniform(0.005, 0.05)),
                      "seed": k + 1, "particles": n})
    return specs


class GemInstance:
    """One instance of the batching form, generated on the device: FP32
    fields and particles, an engine with the given tile shape."""

    def __init__(self, spec: dict, tile, stream):
        import sputnipic_amd as S
        self.grd = S.Grid(*spec["grid"], *BOX, "f32")
        gem = S.GemParams(pertX=spec["pertX"])
        self.field = S.gem_fields(self.grd, gem)
        species = S.harris_species()
        for sp, (px, py) in zip(species, spec["ppc"]):
            sp.npcelx, sp.npcely, sp.npcelz = px, py, py
        self.eng = S.Engine(self.grd, len(species), tile=tile)
        self.n = 0
        for s, sp in enumerate(species):
            p = S.gem_particles(self.grd, gem, sp, s, seed=spec["seed"])
            self.n += p.npmax
            self.eng.load(s, p, stream)
            stream.synchronize()
            for d in p.d.values():
                d.free()
        self.ids = [S.InterpDensSpecies(self.grd) for _ in species]
        self.param = S.Parameters(dt=0.25, c=1.0, NiterMover=3)
        self.ns = len(species)

    def member(self):
        return (self.eng, self.field, self.ids, self.param)

    def zero(self, stream):
        import sputnipic_amd as S
        S.setZeroSpeciesDensities(self.ids, self.grd, self.ns, stream)

    def close(self):
        self.eng.close()
        for m in self.ids:
            for d in m.d.values():
                d.free()


def run_batch(steps: int, warmup: int) -> dict:
    """The batching form: 16 independent Harris-sheet instances (batch_specs)
    in FP32, advanced by one batch (one launch per kernel and phase); tile
    shape and sort schedule from the autotuner, timed on the same instances
    built afresh; then `steps` cycles captured in one HIP graph and timed as
    one launch.  A cycle is one batch phase that writes every instance's
    moments (the zeroing before the deposit folded into the reduce)."""
    import sputnipic_amd as S
    from sputnipic_amd.tune import tune_batch

    specs = batch_specs()
    n_total = sum(sp["particles"] for sp in specs)
    log(f"batch: {len(specs)} instances, {n_total} particles, grids "
        f"{sorted(set(sp['grid'] for sp in specs))}")
    stream = S.Stream()
    t0 = time.perf_counter()
    tuned = tune_batch(lambda tile: [GemInstance(sp, tile, stream) for sp in specs],
                       phases=10, warmup=1, stream=stream, set_moments=True)
    t_tune = time.perf_counter() - t0
    log(f"batch: tuned in {t_tune:.1f} s: " + ", ".join(
        f"{'x'.join(map(str, t.tile))}/sort {t.sort_every}: {t.ms_per_phase:.3f} ms"
        for t in tuned.trials))
    insts = [GemInstance(sp, tuned.tile, stream) for sp in specs]
    assert sum(x.n for x in insts) == n_total
    batch = S.Batch([x.member() for x in insts], sort_every=tuned.sort_every,
                    set_moments=True)

    def cycle():
        batch.phase(stream)

    S.set_async(True)
    try:
        for _ in range(warmup):
            cycle()
        stream.synchronize()
        S.check_domain(stream)
        S.set_async(True)
        with S.Graph(stream) as graph:
            for _ in range(steps):
                cycle()
        a, b = S.Event(), S.Event()
        stream.synchronize()
        t0 = time.perf_counter()
        a.record(stream)
        graph.launch()
        b.record(stream)
        stream.synchronize()
        wall = time.perf_counter() - t0
        S.check_domain(stream)
    finally:
        S.set_async(False)
    gpu_ms = a.elapsed_ms(b)
    stats = [x.eng.stats() for x in insts]
    graph.close()
    batch.close()
    for x in insts:
        x.close()
    per_cycle_bytes = sum(algorithmic_bytes(x.n, x.grd.nn, 4, x.ns) for x in insts)
    value = n_total * steps / wall
    achieved = per_cycle_bytes * steps / (gpu_ms * 1e-3) / 1e9
    log(f"batch: {value / 1e9:.3f} G particles/s, {1e3 * wall / steps:.3f} ms/cycle "
        f"({gpu_ms / steps:.3f} ms/cycle on the GPU clock), {achieved:.1f} GB/s "
        f"({100 * achieved / HBM_PEAK_GBS:.1f}% of HBM peak)")
    return {
        "value": value,
        "ms_per_step": 1e3 * wall / steps,
        "gpu_ms_per_cycle": gpu_ms / steps,
        "particles": n_total,
        "instances": [{"grid": "x".join(map(str, sp["grid"])),
                       "ppc": [px * py * py for px, py in sp["ppc"]],
                       "pertX": round(sp["pertX"], 5), "particles": sp["particles"]}
                      for sp in specs],
        "tuned": tuned.as_dict(),
        "tune_seconds": t_tune,
        "achieved_GBs": achieved,
        "roofline_frac": achieved / HBM_PEAK_GBS,
        "bytes_per_cycle": per_cycle_bytes,
        "overflow": sum(st["overflow"] for st in stats),
        "global_deposits": sum(st["global_deposits"] for st in stats),
    }


def run_order(prec: str, phases: int = 3) -> dict:
    """Input order: the same particles generated in initial-condition order
    (cell-major, as the reference's initGEM lays them out) and randomly
    permuted on the device (an affine bijection of the indices).  For each:
    the engine's load (copy in and bin by tile) and its first phases, and
    the plain caller-order fused kernel (mover_interp, global atomics) on
    species 0.  The engine's binning makes its phases independent of the
    input order; the caller-order kernel is not."""
    import ctypes as C

    import sputnipic_amd as S
    from sputnipic_amd._capi import check, lib

    L = lib()
    grd = S.Grid(*GRID, *BOX, prec)
    gem = S.GemParams()
    field = S.gem_fields(grd, gem)
    param = S.Parameters(dt=0.25, c=1.0, NiterMover=3)
    stream = S.Stream()
    out = {}
    g_c = grd.c_struct()
    for order in ("ic_order", "permuted"):
        parts = []
        for s_, sp in enumerate(S.harris_species(PPC)):
            p = S.gem_particles(grd, gem, sp, s_, seed=0)
            if order == "permuted":
                q = S.DeviceParticles(p.npmax, p.qom, prec)
                a, b = 2654435761 % p.npmax, 12345
                while np_gcd(a, p.npmax) != 1:
                    a += 1
                check(L.spic_permute_species(C.byref(g_c), C.byref(p.c_struct()),
                                             C.byref(q.c_struct()), a, b, None), "permute")
                for d in p.d.values():
                    d.free()
                p = q
            parts.append(p)
        # caller-order fused kernel on species 0
        moms0 = S.InterpDensSpecies(grd)
        ev = [S.Event(), S.Event()]
        S.set_async(True)
        try:
            ev[0].record(stream)
            check(L.spic_mover_interp(C.byref(g_c), C.byref(field.c_struct()),
                                      C.byref(parts[0].c_struct()), C.byref(moms0.c_struct()),
                                      C.byref(param.push()), stream.ptr), "mover_interp")
            ev[1].record(stream)
            stream.synchronize()
            check(L.spic_check_domain(stream.ptr), "domain")
        finally:
            S.set_async(False)
        plain_ms = ev[0].elapsed_ms(ev[1])
        for d in moms0.d.values():
            d.free()
        eng = S.Engine(grd, NS, set_moments=True)
        t0 = time.perf_counter()
        for s_, p in enumerate(parts):
            eng.load(s_, p, stream)
        stream.synchronize()
        t_load = time.perf_counter() - t0
        for p in parts:
            for d in p.d.values():
                d.free()
        moms = [S.InterpDensSpecies(grd) for _ in range(NS)]
        ms = []
        S.set_async(True)
        try:
            for _ in range(phases):
                a_, b_ = S.Event(), S.Event()
                a_.record(stream)
                eng.phase(field, moms, param, stream)
                b_.record(stream)
                stream.synchronize()
                ms.append(a_.elapsed_ms(b_))
            S.check_domain(stream)
        finally:
            S.set_async(False)
        eng.close()
        for m in moms:
            for d in m.d.values():
                d.free()
        out[order] = {"engine_load_and_bin_s": t_load, "engine_phase_ms": ms,
                      "plain_mover_interp_species0_ms": plain_ms}
        log(f"{prec} {order}: engine load+bin {t_load:.3f} s, phases "
            + ", ".join(f"{v:.2f}" for v in ms) + f" ms; caller-order mover_interp of "
            f"species 0 {plain_ms:.1f} ms")
    return out


def np_gcd(a: int, b: int) -> int:
    while b:
        a, b = b, a % b
    return a


def field_bytes(grd, ns: int):
    """Algorithmic bytes of the field-solve operators on grd (FP64): the
    net rhoc step + hat densities read the net rhon, every species' J and
    pressure (9 node arrays), B and rhoc and write rhoc, J hat and rho hat;
    MaxwellImage reads the Krylov vector, B and every species' rhon and
    writes the image."""
    nn = grd.nn
    nc = grd.nxc * grd.nyc * grd.nzc
    m = (grd.nxn - 2) * (grd.nyn - 2) * (grd.nzn - 2)
    hat = 8 * ((7 + 9 * ns) * nn + 3 * nc)
    image = 8 * (6 * m + (3 + ns) * nn)
    return hat, image


def run_field(reps: int, cpu: bool) -> dict:
    """The field-solve consumers of the deposit on the 128x64x64 grid, FP64:
    the driver's net-rhoc step with calculateHatDensities, and MaxwellImage
    (one GMRES operator application), timed with HIP events over `reps`
    calls each; with `cpu`, the reference's own routines (oracle/_ref, -O3)
    on the host cores at the same grid."""
    import numpy as np

    import sputnipic_amd as S

    grd = S.Grid(*GRID, *BOX, "f64")
    field = S.gem_fields(grd, S.GemParams())
    rng = np.random.default_rng(1)
    ids = [S.InterpDensSpecies(grd) for _ in range(NS)]
    for m in ids:
        m.upload({n: rng.normal(0, 1e-2, grd.nn) for n in S.MOMENT_NAMES})
    net = S.InterpDensNet(grd)
    S.sumOverSpecies(net, ids, grd, NS)
    sp = S.SolveParams()
    aux = S.InterpDensAux(grd)
    nvec = 3 * (grd.nxn - 2) * (grd.nyn - 2) * (grd.nzn - 2)
    vec = S.DeviceArray.from_numpy(rng.normal(0, 1e-2, nvec))
    im = S.DeviceArray(nvec, np.float64)
    stream = S.Stream()

    def hat():
        S.interpN2Crho_net(net, grd, sp)
        S.calculateHatDensities(aux, net, ids, field, grd, sp, stream)

    def image():
        S.MaxwellImage(im, vec, field, ids, grd, sp, stream)

    out = {"grid": "x".join(map(str, GRID)), "species": NS, "reps": reps}
    hat_b, image_b = field_bytes(grd, NS)
    S.set_async(True)
    try:
        for name, fn, nbytes in (("hat_densities", hat, hat_b), ("maxwell_image", image, image_b)):
            fn()
            stream.synchronize()
            a, b = S.Event(), S.Event()
            a.record(stream)
            for _ in range(reps):
                fn()
            b.record(stream)
            stream.synchronize()
            ms = a.elapsed_ms(b) / reps
            gbs = nbytes / (ms * 1e-3) / 1e9
            out[name] = {"gpu_ms": ms, "algorithmic_bytes": nbytes, "achieved_GBs": gbs,
                         "roofline_frac": gbs / HBM_PEAK_GBS}
        S.check_domain(stream)
    finally:
        S.set_async(False)
    if cpu:
        import oracle as O
        if O.ref_available("f64_O3"):
            cfg = O.GemConfig(nxc=GRID[0], nyc=GRID[1], nzc=GRID[2], Lx=BOX[0], Ly=BOX[1],
                              Lz=BOX[2], npcelx=[1] * NS, npcely=[1] * NS, npcelz=[1] * NS)
            ref = O.RefState(cfg, "f64_O3")
            creps = 3
            th, ti = ref.time_field(creps)
            out["hat_densities"]["cpu_reference_ms"] = 1e3 * th / creps
            out["maxwell_image"]["cpu_reference_ms"] = 1e3 * ti / creps
            out["cpu_threads"] = ref.max_threads()
            ref.close()
    log("field solve: " + ", ".join(
        f"{k} {v['gpu_ms']:.3f} ms ({v['achieved_GBs']:.0f} GB/s"
        + (f"; reference CPU {v['cpu_reference_ms']:.1f} ms" if "cpu_reference_ms" in v else "")
        + ")" for k, v in out.items() if isinstance(v, dict)))
    return out


class _StdoutToStderr:
    """Route the process's C-level stdout to stderr, so that text the CPU
    baseline's compiled code prints (the reference's initGEM writes a banner
    with std::cout) never mixes into the one JSON line on stdout."""

    def __enter__(self):
        import ctypes
        self.libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self.libc.fflush(None)
        self.saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        sys.stdout.flush()
        self.libc.fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def copy_bandwidth(nbytes: int = 4 << 30, reps: int = 5) -> dict:
    """A plain device-to-device copy of nbytes (hipMemcpy), timed with HIP
    events on the null stream; GB/s counts the bytes read plus the bytes
    written.  The median of `reps` copies, after one untimed copy."""
    import sputnipic_amd as S
    from sputnipic_amd._capi import check, lib

    n = nbytes // 4
    a = S.DeviceArray(n, "float32", fill_zero=True)
    b = S.DeviceArray(n, "float32")
    times = []
    try:
        for r in range(reps + 1):
            e0, e1 = S.Event(), S.Event()
            e0.record(None)
            check(lib().spic_memcpy_d2d(b.ptr, a.ptr, n * 4), "copy")
            e1.record(None)
            check(lib().spic_synchronize(), "synchronize")
            if r:
                times.append(e0.elapsed_ms(e1))
    finally:
        a.free()
        b.free()
    ms = sorted(times)[len(times) // 2]
    gbs = 2 * n * 4 / (ms * 1e-3) / 1e9
    log(f"device copy: {n * 4} bytes in {ms:.3f} ms (median of {reps}): {gbs:.0f} GB/s "
        "read + write")
    return {"bytes": n * 4, "ms": ms, "GBs": gbs, "reps": reps}


def host_cpu() -> dict:
    """The host's CPU model and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


REF_CYCLES = 10


def moment_errors(got, want, shape, wall_planes: int = 3) -> dict:
    """Per moment over the species: max |got - want| / max |want| over all
    nodes, the same without the `wall_planes` node planes next to each
    reflecting y wall, and ||got - want||_2 / ||want||_2 over all nodes."""
    import numpy as np

    out = {}
    for n in want[0]:
        num = num_in = den = l2n = l2d = 0.0
        for g, w in zip(got, want):
            d = np.abs(g[n] - w[n])
            num = max(num, float(d.max()))
            d3 = d.reshape(shape)[:, wall_planes:shape[1] - wall_planes, :]
            num_in = max(num_in, float(d3.max()))
            den = max(den, float(np.abs(w[n]).max()))
            l2n += float(np.dot(d, d))
            l2d += float(np.dot(w[n], w[n]))
        out[n] = {"max_rel": num / den if den else num,
                  "max_rel_interior": num_in / den if den else num_in,
                  "l2_rel": (l2n / l2d) ** 0.5 if l2d else l2n ** 0.5}
    return out


class ParticleErrors:
    """Per particle component, accumulated over species: max and RMS of
    |got - want| over the box length (x and z modulo their period) or the
    largest |want| of the velocity component, and the particles whose
    velocity component has the other sign (reflected in one run only)."""

    def __init__(self, vmax):
        self.vmax = vmax
        self.acc = {k: {"max": 0.0, "ss": 0.0, "n": 0, "flips": 0} for k in "xyzuvw"}

    def add(self, got, want):
        import numpy as np
        for k in "xyzuvw":
            g = got[k].astype(np.float64)
            d = np.abs(g - want[k])
            if k in "xyz":
                L = BOX["xyz".index(k)]
                if k != "y":
                    d = np.minimum(d, L - d)
                scale = L
            else:
                scale = self.vmax[k]
                self.acc[k]["flips"] += int(np.count_nonzero(
                    (np.sign(g) * np.sign(want[k]) < 0) & (d > 1e-3 * scale)))
            a = self.acc[k]
            a["max"] = max(a["max"], float(d.max()) / scale)
            a["ss"] += float(np.dot(d, d)) / scale ** 2
            a["n"] += d.size

    def result(self):
        out = {}
        for k, a in self.acc.items():
            out[k] = {"max": a["max"], "rms": (a["ss"] / max(a["n"], 1)) ** 0.5}
            if k in "uvw":
                out[k]["flips"] = a["flips"]
        return out


def reference_leg(cycles: int = REF_CYCLES) -> dict:
    """The CPU reference on the benchmark's own workload, and the GPU forms
    against it.

    The 262,144,000-particle Harris sheet is generated on the device in FP64
    (the same generator as the timed runs) and copied to the host, into the
    reference's own state (oracle/_ref/libref_f64_O3.so: src/Mover.cpp etc.
    built at the optimisation level of the reference's Makefile, FP64,
    without initGEM; the oracle restatement where that build is absent).
    One cycle of its mover_PC + zero + interpP2G with the reference
    driver's OpenMP structure is the timed CPU baseline; it then runs to
    `cycles` cycles.  The FP64 form and the FP32 form (fields rounded to
    float; the FP64 particle arrays loaded into fixed-point positions and
    float velocities and charges) run `cycles` engine phases on the device
    from the same initial condition, are stored back into FP64 arrays and
    compared with it after the ghost-node fold (applyBCids on both sides):
    node moments per moment, particles per component in load order."""
    import numpy as np

    import oracle as O
    import sputnipic_amd as S

    S.set_async(False)
    S.set_strict_fp(False)
    grd64 = S.Grid(*GRID, *BOX, "f64")
    gem = S.GemParams()
    field64 = S.gem_fields(grd64, gem)
    species = S.harris_species(PPC)
    ic = [S.gem_particles(grd64, gem, sp, s, seed=0) for s, sp in enumerate(species)]
    counts = [p.npmax for p in ic]
    n = sum(counts)
    cfg = O.GemConfig(nxc=GRID[0], nyc=GRID[1], nzc=GRID[2], Lx=BOX[0], Ly=BOX[1],
                      Lz=BOX[2], npcelx=[PPC[0]] * NS, npcely=[PPC[1]] * NS,
                      npcelz=[PPC[2]] * NS)
    fields = {f: field64.d[f].download() for f in O.FIELD_NAMES}
    if O.ref_available("f64_O3"):
        kind = "reference"
        ref = O.RefState(cfg, "f64_O3", ic=False)
        for s, p in enumerate(ic):
            assert ref.nop(s) == counts[s]
            ref.set_particles(s, p.download())
        for w, f in enumerate(O.FIELD_NAMES):
            ref.set_field(w, fields[f])
        cores = ref.max_threads()
        timer = ref.time_cycles
    else:
        kind = "port"
        ref = O.State(cfg, "f64", build_ic=False)
        for f in O.FIELD_NAMES:
            ref.fields[f][:] = fields[f]
        ref.parts = [p.download() for p in ic]
        cores = O.State.max_threads()
        timer = ref.time_cycles
    log(f"cpu {kind}: {n} particles on the host ({cores} threads)")
    t1, tm1, td1 = timer(1)
    log(f"cpu {kind}: cycle 1 in {t1:.2f} s (mover {tm1:.2f}, deposit {td1:.2f}): "
        f"{n / t1 / 1e6:.1f} M particles/s")
    t_rest = timer(cycles - 1)[0] if cycles > 1 else 0.0
    log(f"cpu {kind}: cycles 2..{cycles} in {t_rest:.1f} s")
    if kind == "reference":
        for s in range(NS):
            ref.apply_bc_ids(s)
        want_m = [{m: ref.moment(s, w) for w, m in enumerate(O.MOMENT_NAMES)}
                  for s in range(NS)]
        want_p = lambda s: ref.particles(s)  # noqa: E731
    else:
        for s in range(NS):
            ref.apply_bc_ids(s)
        want_m = ref.moments
        want_p = lambda s: ref.parts[s]  # noqa: E731
    vmax = {}
    for s in range(NS):
        p = want_p(s)
        for k in "uvw":
            vmax[k] = max(vmax.get(k, 0.0), float(np.abs(p[k]).max()))
        del p
    param = S.Parameters(dt=0.25, c=1.0, NiterMover=3)
    errors = {}
    for prec in ("f64", "f32"):
        grd = S.Grid(*GRID, *BOX, prec)
        if prec == "f64":
            field = field64
        else:
            field = S.EMField(grd, {f: fields[f].astype(np.float32) for f in O.FIELD_NAMES})
        eng = S.Engine(grd, NS, set_moments=True)
        for s, p in enumerate(ic):
            eng.load(s, p)  # the FP32 form takes the FP64 arrays (spic_engine_load_as)
        moms = [S.InterpDensSpecies(grd) for _ in range(NS)]
        for _ in range(cycles):
            eng.phase(field, moms, param)
        S.applyBCids_all(moms, grd, param)
        got_m = [m.download() for m in moms]
        me = moment_errors(got_m, want_m, (grd.nxn, grd.nyn, grd.nzn))
        del got_m
        pe = ParticleErrors(vmax)
        for s in range(NS):
            out = S.DeviceParticles(counts[s], species[s].qom, "f64")
            eng.store(s, out)  # into FP64 arrays from either form
            got = {k: out.d[k].download() for k in "xyzuvw"}
            for d in out.d.values():
                d.free()
            pe.add(got, want_p(s))
            del got
        eng.close()
        for m in moms:
            for d in m.d.values():
                d.free()
        if prec == "f32":
            for d in field.d.values():
                d.free()
        errors[prec] = {"moments": me, "particles": pe.result()}
        log(f"{prec} against the {kind} after {cycles} cycles: moments max rel "
            + ", ".join(f"{k} {v['max_rel']:.1e}" for k, v in me.items())
            + "; interior " + ", ".join(f"{k} {v['max_rel_interior']:.1e}" for k, v in me.items())
            + "; L2 " + ", ".join(f"{k} {v['l2_rel']:.1e}" for k, v in me.items()))
        log(f"{prec} against the {kind}: particles (max, rms, flips) " + ", ".join(
            f"{k} {v['max']:.1e} {v['rms']:.1e} {v.get('flips', '')}"
            for k, v in errors[prec]["particles"].items()))
    for p in ic:
        for d in p.d.values():
            d.free()
    if kind == "reference":
        ref.close()
    cpu = {
        "value": n / t1,
        "unit": "particles/s",
        "cores": cores,
        "kind": kind,
        "sample": f"the benchmark's workload itself: 128x64x64 Harris sheet, 4 species x "
                  f"5x5x5 per cell = {n} particles, FP64, one cycle of mover_PC + zero + "
                  f"interpP2G (OpenMP over particles in the mover, over species in the "
                  f"deposit), from the device generator's initial condition",
        "seconds": t1,
        "mover_seconds": tm1,
        "deposit_seconds": td1,
        "host": host_cpu(),
    }
    return {"cpu_baseline": cpu, "errors": errors, "cycles": cycles, "kind": kind,
            "cpu_seconds_all_cycles": t1 + t_rest}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ref-cycles", type=int, default=REF_CYCLES,
                    help="cycles of the CPU reference leg (the first one is the timed "
                         "baseline; the GPU forms are compared with it after all of them)")
    ap.add_argument("--forms", default="f64,f32")
    ap.add_argument("--engine", choices=("tiled", "plain"), default="tiled")
    ap.add_argument("--sort-every", type=int, default=0)
    ap.add_argument("--tile", default="4x4x4")
    ap.add_argument("--ppt", type=int, choices=(1, 2), default=1,
                    help="particles per thread and batch of the FP32 form's phase "
                         "(with --no-tune; the autotuner otherwise chooses)")
    ap.add_argument("--order", choices=("load", "engine"), default="load",
                    help="engine particle order of the timed forms: load (a load index per "
                         "particle, store in load order; default) or engine (no index)")
    ap.add_argument("--no-tune", action="store_true",
                    help="time the forms with --tile and --sort-every instead of the "
                         "autotuner's choice")
    ap.add_argument("--no-engine-order", action="store_true",
                    help="skip the extra timing of both forms without the load index")
    ap.add_argument("--sort-mode", type=int, choices=(0, 1), default=0,
                    help="stress: re-sorts by tile sort (0) or radix sort (1)")
    ap.add_argument("--workload", choices=("gem", "stress", "batch", "field"), default="gem",
                    help="gem: the 262M-particle Harris sheet (default); stress: 8 "
                         "species, 314.6M particles, FP32, 50 cycles in one HIP graph; "
                         "batch: 16 independent Harris-sheet instances in FP32, one "
                         "launch per kernel and phase, autotuned, in one HIP graph; "
                         "field: only the field-solve operators on the 128x64x64 grid")
    ap.add_argument("--no-order", action="store_true",
                    help="skip the IC-order against permuted-input comparison")
    ap.add_argument("--no-field", action="store_true",
                    help="skip the field-solve operators (hat densities, MaxwellImage)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES once device to device (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import sputnipic_amd as S
        from sputnipic_amd._capi import check, lib
        n = args.copy_probe // 4
        a = S.DeviceArray(n, "float32", fill_zero=True)
        b = S.DeviceArray(n, "float32")
        check(lib().spic_memcpy_d2d(b.ptr, a.ptr, n * 4), "copy probe")
        print(f"copy probe: {n * 4} bytes read and {n * 4} bytes written", file=sys.stderr)
        return 0
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus}: only one GPU is measured; the path would "
              "shard by particles (one moment all-reduce per cycle) or by x-slabs (halo "
              "folds and particle migration), which is not built.", file=sys.stderr)
        return 2
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: multi-rank launch is not supported (one GPU only)", file=sys.stderr)
        return 2
    import sputnipic_amd as S
    if S.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    if args.workload == "stress":
        f = run_stress(args.warmup, args.sort_mode)
        line = {
            "metric": METRIC,
            "value": f["value"],
            "unit": "particles/s",
            "n_gpus": 1,
            "steps": STRESS_CYCLES,
            "warmup": args.warmup,
            "ms_per_step": f["ms_per_step"],
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic",
            "config": {
                "workload": "stress: Harris-sheet GEM 3D, 128x64x64 cells, 4 Harris species "
                            "at 5x5x5 per cell + 4 x-beams (+-0.1 drift, uth 0.02) at 1x5x5 "
                            "per cell = 314,572,800 particles, FP32, tile segments rebuilt "
                            "every 5 phases, 50 cycles captured in one HIP graph",
                "particles": f["particles"],
                "sort_mode": args.sort_mode,
            },
            "roofline": {
                "bound": "hbm",
                "achieved": f["achieved_GBs"],
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": f["roofline_frac"],
                "traffic": None,
                "kernel": "whole captured cycle (zero, phase, overflow, reduce, sorts)",
                "algorithmic_bytes_per_launch": f["bytes_per_cycle"],
            },
            "stress": f,
            "cpu_baseline": None,
        }
        print(json.dumps(line), flush=True)
        return 0
    if args.workload == "field":
        with _StdoutToStderr():
            f = run_field(args.steps, cpu=not args.no_cpu_baseline)
        print(json.dumps({"metric": "field-solve operators, ms per call", "field_solve": f}),
              flush=True)
        return 0
    if args.workload == "batch":
        f = run_batch(args.steps, args.warmup)
        line = {
            "metric": METRIC,
            "value": f["value"],
            "unit": "particles/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": f["ms_per_step"],
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic",
            "config": {
                "workload": "batch: 16 independent Harris-sheet GEM 3D instances, grids "
                            "32x16x16..96x48x48 cells on a 20x10x10 box, 4 species at 1..64 "
                            "particles per cell (0.5M..60M per instance), pertX in "
                            "[0.005, 0.05], FP32, one launch per kernel and phase, tile "
                            "shape and sort schedule autotuned, cycles in one HIP graph",
                "particles": f["particles"],
            },
            "roofline": {
                "bound": "hbm",
                "achieved": f["achieved_GBs"],
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": f["roofline_frac"],
                "traffic": None,
                "kernel": "whole captured cycle (zero, batched phase, overflow, reduce, "
                          "sorts)",
                "algorithmic_bytes_per_launch": f["bytes_per_cycle"],
            },
            "batch": f,
            "cpu_baseline": None,
        }
        print(json.dumps(line), flush=True)
        return 0
    forms = {}
    tuning = {}
    settings = {}
    for prec in args.forms.split(","):
        tile = tuple(int(t) for t in args.tile.split("x"))
        sort_every = args.sort_every
        ppt = args.ppt if prec == "f32" else 1
        if args.engine == "tiled" and not args.no_tune:
            with _StdoutToStderr():
                tuning[prec] = tune_form(prec)
            tile = tuple(int(t) for t in tuning[prec]["tile"].split("x"))
            sort_every = tuning[prec]["sort_every"]
            ppt = tuning[prec]["particles_per_thread"]
        settings[prec] = (tile, sort_every, ppt)
        forms[prec] = run_form(prec, args.steps, args.warmup, args.engine, sort_every,
                               tile, args.order, ppt=ppt)
        forms[prec]["tile"] = "x".join(map(str, tile))
        forms[prec]["particles_per_thread"] = ppt
    # the same forms without the per-particle load index and with an overflow
    # region of LEAN_OVERFLOW_PCT percent (the smallest engine; reported beside
    # the headline, which keeps the default engine)
    forms_engine_order = {}
    if args.engine == "tiled" and args.order == "load" and not args.no_engine_order:
        for prec in args.forms.split(","):
            tile, sort_every, ppt = settings[prec]
            forms_engine_order[prec] = run_form(prec, args.steps, args.warmup, args.engine,
                                                sort_every, tile, "engine", LEAN_OVERFLOW_PCT,
                                                ppt=ppt)
    primary = "f64" if "f64" in forms else list(forms)[0]
    f = forms[primary]
    cpu = refleg = None
    if not args.no_cpu_baseline:
        with _StdoutToStderr():
            refleg = reference_leg(args.ref_cycles)
        cpu = refleg["cpu_baseline"]
    order = None
    if not args.no_order and args.engine == "tiled":
        with _StdoutToStderr():
            order = run_order("f32" if "f32" in forms else primary)
    field = None
    if not args.no_field:
        with _StdoutToStderr():
            field = run_field(20, cpu=not args.no_cpu_baseline)
    with _StdoutToStderr():
        copy = copy_bandwidth()
    line = {
        "metric": METRIC,
        "value": f["value"],
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": primary,
        "data": "synthetic",
        "config": {
            "workload": "Harris-sheet GEM 3D, 128x64x64 cells, 20x10x10 box, 4 species x "
                        "65,536,000 particles (5x5x5 per cell), dt 0.25, NiterMover 3, "
                        "fields frozen at the initial condition, all particles in HBM",
            "form": "plain FP64" if primary == "f64" else "FP32",
            "particles": f["particles"],
            "cycles_timed": args.steps,
            "tile": f.get("tile"),
            "sort_every": f.get("sort_every"),
            "particles_per_thread": f.get("particles_per_thread"),
            "settings_from": "autotuner" if tuning else "command line",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": f["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": f["roofline_frac"],
            "traffic": measured_traffic(primary) if args.engine == "tiled" else None,
            "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT),
            "kernel": f["kernel"],
            "avg_kernel_ms": f["avg_kernel_ms"],
            "algorithmic_bytes_per_launch": f["bytes_per_launch"],
            "device_copy_GBs": copy["GBs"],
            "frac_of_device_copy": f["achieved_GBs"] / copy["GBs"],
        },
        "device_copy": copy,
        "forms": {k: {kk: vv for kk, vv in v.items()} for k, v in forms.items()},
        "forms_engine_order": forms_engine_order or None,
        "tuning": tuning or None,
        "cpu_baseline": cpu,
        "field_solve": field,
        "input_order": order,
    }
    if "f64" in forms and "f32" in forms:
        line["fp32_over_fp64"] = forms["f32"]["value"] / forms["f64"]["value"]
    if refleg is not None:
        definition = (
            "against the reference's own mover_PC + interpP2G (FP64, CPU) after the same "
            "cycles from the same initial condition (the device generator's, FP64; the "
            "FP32 form gets it rounded to float), frozen fields, production arithmetic on "
            "the GPU.  Moments after the ghost-node fold (applyBCids) on both sides: per "
            "moment max |GPU - ref| over species and nodes / max |ref|, the same without "
            "the 3 node planes next to each reflecting y wall (interior), and the "
            "relative L2 norm over all nodes.  Particles in load order: max and rms of "
            "|dx| / box length (x, z modulo the period) and |du| / max |u|; flips: "
            "particles whose velocity component has the other sign (reflected at a wall "
            "in one run only)")
        for prec, key in (("f32", "fp32_error"), ("f64", "fp64_error")):
            line[key] = {"reference": "oracle" if refleg["kind"] == "port" else
                         "reference (oracle/_ref, the reference's own sources)",
                         "cycles": refleg["cycles"], **refleg["errors"][prec],
                         "definition": definition}
        line["reference_leg_cpu_seconds"] = refleg["cpu_seconds_all_cycles"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
