#!/usr/bin/env python3
"""Throughput of sputniPIC's particle hot path (mover_PC + interpP2G) on one
MI355X, FP64 form and FP32 form, with its HBM-roofline fraction.

Workload: the 3-D Harris sheet (GEM) of inputfiles/GEM_3D.inp at the
published run's shape -- 128x64x64 cells on a 20x10x10 box, dt 0.25,
NiterMover 3, four species (qom -64, 1, -64, 1) at 5x5x5 particles per cell
= 65,536,000 per species, 262,144,000 in total, all resident in HBM.  The
fields are frozen at the initial condition (no field solve between cycles,
which is not on this path).  A step is one cycle: zero the node moments,
then push and deposit every species.

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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(n_particles: int, nn: int, w: int, n_species: int) -> int:
    """Minimum HBM traffic of fused mover + deposit: each particle's
    x,y,z,u,v,w,q read once and x..w written once (13 words), the six
    fields read once, the ten FP64 node moments of each species written
    once."""
    return n_particles * 13 * w + 6 * nn * w + 10 * n_species * nn * 8


def run_form(prec: str, steps: int, warmup: int) -> dict:
    import ctypes as C

    import sputnipic_amd as S
    from sputnipic_amd._capi import check, lib

    L = lib()
    w = 8 if prec == "f64" else 4
    grd = S.Grid(*GRID, *BOX, prec)
    gem = S.GemParams()
    field = S.gem_fields(grd, gem)
    species = S.harris_species(PPC)
    parts = [S.gem_particles(grd, gem, sp, s, seed=0) for s, sp in enumerate(species)]
    moms = [S.InterpDensSpecies(grd) for _ in range(NS)]
    n_total = sum(p.npmax for p in parts)
    log(f"{prec}: {n_total} particles resident, grid {grd.shape}")
    param = S.Parameters(dt=0.25, c=1.0, NiterMover=3)
    stream = S.Stream()
    g_c = grd.c_struct()
    f_c = field.c_struct()
    p_c = param.push()
    s_c = [p.c_struct() for p in parts]
    m_c = [m.c_struct() for m in moms]
    ev = [[S.Event(), S.Event()] for _ in range(NS)]
    S.set_async(True)
    kernel_ms = []

    def cycle(timed: bool):
        for s in range(NS):
            check(L.spic_zero_moments(C.byref(g_c), C.byref(m_c[s]), stream.ptr), "zero")
        for s in range(NS):
            if timed:
                ev[s][0].record(stream)
            check(L.spic_mover_interp(C.byref(g_c), C.byref(f_c), C.byref(s_c[s]),
                                      C.byref(m_c[s]), C.byref(p_c), stream.ptr),
                  "mover_interp")
            if timed:
                ev[s][1].record(stream)
        if timed:
            for s in range(NS):
                kernel_ms.append((s, ev[s][0].elapsed_ms(ev[s][1])))

    try:
        for _ in range(warmup):
            cycle(False)
        stream.synchronize()
        check(L.spic_check_domain(stream.ptr), "domain check after warmup")
        t0 = time.perf_counter()
        for _ in range(steps):
            cycle(True)
        stream.synchronize()
        t1 = time.perf_counter()
        check(L.spic_check_domain(stream.ptr), "domain check")
    finally:
        S.set_async(False)
    wall = t1 - t0
    # charge conservation on the last cycle: sum(rhon)/invVOL == sum(q)
    q_err = 0.0
    for s in range(NS):
        rho = moms[s].d["rhon"].download()
        q = parts[s].d["q"].download().astype("float64")
        tot = float(rho.sum()) / grd.invVOL
        q_err = max(q_err, abs(tot - float(q.sum())) / float(abs(q).sum()))
    avg_kernel_ms = sum(t for _, t in kernel_ms) / len(kernel_ms)
    per_launch_bytes = sum(algorithmic_bytes(p.npmax, grd.nn, w, 1) for p in parts) / NS
    achieved = per_launch_bytes / (avg_kernel_ms * 1e-3) / 1e9
    value = n_total * steps / wall
    log(f"{prec}: {value / 1e9:.3f} G particles/s, {1e3 * wall / steps:.3f} ms/cycle, "
        f"kernel {avg_kernel_ms:.3f} ms avg, {achieved:.1f} GB/s, charge err {q_err:.1e}")
    out = {
        "value": value,
        "ms_per_step": 1e3 * wall / steps,
        "particles": n_total,
        "kernel": "mover_interp (plain fused, one launch per species)",
        "avg_kernel_ms": avg_kernel_ms,
        "bytes_per_launch": per_launch_bytes,
        "achieved_GBs": achieved,
        "roofline_frac": achieved / HBM_PEAK_GBS,
        "charge_rel_err": q_err,
    }
    for p in parts:
        for d in p.d.values():
            d.free()
    return out


def cpu_baseline(cycles: int) -> dict:
    """The reference's own CPU mover_PC + interpP2G (its source compiled at
    the optimisation level of its Makefile, oracle/_ref/libref_f64_O3.so),
    or the oracle restatement if that build is absent, timed on this host's
    cores with the reference driver's loop structure, on a bounded sample
    of the same workload: the 128x64x64 grid with 2x2x2 particles per cell
    (16,777,216 particles)."""
    import oracle as O

    cfg = O.GemConfig(nxc=GRID[0], nyc=GRID[1], nzc=GRID[2], Lx=BOX[0], Ly=BOX[1],
                      Lz=BOX[2], npcelx=[2] * NS, npcely=[2] * NS, npcelz=[2] * NS)
    if O.ref_available("f64_O3"):
        st = O.RefState(cfg, "f64_O3")
        kind = "reference"
        n = sum(st.nop(s) for s in range(NS))
        cores = st.max_threads()
    else:
        st = O.State(cfg, "f64")
        kind = "port"
        n = sum(len(p["x"]) for p in st.parts)
        cores = O.State.max_threads()
    total, tm, td = st.time_cycles(cycles)
    if kind == "reference":
        st.close()
    value = n * cycles / total
    log(f"cpu baseline ({kind}, {cores} threads): {value / 1e6:.1f} M particles/s "
        f"(mover {tm:.2f} s, deposit {td:.2f} s over {cycles} cycles)")
    return {
        "value": value,
        "unit": "particles/s",
        "cores": cores,
        "kind": kind,
        "sample": f"128x64x64 Harris sheet, 2x2x2 particles/cell x 4 species = {n} "
                  f"particles, FP64, {cycles} cycles of mover_PC + zero + interpP2G "
                  f"(OpenMP over particles in the mover, over species in the deposit)",
        "seconds": total,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cycles", type=int, default=10)
    ap.add_argument("--forms", default="f64,f32")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus}: this build measures one GPU only; the path "
              "would shard by particles or by x-slabs (DESIGN.md), which is not built "
              "here.", file=sys.stderr)
        return 2
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: multi-rank launch is not supported (one GPU only)", file=sys.stderr)
        return 2
    import sputnipic_amd as S
    if S.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    forms = {}
    for prec in args.forms.split(","):
        forms[prec] = run_form(prec, args.steps, args.warmup)
    primary = "f64" if "f64" in forms else list(forms)[0]
    f = forms[primary]
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_cycles)
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
        },
        "roofline": {
            "bound": "hbm",
            "achieved": f["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": f["roofline_frac"],
            "traffic": None,
            "kernel": f["kernel"],
            "avg_kernel_ms": f["avg_kernel_ms"],
            "algorithmic_bytes_per_launch": f["bytes_per_launch"],
        },
        "forms": {k: {kk: vv for kk, vv in v.items()} for k, v in forms.items()},
        "cpu_baseline": cpu,
    }
    if "f64" in forms and "f32" in forms:
        line["fp32_over_fp64"] = forms["f32"]["value"] / forms["f64"]["value"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
