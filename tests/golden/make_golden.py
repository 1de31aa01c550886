"""Write the golden fixtures in this directory from the reference's own code.

Runs only where the reference checkout exists (the build container): it
loads ``oracle/_ref/libref_{f64,f32}.so``, which ``oracle/Makefile`` compiles
from the reference's unmodified src/Mover.cpp, Grid.cpp, Particles.cpp,
IC.cpp, InterpDensSpecies.cpp, InterpDensNet.cpp, EMfield.cpp and BC.cpp,
and records inputs and outputs of mover_PC, interpP2G, mover_interp,
particle_sort and applyBCids on two small Harris-sheet cases:

* ``gem_ref``  -- the reference's GEM_3D parameters (inputfiles/GEM_3D.inp)
  on an 8x4x4-cell 20x10x10 box, 2x1x1 particles per cell, dt 0.25,
  initGEM fields and particles (srand(species), src/IC.cpp:121-122).
* ``gem_hot``  -- the same initGEM on an 8x4x4-cell 2x1x1 box with hot
  species (uth = 0.05), B0x = 0.2, dt = 1.0 and a nonzero smooth electric
  field written over the initGEM one, so that particles cross cells, wrap
  through the periodic x/z faces and reflect off the y faces, and the E
  terms of the push are exercised.

Each case runs, per species: mover_PC, zero + interpP2G (cycle 1),
applyBCids, then mover_PC + interpP2G again (cycle 2), then mover_interp
(cycle 3, fused), then particle_sort with sort_cps 1.  Arrays are stored
in the reference precision (float64 / float32) in ``<case>_<prec>.npz``.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

CASES = {
    "gem_ref": dict(nxc=8, nyc=4, nzc=4, Lx=20.0, Ly=10.0, Lz=10.0, dt=0.25,
                    npcelx=[2] * 4, npcely=[1] * 4, npcelz=[1] * 4, sort_cps=1),
    "gem_hot": dict(nxc=8, nyc=4, nzc=4, Lx=2.0, Ly=1.0, Lz=1.0, dt=1.0,
                    npcelx=[2] * 4, npcely=[1] * 4, npcelz=[1] * 4, sort_cps=1,
                    uth=[0.05] * 4, vth=[0.05] * 4, wth=[0.05] * 4, B0x=0.2,
                    w0=[0.01, -0.01, 0.0, 0.0]),
}


def case_config(name: str) -> O.GemConfig:
    return O.GemConfig(**CASES[name])


def hot_efield(XN: np.ndarray, YN: np.ndarray, ZN: np.ndarray, cfg: O.GemConfig):
    """Smooth electric field written over initGEM's E = 0 for gem_hot."""
    kx, ky, kz = 2 * np.pi / cfg.Lx, np.pi / cfg.Ly, 2 * np.pi / cfg.Lz
    Ex = 0.002 * np.sin(kx * XN) * np.cos(ky * YN)
    Ey = 0.001 * np.cos(kz * ZN) + 0.0005
    Ez = -0.0015 * np.sin(kx * XN + kz * ZN)
    return Ex, Ey, Ez


def run_case(name: str, prec: str) -> dict:
    cfg = case_config(name)
    ref = O.RefState(cfg, prec)
    dt = np.float64 if prec == "f64" else np.float32
    out = {}
    XN, YN, ZN = ref.field(6), ref.field(7), ref.field(8)
    if name == "gem_hot":
        for w, arr in enumerate(hot_efield(XN, YN, ZN, cfg)):
            ref.set_field(w, arr)
    for w, n in enumerate(O.FIELD_NAMES):
        out[f"field_{n}"] = ref.field(w).astype(dt)
    out["XN"], out["YN"], out["ZN"] = XN.astype(dt), YN.astype(dt), ZN.astype(dt)
    out["grid_ints"] = np.array(ref.grid_ints, np.int64)
    out["grid_dbl"] = np.array(ref.grid_dbl, np.float64)
    ns = cfg.ns
    for s in range(ns):
        for k, a in ref.particles(s).items():
            out[f"ic_s{s}_{k}"] = a.astype(dt)
        out[f"ic_s{s}_rhoc"] = ref.moment(s, 10).astype(dt)
    # cycle 1: mover_PC, then zero + interpP2G
    for s in range(ns):
        assert ref.mover_pc(s) == 0
    ref.zero_moments()
    for s in range(ns):
        ref.interp_p2g(s)
        for k, a in ref.particles(s).items():
            out[f"c1_s{s}_{k}"] = a.astype(dt)
        for w, n in enumerate(O.MOMENT_NAMES):
            out[f"c1_s{s}_{n}"] = ref.moment(s, w).astype(dt)
    # applyBCids on cycle-1 moments
    for s in range(ns):
        ref.apply_bc_ids(s)
        for w, n in enumerate(O.MOMENT_NAMES):
            out[f"bc1_s{s}_{n}"] = ref.moment(s, w).astype(dt)
    # cycle 2: mover_PC + zero + interpP2G
    for s in range(ns):
        ref.mover_pc(s)
    ref.zero_moments()
    for s in range(ns):
        ref.interp_p2g(s)
        for k, a in ref.particles(s).items():
            out[f"c2_s{s}_{k}"] = a.astype(dt)
        for w, n in enumerate(O.MOMENT_NAMES):
            out[f"c2_s{s}_{n}"] = ref.moment(s, w).astype(dt)
    # cycle 3: fused mover_interp
    ref.zero_moments()
    for s in range(ns):
        ref.mover_interp(s)
        for k, a in ref.particles(s).items():
            out[f"c3_s{s}_{k}"] = a.astype(dt)
        for w, n in enumerate(O.MOMENT_NAMES):
            out[f"c3_s{s}_{n}"] = ref.moment(s, w).astype(dt)
    # particle_sort
    for s in range(ns):
        ref.particle_sort(s)
        for k, a in ref.particles(s).items():
            out[f"sort_s{s}_{k}"] = a.astype(dt)
    ref.close()
    return out


def main():
    if not (O.ref_available("f64") and O.ref_available("f32")):
        raise SystemExit("oracle/_ref not built (needs the reference checkout): make -C oracle ref")
    for name in CASES:
        for prec in ("f64", "f32"):
            data = run_case(name, prec)
            path = os.path.join(HERE, f"{name}_{prec}.npz")
            np.savez_compressed(path, **data)
            print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(data)} arrays)")


if __name__ == "__main__":
    main()
