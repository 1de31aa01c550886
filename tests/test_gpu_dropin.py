"""The reference's C++ entry points through the drop-in library on the GPU.

A C++ driver (tests/dropin/dropin_driver.cpp, built into
sputnipic_amd/lib/spic_dropin_driver_{f32,f64}) fills host structs laid out
as the reference's, calls mover_PC / interpP2G / mover_interp /
particle_sort in the order of the reference's CPU driver, and writes the
host arrays back out.  Compared with the oracle on the same inputs, in
strict arithmetic: particle arrays bit for bit; node moments (accumulated
in FP64 on the device, added once into the host arrays) within 1e-13 of the
FP64 reference and 1e-5 of the FP32 reference, which accumulates in float.
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from helpers import assert_bitwise, assert_close

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "sputnipic_amd", "lib")


def run_driver(st: O.State, prec: str, ops, strict=True, resident=False):
    c = st.cfg
    dt = np.float64 if prec == "f64" else np.float32
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("12i", c.nxc, c.nyc, c.nzc, c.ns, c.NiterMover, len(ops),
                                int(strict), int(resident), c.sort_cps, int(c.PERIODICX),
                                int(c.PERIODICY), int(c.PERIODICZ)))
            f.write(struct.pack("5d", c.Lx, c.Ly, c.Lz, c.dt, c.c))
            f.write(struct.pack(f"{len(ops)}i", *ops))
            for n in O.FIELD_NAMES:
                f.write(np.ascontiguousarray(st.fields[n], dt).tobytes())
            for s in range(c.ns):
                p = st.parts[s]
                f.write(struct.pack("q", len(p["x"])))
                f.write(struct.pack("d", c.qom[s]))
                for k in O.PARTICLE_NAMES:
                    f.write(np.ascontiguousarray(p[k], dt).tobytes())
        exe = os.path.join(LIB, f"spic_dropin_driver_{prec}")
        r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        raw = np.fromfile(fout, dtype=dt)
    parts, moms, off = [], [], 0
    for s in range(c.ns):
        n = len(st.parts[s]["x"])
        parts.append({k: raw[off + i * n: off + (i + 1) * n] for i, k in
                      enumerate(O.PARTICLE_NAMES)})
        off += 7 * n
    for s in range(c.ns):
        moms.append({m: raw[off + i * st.nn: off + (i + 1) * st.nn] for i, m in
                     enumerate(O.MOMENT_NAMES)})
        off += 10 * st.nn
    assert off == raw.size
    return parts, moms


def small_state(prec):
    cfg = O.GemConfig(nxc=16, nyc=8, nzc=8, npcelx=[3] * 4, npcely=[2] * 4,
                      npcelz=[2] * 4, sort_cps=2)
    return O.State(cfg, prec)


@pytest.mark.parametrize("resident", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dropin_reference_cycle(prec, resident):
    """Two reference cycles (mover_PC for every species, zero, interpP2G
    for every species) through the C++ entry points."""
    st = small_state(prec)
    parts, moms = run_driver(st, prec, [1, 2, 1, 2], resident=resident)
    st.cycle()
    st.cycle()
    tol = 1e-13 if prec == "f64" else 1e-5
    for s in range(st.cfg.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(parts[s][k], st.parts[s][k], f"s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(moms[s][n], st.moments[s][n], tol, f"s{s} {n}")


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dropin_fused_and_sort(prec):
    st = small_state(prec)
    parts, moms = run_driver(st, prec, [3, 4])
    for s in range(st.cfg.ns):
        st.zero_moments(s)
        st.mover_interp(s)
        st.particle_sort(s)
    tol = 1e-13 if prec == "f64" else 1e-5
    for s in range(st.cfg.ns):
        for k in O.PARTICLE_NAMES:
            assert_bitwise(parts[s][k], st.parts[s][k], f"s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_close(moms[s][n], st.moments[s][n], tol, f"s{s} {n}")
