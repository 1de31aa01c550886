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


@pytest.mark.parametrize("resident", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dropin_interp_from_a_thread_team(prec, resident):
    """The reference cycle with interpP2G called from one thread per species
    at once, as the reference's OpenMP team calls it: the same results as
    the sequential calls."""
    st = small_state(prec)
    parts, moms = run_driver(st, prec, [1, 5, 1, 5], resident=resident)
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


def run_field_driver(st: O.State, prec: str, rhoc, vec):
    c = st.cfg
    dt = np.float64 if prec == "f64" else np.float32
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("7i", c.nxc, c.nyc, c.nzc, c.ns, int(c.PERIODICX),
                                int(c.PERIODICY), int(c.PERIODICZ)))
            f.write(struct.pack(f"{7 + c.ns}d", c.Lx, c.Ly, c.Lz, c.dt, c.c, c.th, c.fourpi,
                                *c.qom[:c.ns]))
            for n in ("Bxn", "Byn", "Bzn"):
                f.write(np.ascontiguousarray(st.fields[n], dt).tobytes())
            for s in range(c.ns):
                for m in O.MOMENT_NAMES:
                    f.write(np.ascontiguousarray(st.moments[s][m], dt).tobytes())
            f.write(np.ascontiguousarray(rhoc, dt).tobytes())
            f.write(np.ascontiguousarray(vec, dt).tobytes())
        exe = os.path.join(LIB, f"spic_dropin_field_driver_{prec}")
        r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        raw = np.fromfile(fout, dtype=dt)
    nn, nc = st.nn, st.ncell
    out, off = {}, 0
    for name, n in (("rhoh", nc), ("Jxh", nn), ("Jyh", nn), ("Jzh", nn), ("im", len(vec))):
        out[name] = raw[off:off + n]
        off += n
    assert off == raw.size
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dropin_field_solve(prec):
    """calculateHatDensities and MaxwellImage through the C++ entry points,
    on the moments of one cycle.  FP64: bit-identical to the oracle (and so
    to the reference).  Float build: the library widens the inputs, works in
    FP64 and rounds once, while the float reference rounds every operation:
    within 1e-5 of each array's largest magnitude."""
    st = small_state(prec)
    st.cycle()
    net = st.sum_species()
    _, rhoc = st.rhoc_from_rhon(net["rhon"])
    g = st.grid
    nvec = 3 * (g.nxn - 2) * (g.nyn - 2) * (g.nzn - 2)
    vec = np.random.default_rng(4).normal(0, 1e-2, nvec).astype(st.dt)
    got = run_field_driver(st, prec, rhoc, vec)
    want = st.hat_densities(rhoc)
    want["im"] = st.maxwell_image(vec)
    for n in ("rhoh", "Jxh", "Jyh", "Jzh", "im"):
        if prec == "f64":
            assert_bitwise(got[n], want[n], n)
        else:
            assert_close(got[n], want[n], 1e-5, n)
