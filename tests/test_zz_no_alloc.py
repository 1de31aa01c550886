"""spic_particle_sort_noalloc (GPU): the allocation-free particle_sort,
through the C ABI, against spic_particle_sort on the same arrays.

Two calls in a row (the second with fewer particles, so that the shared
fixed scratch holds the first call's larger contents) must each give the
same seven arrays, bit for bit, as spic_particle_sort, whose output order
is the reference's (src/Particles.cpp:155-233, checked against the oracle
in test_gpu_parity.py); and a particle count beyond the fixed scratch is
refused with SPIC_EINVAL before anything runs.  Particles are the Harris
sheet's with positions shuffled over the box (plus a few outside it, the
sort's "lost" bin), in both precisions.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import sputnipic_amd as S
from sputnipic_amd._capi import SPIC_EINVAL, lib

pytestmark = pytest.mark.gpu

GRID = (32, 16, 16)
BOX = (20.0, 10.0, 10.0)
NAMES = ("x", "y", "z", "u", "v", "w", "q")


def _arrays(n, seed, prec):
    rng = np.random.default_rng(seed)
    dt = np.float64 if prec == "f64" else np.float32
    a = {"x": rng.uniform(0, BOX[0], n), "y": rng.uniform(0, BOX[1], n),
         "z": rng.uniform(0, BOX[2], n)}
    for k in "uvw":
        a[k] = rng.normal(0, 0.05, n)
    a["q"] = rng.uniform(-1e-3, 1e-3, n)
    # a few particles outside the box: the trailing "lost" bin
    k = max(1, n // 1000)
    a["x"][:k] = -0.5
    a["y"][k:2 * k] = BOX[1] + 0.5
    return {key: v.astype(dt) for key, v in a.items()}


def _sort(fn, grd, arrays, sort_cps):
    part = S.Particles({k: v.copy() for k, v in arrays.items()}, -64.0, grd.prec)
    g, s = grd.c_struct(), part.c_struct()
    rc = getattr(lib(), fn)(C.byref(g), sort_cps, C.byref(s), None)
    return rc, part


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_sort_matches_particle_sort_twice(prec):
    S.set_async(False)
    grd = S.Grid(*GRID, *BOX, prec)
    for call, (n, seed, cps) in enumerate(((1_200_000, 1, 2), (300_001, 2, 1))):
        print(f"[noalloc] {prec} call {call}: n={n} sort_cps={cps}", file=sys.stderr, flush=True)
        arrays = _arrays(n, seed, prec)
        rc, want = _sort("spic_particle_sort", grd, arrays, cps)
        assert rc == 0, lib().spic_last_error()
        rc, got = _sort("spic_particle_sort_noalloc", grd, arrays, cps)
        assert rc == 0, lib().spic_last_error()
        w, g = want.download(), got.download()
        for k in NAMES:
            assert np.array_equal(g[k], w[k]), (prec, call, k)
        # it is a sort: the input was not already in order
        assert not np.array_equal(w["x"], arrays["x"])


def test_noalloc_sort_refuses_beyond_its_scratch():
    S.set_async(False)
    grd = S.Grid(*GRID, *BOX, "f64")
    mib = 64
    n = (mib << 20) // 8 + 1  # more FP64 particles than the scratch holds bytes / 8
    part = S.DeviceParticles(n, -64.0, "f64")
    g, s = grd.c_struct(), part.c_struct()
    rc = lib().spic_particle_sort_noalloc(C.byref(g), 2, C.byref(s), None)
    assert rc == SPIC_EINVAL
    assert b"scratch" in lib().spic_last_error()
