"""Shared test helpers: golden fixtures, oracle states built from them, and
comparison functions with the tolerances the tests state."""
from __future__ import annotations

import os

import numpy as np

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("gem_ref", "gem_hot")
PRECS = ("f64", "f32")

# Configurations of the golden cases (tests/golden/make_golden.py).
CASE_CFG = {
    "gem_ref": dict(nxc=8, nyc=4, nzc=4, Lx=20.0, Ly=10.0, Lz=10.0, dt=0.25,
                    npcelx=[2] * 4, npcely=[1] * 4, npcelz=[1] * 4, sort_cps=1),
    "gem_hot": dict(nxc=8, nyc=4, nzc=4, Lx=2.0, Ly=1.0, Lz=1.0, dt=1.0,
                    npcelx=[2] * 4, npcely=[1] * 4, npcelz=[1] * 4, sort_cps=1,
                    uth=[0.05] * 4, vth=[0.05] * 4, wth=[0.05] * 4, B0x=0.2,
                    w0=[0.01, -0.01, 0.0, 0.0]),
}


def load_case(name: str, prec: str) -> dict:
    with np.load(os.path.join(GOLDEN, f"{name}_{prec}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def case_config(name: str) -> O.GemConfig:
    return O.GemConfig(**CASE_CFG[name])


def oracle_from_fixture(name: str, prec: str, data: dict) -> O.State:
    """Oracle state holding the fixture's IC fields and particles."""
    cfg = case_config(name)
    st = O.State(cfg, prec, build_ic=False)
    for n in O.FIELD_NAMES:
        st.fields[n][:] = data[f"field_{n}"]
    st.parts = [{k: data[f"ic_s{s}_{k}"].copy() for k in O.PARTICLE_NAMES}
                for s in range(cfg.ns)]
    return st


def assert_bitwise(got, want, what: str) -> None:
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{what}: shape {got.shape} != {want.shape}"
    if not np.array_equal(got, want):
        bad = np.flatnonzero(got != want)
        i = int(bad[0])
        raise AssertionError(f"{what}: {bad.size}/{got.size} elements differ; first at {i}: "
                             f"{got[i]!r} != {want[i]!r}")


def max_rel_err(got, want) -> float:
    """max |got - want| / max |want| (0 if both are zero)."""
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    scale = np.max(np.abs(want)) if want.size else 0.0
    diff = np.max(np.abs(got - want)) if want.size else 0.0
    if scale == 0.0:
        return float(diff)
    return float(diff / scale)


def assert_close(got, want, rel: float, what: str) -> None:
    e = max_rel_err(got, want)
    assert e <= rel, f"{what}: max relative error {e:.3e} > {rel:.1e}"
