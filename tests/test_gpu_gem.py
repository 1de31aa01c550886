"""The device Harris-sheet generator (spic_gem_fields, spic_gem_particles)
against the oracle's restatement of the reference's initGEM
(src/IC.cpp:38-177, oracle_gem_fields / oracle_gem_particles, which are
bit-identical to the reference's own build), on the benchmark's grid and a
small one, FP64 and FP32.

* Fields: the same expressions in the same order; the device's tanh, exp,
  cos and sin may differ from glibc's in the last bits, so in FP64 every
  node is within 8 ulp of the largest |B| (3.5e-17 for B0x = 0.0195), and
  in FP32, whose values are those doubles rounded, within one float ulp of
  the element.  E and Bz are exact.
* Particles: positions on the reference's sub-lattice in its cell-major
  order (with the kk < npcely bound of src/IC.cpp:128) and charges from the
  interpN2Crho row table are bit for bit the oracle's.  Velocities come
  from another random source on purpose (a counter-based SplitMix64 instead
  of glibc rand()): they are checked as a distribution -- mean drift and
  thermal spread of each component within five standard errors.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S

pytestmark = pytest.mark.gpu

CASES = [((32, 16, 16), (3, 3, 3), 0.4), ((128, 64, 64), (1, 1, 1), 0.4),
         ((48, 24, 24), (2, 2, 2), 0.0226)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape,ppc,pert", CASES)
def test_device_generator_matches_oracle_initgem(prec, shape, ppc, pert):
    S.set_async(False)
    cfg = O.GemConfig(nxc=shape[0], nyc=shape[1], nzc=shape[2], npcelx=[ppc[0]] * 4,
                      npcely=[ppc[1]] * 4, npcelz=[ppc[2]] * 4, pertX=pert)
    st = O.State(cfg, prec, max_per_species=0)
    grd = S.Grid(*shape, cfg.Lx, cfg.Ly, cfg.Lz, prec)
    gem = S.GemParams(pertX=pert)
    field = S.gem_fields(grd, gem)
    bmax = max(float(np.abs(st.fields[n]).max()) for n in ("Bxn", "Byn"))
    for n in O.FIELD_NAMES:
        got = field.d[n].download()
        want = st.fields[n]
        if n in ("Ex", "Ey", "Ez", "Bzn"):
            assert np.array_equal(got, want), n
        elif prec == "f64":
            err = float(np.abs(got - want).max())
            assert err <= 8 * np.spacing(bmax), (n, err)
        else:
            assert np.all(np.abs(got - want) <= np.spacing(np.abs(want))), n
    for s, sp in enumerate(S.harris_species(ppc)):
        p = S.gem_particles(grd, gem, sp, s, seed=0)
        got = p.download()
        want = st.parts[s]
        assert len(got["x"]) == len(want["x"])
        for k in ("x", "y", "z", "q"):
            assert np.array_equal(got[k], want[k]), f"s{s} {k}"
        n = len(got["u"])
        for k, mean, th in (("u", sp.u0, sp.uth), ("v", sp.v0, sp.vth), ("w", sp.w0, sp.wth)):
            a = got[k].astype(np.float64)
            assert abs(a.mean() - mean) <= 5 * th / np.sqrt(n), (s, k, a.mean())
            # Box-Muller from 1 - .999999 h: the spread is th to 1e-6
            assert abs(a.std() / th - 1.0) <= 5 * np.sqrt(0.5 / n) + 1e-5, (s, k, a.std())
        for d in p.d.values():
            d.free()
