"""The batching form at its configured size (GPU): 16 independent
Harris-sheet instances with grids from 32x16x16 to 96x48x48 cells, four
species each at 1..64 particles per cell (0.6M..13.3M particles per
instance, 71.7M in all), perturbation amplitudes in [0.005, 0.05], each
instance from its own seed (bench.batch_specs), FP32, advanced by one
batch: one eager phase, then two phases captured in a HIP graph (with a
batched tile re-sort of every species of every instance before the third
phase) and replayed once.

* Every instance, every particle and node: each particle comes back at its
  load index exactly once (q, which the push leaves alone, returns bit for
  bit into NaN-filled arrays); charge, the x current and the xx pressure sum
  to their particle sums (sum(rhon)/invVOL = sum(q), sum(Jx)/invVOL =
  sum(q u), sum(pxx)/invVOL = sum(q u^2)); linear precision along x.  Within
  2e-6 of the sum of magnitudes (FP32 weights and register sums).
* Instances under 2M particles: in strict arithmetic, the particle state is
  bit-identical to the FP32 oracle after the same three cycles, and the
  node moments within 1e-5 of each array's largest magnitude (the oracle
  accumulates in float in another order).
"""
from __future__ import annotations

import numpy as np
import pytest

import bench
import oracle as O
import sputnipic_amd as S

pytestmark = pytest.mark.gpu

SMALL = 2_000_000


class Instance:
    def __init__(self, spec, stream):
        self.spec = spec
        self.grd = S.Grid(*spec["grid"], *bench.BOX, "f32")
        self.gem = S.GemParams(pertX=spec["pertX"])
        self.field = S.gem_fields(self.grd, self.gem)
        self.species = S.harris_species()
        for sp, (px, py) in zip(self.species, spec["ppc"]):
            sp.npcelx, sp.npcely, sp.npcelz = px, py, py
        self.eng = S.Engine(self.grd, len(self.species))
        self.q0, self.ic, self.counts = [], [], []
        for s, sp in enumerate(self.species):
            p = S.gem_particles(self.grd, self.gem, sp, s, seed=spec["seed"])
            self.counts.append(p.npmax)
            self.q0.append(p.d["q"].download())
            if spec["particles"] < SMALL:
                self.ic.append(p.download())
            self.eng.load(s, p, stream)
            stream.synchronize()
            for d in p.d.values():
                d.free()
        assert sum(self.counts) == spec["particles"]
        self.ids = [S.InterpDensSpecies(self.grd) for _ in self.species]
        self.param = S.Parameters(dt=0.25, c=1.0, NiterMover=3)

    def member(self):
        return (self.eng, self.field, self.ids, self.param)

    def close(self):
        self.eng.close()


def _close(got, want, scale, tol, what):
    assert abs(got - want) <= tol * scale, f"{what}: {got} vs {want} (scale {scale})"


@pytest.mark.parametrize("strict", [True, False])
def test_sixteen_instance_batch_at_configured_size(strict):
    S.set_async(False)
    S.set_strict_fp(strict)
    stream = S.Stream()
    specs = bench.batch_specs()
    assert len(specs) == 16
    insts = []
    try:
        for sp in specs:
            insts.append(Instance(sp, stream))
        batch = S.Batch([x.member() for x in insts], sort_every=2, set_moments=True)
        try:
            batch.phase(stream)
            stream.synchronize()
            S.set_async(True)
            with S.Graph(stream) as graph:
                batch.phase(stream)
                batch.phase(stream)
            graph.launch()
            stream.synchronize()
            S.check_domain(stream)
            graph.close()
        finally:
            S.set_async(False)
            batch.close()
        n_small = 0
        for x in insts:
            _check_instance(x)
            if x.ic and strict:
                _check_against_oracle(x)
                n_small += 1
        if strict:
            assert n_small == sum(1 for sp in specs if sp["particles"] < SMALL) == 4
    finally:
        S.set_strict_fp(False)
        for x in insts:
            x.close()


def _check_instance(x):
    grd = x.grd
    iv = float(grd.invVOL)
    dx = bench.BOX[0] / x.spec["grid"][0]
    xn = (np.arange(grd.nxn, dtype=np.float64) - 1.0) * dx
    tag = "x".join(map(str, x.spec["grid"])) + f" seed {x.spec['seed']}"
    x.out = []
    for s, sp in enumerate(x.species):
        out = S.DeviceParticles(x.counts[s], sp.qom, "f32")
        for d in out.d.values():
            d.upload(np.full(x.counts[s], np.nan, np.float32))
        x.eng.store(s, out)
        got = {k: out.d[k].download() for k in O.PARTICLE_NAMES}
        for d in out.d.values():
            d.free()
        assert np.array_equal(got["q"], x.q0[s]), f"{tag} s{s}: q not returned"
        for k in ("x", "y", "z", "u", "v", "w"):
            assert np.all(np.isfinite(got[k])), f"{tag} s{s} {k}: slot not written"
        if x.ic:
            x.out.append(got)
        qd = got["q"].astype(np.float64)
        u = got["u"].astype(np.float64)
        m = x.ids[s].download()
        _close(float(m["rhon"].sum()) / iv, float(qd.sum()), float(np.abs(qd).sum()), 2e-6,
               f"{tag} s{s} charge")
        qu = qd * u
        _close(float(m["Jx"].sum()) / iv, float(qu.sum()), float(np.abs(qu).sum()), 2e-6,
               f"{tag} s{s} Jx")
        quu = qu * u
        _close(float(m["pxx"].sum()) / iv, float(quu.sum()), float(np.abs(quu).sum()), 2e-6,
               f"{tag} s{s} pxx")
        rx = m["rhon"].reshape(grd.nxn, grd.nyn, grd.nzn).sum(axis=(1, 2))
        qx = qd * got["x"].astype(np.float64)
        _close(float((xn * rx).sum()) / iv, float(qx.sum()), float(np.abs(qx).sum()), 2e-6,
               f"{tag} s{s} linear precision")


def _check_against_oracle(x):
    g = x.spec["grid"]
    cfg = O.GemConfig(nxc=g[0], nyc=g[1], nzc=g[2], Lx=bench.BOX[0], Ly=bench.BOX[1],
                      Lz=bench.BOX[2])
    st = O.State(cfg, "f32", build_ic=False)
    for n in O.FIELD_NAMES:
        st.fields[n][:] = x.field.d[n].download()
    st.parts = [{k: v.copy() for k, v in p.items()} for p in x.ic]
    st.cfg.qom = [sp.qom for sp in x.species]
    for _ in range(3):
        st.cycle()
    tag = "x".join(map(str, g)) + f" seed {x.spec['seed']}"
    for s in range(len(x.species)):
        for k in O.PARTICLE_NAMES:
            assert np.array_equal(x.out[s][k], st.parts[s][k]), f"{tag} s{s} {k}"
        m = x.ids[s].download()
        for n in O.MOMENT_NAMES:
            want = st.moments[s][n].astype(np.float64)
            err = float(np.abs(m[n] - want).max()) / float(np.abs(want).max())
            assert err <= 1e-5, f"{tag} s{s} {n}: {err:.2e}"
