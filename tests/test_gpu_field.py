"""GPU parity of the field-solve side (spic_bc_scalar, spic_rhoc_from_rhon,
spic_hat_densities, spic_maxwell_image through the C ABI) against the
oracle, which is bit-identical to the reference's FP64 build
(tests/test_oracle_field.py).  Every kernel keeps the reference's operation
order without FMA contraction, so the device results are bit-identical to
the oracle's: rhon after its ghost update, rhoc, rho hat and J hat with
their ghosts, the Maxwell image, and every ghost-update kind, for every
combination of field periodicity, on a grid with inexact spacings."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import sputnipic_amd as S
from helpers import assert_bitwise

pytestmark = pytest.mark.gpu

PERIODICITIES = [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]


@pytest.fixture(autouse=True)
def _modes():
    S.set_async(False)
    S.set_strict_fp(False)
    yield


def _state(per, shape=(8, 6, 4)):
    cfg = O.GemConfig(nxc=shape[0], nyc=shape[1], nzc=shape[2], npcelx=[2] * 4,
                      npcely=[1] * 4, npcelz=[1] * 4, PERIODICX=bool(per[0]),
                      PERIODICY=bool(per[1]), PERIODICZ=bool(per[2]))
    st = O.State(cfg, "f64")
    st.cycle()
    return st


def _device(st):
    c = st.cfg
    grd = S.Grid(c.nxc, c.nyc, c.nzc, c.Lx, c.Ly, c.Lz, "f64")
    field = S.EMField(grd, {n: st.fields[n] for n in O.FIELD_NAMES})
    ids = [S.InterpDensSpecies(grd) for _ in range(c.ns)]
    for s in range(c.ns):
        ids[s].upload(st.moments[s])
    sp = S.SolveParams(dt=c.dt, c=c.c, th=c.th, fourpi=c.fourpi, PERIODICX=c.PERIODICX,
                       PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ, qom=c.qom[:c.ns])
    param = S.Parameters(dt=c.dt, c=c.c, NiterMover=c.NiterMover, PERIODICX=c.PERIODICX,
                         PERIODICY=c.PERIODICY, PERIODICZ=c.PERIODICZ)
    return grd, field, ids, sp, param


@pytest.mark.parametrize("per", PERIODICITIES)
def test_hat_densities_bitwise(per):
    st = _state(per, shape=(8, 6, 4) if per != (1, 0, 1) else (12, 7, 5))
    grd, field, ids, sp, param = _device(st)
    net = S.InterpDensNet(grd)
    S.sumOverSpecies(net, ids, grd, st.cfg.ns)
    S.interpN2Crho_net(net, grd, sp)
    aux = S.InterpDensAux(grd)
    S.calculateHatDensities(aux, net, ids, field, grd, sp)
    want_net = st.sum_species()
    rhon, rhoc = st.rhoc_from_rhon(want_net["rhon"])
    assert_bitwise(net.d["rhon"].download(), rhon, "rhon after applyBCscalarDensN")
    assert_bitwise(net.rhoc.download(), rhoc, "rhoc")
    want = st.hat_densities(rhoc)
    got = aux.download()
    for n in ("rhoh", "Jxh", "Jyh", "Jzh"):
        assert_bitwise(got[n], want[n], f"{per} {n}")


@pytest.mark.parametrize("per", PERIODICITIES)
def test_maxwell_image_bitwise(per):
    st = _state(per)
    grd, field, ids, sp, param = _device(st)
    g = st.grid
    n = 3 * (g.nxn - 2) * (g.nyn - 2) * (g.nzn - 2)
    rng = np.random.default_rng(17 + per[0] + 2 * per[1] + 4 * per[2])
    vec = rng.normal(0, 1e-2, n)
    dv = S.DeviceArray.from_numpy(vec)
    im = S.DeviceArray(n, np.float64)
    for _ in range(2):  # the workspace keeps the first call's contents
        S.MaxwellImage(im, dv, field, ids, grd, sp)
        assert_bitwise(im.download(), st.maxwell_image(vec), f"{per} MaxwellImage")


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("per", [(0, 0, 0), (1, 0, 1), (0, 1, 0), (1, 1, 1)])
def test_bc_scalar_bitwise(per, kind):
    st = _state(per)
    grd, field, ids, sp, param = _device(st)
    n = st.nn if kind <= 1 else st.ncell
    a = np.random.default_rng(kind + 10 * per[0] + 20 * per[1] + 40 * per[2]).normal(0, 1, n)
    d = S.DeviceArray.from_numpy(a)
    S.applyBCscalar(d, grd, sp, kind)
    assert_bitwise(d.download(), st.bc_scalar(a, kind), f"{per} kind {kind}")


def test_field_solve_rejects_fp32_grid():
    grd = S.Grid(8, 6, 4, 20.0, 10.0, 10.0, "f32")
    d = S.DeviceArray(grd.nn, np.float64)
    with pytest.raises(S.SpicError) as ei:
        S.applyBCscalar(d, grd, S.SolveParams(), 0)
    assert ei.value.code == S.SPIC_EINVAL
