"""Pins the CPU oracle (oracle/spic_oracle.c) before anything is checked
against it.

* Every function against the golden fixtures written by the reference's own
  translation units (tests/golden/make_golden.py), bit for bit, in FP64
  and FP32.
* Against the live reference build (oracle/_ref) on the full small
  Harris-sheet case (32x16x16 cells, 884,736 particles), where that build
  exists.
* Closed forms and invariants that a plausible slip (dropped term, wrong
  sign, swapped weight/node pairing, transposed index) would break.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import (CASES, PRECS, assert_bitwise, case_config, load_case,
                     oracle_from_fixture)


# --------------------------------------------------------------------------
# Golden fixtures: inputs and outputs of the reference's own code.
# --------------------------------------------------------------------------
@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", CASES)
def test_oracle_ic_matches_reference_fixture(name, prec):
    data = load_case(name, prec)
    cfg = case_config(name)
    st = O.State(cfg, prec)
    assert_bitwise(st.XN, data["XN"], "XN")
    assert_bitwise(st.YN, data["YN"], "YN")
    assert_bitwise(st.ZN, data["ZN"], "ZN")
    g = st.grid
    assert [g.nxc, g.nyc, g.nzc, g.nxn, g.nyn, g.nzn] == list(data["grid_ints"])
    dbl = [g.dx, g.dy, g.dz, g.invdx, g.invdy, g.invdz, g.invVOL, g.xStart, g.yStart,
           g.zStart, g.Lx, g.Ly, g.Lz]
    assert_bitwise(np.array(dbl), data["grid_dbl"], "grid scalars")
    fields = O.FIELD_NAMES if name == "gem_ref" else ("Bxn", "Byn", "Bzn")
    for n in fields:
        assert_bitwise(st.fields[n], data[f"field_{n}"], f"field {n}")
    for s in range(cfg.ns):
        assert_bitwise(st.rhoc[s], data[f"ic_s{s}_rhoc"], f"rhoc s{s}")
        for k in O.PARTICLE_NAMES:
            assert_bitwise(st.parts[s][k], data[f"ic_s{s}_{k}"], f"IC s{s} {k}")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", CASES)
def test_oracle_path_matches_reference_fixture(name, prec):
    data = load_case(name, prec)
    cfg = case_config(name)
    st = oracle_from_fixture(name, prec, data)
    ns = cfg.ns
    # cycle 1: mover_PC + interpP2G
    for s in range(ns):
        assert st.mover_pc(s) == 0
    for s in range(ns):
        st.zero_moments(s)
        st.interp_p2g(s)
        for k in O.PARTICLE_NAMES:
            assert_bitwise(st.parts[s][k], data[f"c1_s{s}_{k}"], f"c1 s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_bitwise(st.moments[s][n], data[f"c1_s{s}_{n}"], f"c1 s{s} {n}")
    # applyBCids
    for s in range(ns):
        st.apply_bc_ids(s)
        for n in O.MOMENT_NAMES:
            assert_bitwise(st.moments[s][n], data[f"bc1_s{s}_{n}"], f"bc1 s{s} {n}")
    # cycle 2
    for s in range(ns):
        st.mover_pc(s)
    for s in range(ns):
        st.zero_moments(s)
        st.interp_p2g(s)
        for k in O.PARTICLE_NAMES:
            assert_bitwise(st.parts[s][k], data[f"c2_s{s}_{k}"], f"c2 s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_bitwise(st.moments[s][n], data[f"c2_s{s}_{n}"], f"c2 s{s} {n}")
    # cycle 3: fused mover_interp
    for s in range(ns):
        st.zero_moments(s)
        assert st.mover_interp(s) == 0
        for k in O.PARTICLE_NAMES:
            assert_bitwise(st.parts[s][k], data[f"c3_s{s}_{k}"], f"c3 s{s} {k}")
        for n in O.MOMENT_NAMES:
            assert_bitwise(st.moments[s][n], data[f"c3_s{s}_{n}"], f"c3 s{s} {n}")
    # particle_sort
    for s in range(ns):
        assert st.particle_sort(s) == 0
        for k in O.PARTICLE_NAMES:
            assert_bitwise(st.parts[s][k], data[f"sort_s{s}_{k}"], f"sort s{s} {k}")


def test_gem_hot_fixture_exercises_boundaries():
    """The hot case must actually cross cells and both kinds of face."""
    data = load_case("gem_hot", "f64")
    cfg = case_config("gem_hot")
    dx = cfg.Lx / cfg.nxc
    crossed = wrapped = reflected = 0
    for s in range(cfg.ns):
        x0, x1 = data[f"ic_s{s}_x"], data[f"c1_s{s}_x"]
        v0, v1 = data[f"ic_s{s}_v"], data[f"c2_s{s}_v"]
        crossed += int(np.sum(np.floor(x0 / dx) != np.floor(x1 / dx)))
        wrapped += int(np.sum(np.abs(x1 - x0) > cfg.Lx / 2))
        y1, y2 = data[f"c1_s{s}_y"], data[f"c2_s{s}_y"]
        reflected += int(np.sum((np.sign(v1) != np.sign(v0)) & ((y1 < 0.1) | (y1 > cfg.Ly - 0.1)))
                         + np.sum((y2 < 0.05) | (y2 > cfg.Ly - 0.05)))
    assert crossed > 50 and wrapped > 5 and reflected > 5, (crossed, wrapped, reflected)


# --------------------------------------------------------------------------
# The live reference build (build container only).
# --------------------------------------------------------------------------
@pytest.mark.parametrize("prec", PRECS)
def test_oracle_matches_live_reference_full_small_case(prec):
    if not O.ref_available(prec):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    cfg = O.GemConfig()  # 32x16x16, 3x3x3 ppc, 4 species: 884,736 particles
    ref = O.RefState(cfg, prec)
    st = O.State(cfg, prec)
    for s in range(cfg.ns):
        rp = ref.particles(s)
        for k in O.PARTICLE_NAMES:
            assert_bitwise(st.parts[s][k], rp[k], f"IC s{s} {k}")
    for s in range(cfg.ns):
        ref.mover_pc(s)
        st.mover_pc(s)
    ref.zero_moments()
    for s in range(cfg.ns):
        st.zero_moments(s)
        st.interp_p2g(s)
        ref.interp_p2g(s)
        rp = ref.particles(s)
        for k in O.PARTICLE_NAMES:
            assert_bitwise(st.parts[s][k], rp[k], f"mover s{s} {k}")
        for w, n in enumerate(O.MOMENT_NAMES):
            assert_bitwise(st.moments[s][n], ref.moment(s, w), f"deposit s{s} {n}")
    ref.close()


@pytest.mark.parametrize("prec", ["f64", "f64_O3"])
def test_reference_state_from_external_inputs_matches_oracle(prec):
    """The benchmark's CPU reference leg builds the reference's state
    without initGEM and copies in inputs made elsewhere (the device
    generator's, here perturbed fields and particles that wrap and reflect):
    two cycles of the reference's own mover_PC + interpP2G, at its timing
    build's -O3 too, equal the oracle's bit for bit."""
    if not O.ref_available(prec):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    cfg = O.GemConfig(nxc=8, nyc=6, nzc=4, Lx=2.0, Ly=1.5, Lz=1.0, npcelx=[2] * 4,
                      npcely=[2] * 4, npcelz=[2] * 4, dt=0.4, uth=[0.1] * 4, vth=[0.1] * 4,
                      wth=[0.1] * 4)
    st = O.State(cfg, "f64")
    rng = np.random.default_rng(8)
    for n in O.FIELD_NAMES:
        st.fields[n] += rng.normal(0, 1e-4, st.nn)
    # at most 5.3 thermal speeds: under one cell (0.25) per cycle, so that
    # the reference, which has no bounds checks, stays inside its arrays
    for p in st.parts:
        assert max(np.abs(p[k]).max() for k in "uvw") * cfg.dt < 0.25
    ref = O.RefState(cfg, prec, ic=False)
    for s in range(cfg.ns):
        assert ref.nop(s) == len(st.parts[s]["x"])
        ref.set_particles(s, st.parts[s])
    for w, n in enumerate(O.FIELD_NAMES):
        ref.set_field(w, st.fields[n])
    for _ in range(2):
        ref.time_cycles(1)
        st.cycle()
    for s in range(cfg.ns):
        rp = ref.particles(s)
        for k in O.PARTICLE_NAMES:
            assert_bitwise(st.parts[s][k], rp[k], f"s{s} {k}")
        for w, n in enumerate(O.MOMENT_NAMES):
            assert_bitwise(st.moments[s][n], ref.moment(s, w), f"s{s} {n}")
    ref.close()


# --------------------------------------------------------------------------
# Closed forms for the mover.
# --------------------------------------------------------------------------
def _uniform_state(prec="f64", npart=2000, seed=1, qom=-64.0, niter=3, dt=0.25):
    cfg = O.GemConfig(nxc=8, nyc=8, nzc=8, Lx=4.0, Ly=4.0, Lz=4.0, dt=dt, ns=1,
                      npcelx=[1], npcely=[1], npcelz=[1], qom=[qom], NiterMover=niter,
                      PERIODICY=True)
    st = O.State(cfg, prec, build_ic=False)
    rng = np.random.default_rng(seed)
    dtp = st.dt
    # keep particles away from the faces so the BC never triggers
    st.parts = [{
        "x": rng.uniform(0.5, 3.5, npart).astype(dtp),
        "y": rng.uniform(0.5, 3.5, npart).astype(dtp),
        "z": rng.uniform(0.5, 3.5, npart).astype(dtp),
        "u": rng.normal(0, 0.05, npart).astype(dtp),
        "v": rng.normal(0, 0.05, npart).astype(dtp),
        "w": rng.normal(0, 0.05, npart).astype(dtp),
        "q": rng.uniform(0.5, 1.5, npart).astype(dtp),
    }]
    return st


def test_mover_free_streaming_is_exact():
    st = _uniform_state()
    p0 = {k: v.copy() for k, v in st.parts[0].items()}
    st.mover_pc(0)
    p = st.parts[0]
    for k, vk in (("x", "u"), ("y", "v"), ("z", "w")):
        assert_bitwise(p[vk], p0[vk], f"{vk} unchanged with E = B = 0")
        assert_bitwise(p[k], p0[k] + p0[vk] * 0.25, f"{k} = {k}0 + {vk}*dt")


def test_mover_uniform_e_kick():
    st = _uniform_state()
    E = (0.003, -0.002, 0.001)
    for n, e in zip(("Ex", "Ey", "Ez"), E):
        st.fields[n][:] = e
    p0 = {k: v.copy() for k, v in st.parts[0].items()}
    st.mover_pc(0)
    qomdt2 = -64.0 * 0.125 / 1.0
    for (k, vk), e in zip((("x", "u"), ("y", "v"), ("z", "w")), E):
        # trilinear weights sum to 1 only to round-off, hence the tolerance
        np.testing.assert_allclose(st.parts[0][vk], p0[vk] + 2 * qomdt2 * e, rtol=0, atol=1e-13)
        np.testing.assert_allclose(st.parts[0][k], p0[k] + (p0[vk] + qomdt2 * e) * 0.25,
                                   rtol=0, atol=1e-13)


def test_mover_uniform_b_is_cayley_rotation():
    st = _uniform_state(niter=1)
    B = np.array([0.02, -0.01, 0.015])
    for n, b in zip(("Bxn", "Byn", "Bzn"), B):
        st.fields[n][:] = b
    p0 = {k: v.copy() for k, v in st.parts[0].items()}
    st.mover_pc(0)
    u0 = np.stack([p0["u"], p0["v"], p0["w"]], axis=1)
    u1 = np.stack([st.parts[0]["u"], st.parts[0]["v"], st.parts[0]["w"]], axis=1)
    np.testing.assert_allclose(np.linalg.norm(u1, axis=1), np.linalg.norm(u0, axis=1),
                               rtol=1e-13)
    # implicit midpoint: rotation by 2*atan(|qomdt2 B|) about -sign(qom) B
    t = (-64.0 * 0.125) * B
    tn = np.linalg.norm(t)
    theta = 2 * np.arctan(tn)
    axis = -t / tn  # u x B force with negative qom rotates about -B
    c, s_ = np.cos(theta), np.sin(theta)
    rot = (u0 * c + np.cross(axis, u0) * s_ +
           axis[None, :] * (u0 @ axis)[:, None] * (1 - c))
    np.testing.assert_allclose(u1, rot, rtol=0, atol=1e-14)


def test_mover_gather_reproduces_linear_fields():
    """Trilinear weights reproduce a linear field exactly: pins the
    weight<->node pairing and the sign of every weight."""
    st = _uniform_state(niter=1)
    a = np.array([0.001, -0.002, 0.0015])
    b = np.array([[0.0004, -0.0003, 0.0002], [0.0001, 0.0005, -0.0002],
                  [-0.0003, 0.0002, 0.0004]])
    X, Y, Z = st.XN, st.YN, st.ZN
    for i, n in enumerate(("Ex", "Ey", "Ez")):
        st.fields[n][:] = a[i] + b[i, 0] * X + b[i, 1] * Y + b[i, 2] * Z
    p0 = {k: v.copy() for k, v in st.parts[0].items()}
    st.mover_pc(0)
    qomdt2 = -64.0 * 0.125
    for i, vk in enumerate(("u", "v", "w")):
        e_at_p = a[i] + b[i, 0] * p0["x"] + b[i, 1] * p0["y"] + b[i, 2] * p0["z"]
        np.testing.assert_allclose(st.parts[0][vk], p0[vk] + 2 * qomdt2 * e_at_p, rtol=0,
                                   atol=1e-13)


def test_mover_particle_bc_periodic_and_reflecting():
    cfg = O.GemConfig(nxc=4, nyc=4, nzc=4, Lx=1.0, Ly=1.0, Lz=1.0, dt=1.0, ns=1,
                      npcelx=[1], npcely=[1], npcelz=[1], qom=[1.0])
    st = O.State(cfg, "f64", build_ic=False)
    # free streaming across faces: x periodic, y reflecting, z periodic
    x = np.array([0.95, 0.05, 0.5, 0.5, 0.5, 0.5])
    y = np.array([0.5, 0.5, 0.95, 0.05, 0.5, 0.5])
    z = np.array([0.5, 0.5, 0.5, 0.5, 0.95, 0.05])
    u = np.array([0.1, -0.1, 0, 0, 0, 0])
    v = np.array([0, 0, 0.1, -0.1, 0, 0])
    w = np.array([0, 0, 0, 0, 0.1, -0.1])
    st.parts = [dict(x=x.copy(), y=y.copy(), z=z.copy(), u=u.copy(), v=v.copy(), w=w.copy(),
                     q=np.ones(6))]
    st.mover_pc(0)
    p = st.parts[0]
    np.testing.assert_allclose(p["x"][:2], [0.05, 0.95], atol=1e-15)   # wrapped
    np.testing.assert_allclose(p["y"][2:4], [0.95, 0.05], atol=1e-15)  # reflected
    np.testing.assert_allclose(p["v"][2:4], [-0.1, 0.1], atol=0)       # v flipped
    np.testing.assert_allclose(p["z"][4:], [0.05, 0.95], atol=1e-15)   # wrapped
    np.testing.assert_allclose(p["u"][:2], [0.1, -0.1], atol=0)        # u kept


# --------------------------------------------------------------------------
# Invariants of the deposit.
# --------------------------------------------------------------------------
def _deposit_state(npart=3000, seed=7):
    st = _uniform_state(npart=npart, seed=seed)
    st.zero_moments(0)
    st.interp_p2g(0)
    return st


def test_deposit_conserves_charge_current_pressure():
    st = _deposit_state()
    p, m = st.parts[0], st.moments[0]
    iv = st.grid.invVOL
    checks = {
        "rhon": p["q"], "Jx": p["q"] * p["u"], "Jy": p["q"] * p["v"], "Jz": p["q"] * p["w"],
        "pxx": p["q"] * p["u"] ** 2, "pxy": p["q"] * p["u"] * p["v"],
        "pxz": p["q"] * p["u"] * p["w"], "pyy": p["q"] * p["v"] ** 2,
        "pyz": p["q"] * p["v"] * p["w"], "pzz": p["q"] * p["w"] ** 2,
    }
    for n, per_particle in checks.items():
        total = np.sum(m[n]) / iv
        want = np.sum(per_particle)
        assert abs(total - want) <= 1e-12 * np.sum(np.abs(per_particle)), n


def test_deposit_linear_precision_dipole():
    """CIC reproduces the first moment: sum_n X_n rho_n V == sum_p q x_p."""
    st = _deposit_state()
    p, m = st.parts[0], st.moments[0]
    iv = st.grid.invVOL
    for coord, X in (("x", st.XN), ("y", st.YN), ("z", st.ZN)):
        got = np.sum(X * m["rhon"]) / iv
        want = np.sum(p["q"] * p[coord])
        assert abs(got - want) <= 1e-11 * abs(want), coord


def test_deposit_matches_long_double_brute_force():
    st = _deposit_state(npart=400, seed=11)
    p, g = st.parts[0], st.grid
    ld = np.longdouble
    nyn, nzn = g.nyn, g.nzn
    ref = {n: np.zeros(st.nn, ld) for n in O.MOMENT_NAMES}
    iv = ld(g.invVOL)
    for i in range(len(p["x"])):
        x, y, z = ld(p["x"][i]), ld(p["y"][i]), ld(p["z"][i])
        u, v, w, q = ld(p["u"][i]), ld(p["v"][i]), ld(p["w"][i]), ld(p["q"][i])
        ix = 2 + int(np.floor(float(x) * g.invdx))
        iy = 2 + int(np.floor(float(y) * g.invdy))
        iz = 2 + int(np.floor(float(z) * g.invdz))
        xi = [x - ld(st.XN[(ix - 1) * nyn * nzn]), ld(st.XN[ix * nyn * nzn]) - x]
        eta = [y - ld(st.YN[(iy - 1) * nzn]), ld(st.YN[iy * nzn]) - y]
        zeta = [z - ld(st.ZN[iz - 1]), ld(st.ZN[iz]) - z]
        vals = dict(rhon=1, Jx=u, Jy=v, Jz=w, pxx=u * u, pxy=u * v, pxz=u * w, pyy=v * v,
                    pyz=v * w, pzz=w * w)
        for a in range(2):
            for b in range(2):
                for c in range(2):
                    n = ((ix - a) * nyn + (iy - b)) * nzn + (iz - c)
                    wgt = q * xi[a] * eta[b] * zeta[c] * iv * iv
                    for k, val in vals.items():
                        ref[k][n] += val * wgt
    for k in O.MOMENT_NAMES:
        got = st.moments[0][k]
        scale = float(np.max(np.abs(ref[k])))
        err = float(np.max(np.abs(got.astype(ld) - ref[k])))
        assert err <= 1e-13 * scale, (k, err, scale)


def test_deposit_particle_on_node_hits_one_node():
    st = _uniform_state(npart=1)
    g = st.grid
    p = st.parts[0]
    i, j, k = 3, 4, 5  # node indices
    p["x"][0] = st.XN[i * g.nyn * g.nzn]
    p["y"][0] = st.YN[j * g.nzn]
    p["z"][0] = st.ZN[k]
    st.zero_moments(0)
    st.interp_p2g(0)
    rho = st.moments[0]["rhon"]
    n = (i * g.nyn + j) * g.nzn + k
    assert np.count_nonzero(rho) == 1 and rho[n] != 0
    np.testing.assert_allclose(rho[n], p["q"][0] * g.invVOL, rtol=1e-15)


# --------------------------------------------------------------------------
# particle_sort invariants.
# --------------------------------------------------------------------------
@pytest.mark.parametrize("sort_cps", [1, 2, 4])
def test_sort_is_stable_permutation_by_bin(sort_cps):
    st = _uniform_state(npart=5000, seed=3)
    p = st.parts[0]
    p["x"][:20] = -1.0  # out of range: must land in the trailing lost bin
    before = np.stack([p[k] for k in O.PARTICLE_NAMES], axis=1).copy()
    g = st.grid
    nb = [1 + (n - 1) // sort_cps for n in (g.nxc, g.nyc, g.nzc)]
    dist = [L / nbk for L, nbk in zip((g.Lx, g.Ly, g.Lz), nb)]
    pos = [np.floor(before[:, a] / dist[a]).astype(np.int64) for a in range(3)]
    inside = np.all([(pos[a] >= 0) & (pos[a] < nb[a]) for a in range(3)], axis=0)
    key = np.where(inside, pos[2] + pos[1] * nb[2] + pos[0] * nb[2] * nb[1],
                   nb[0] * nb[1] * nb[2])
    order = np.argsort(key, kind="stable")
    assert st.particle_sort(0, sort_cps) == 0
    after = np.stack([p[k] for k in O.PARTICLE_NAMES], axis=1)
    assert_bitwise(after, before[order], "stable bin order")
    assert np.all(np.isin(after[-20:, 0], [-1.0]))


# --------------------------------------------------------------------------
# After the deposit: the ghost-node fold (applyBCids) and the species sum
# (sumOverSpecies).
# --------------------------------------------------------------------------
def _random_moments(st, seed, integers=False):
    rng = np.random.default_rng(seed)
    for s in range(st.cfg.ns):
        for n in O.MOMENT_NAMES:
            if integers:
                st.moments[s][n][:] = rng.integers(-1000, 1000, st.nn).astype(st.dt)
            else:
                st.moments[s][n][:] = rng.normal(0, 1, st.nn).astype(st.dt)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("per", [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)])
def test_oracle_fold_matches_live_reference_all_periodicities(prec, per):
    """The fold of every periodicity combination against the reference's
    applyBCids on the same random moments: edges and corners are folded
    once per axis in the reference's order, so any order slip shows."""
    if not O.ref_available(prec):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    cfg = O.GemConfig(nxc=6, nyc=5, nzc=4, npcelx=[1] * 4, npcely=[1] * 4,
                      npcelz=[1] * 4, PERIODICX=bool(per[0]), PERIODICY=bool(per[1]),
                      PERIODICZ=bool(per[2]))
    ref = O.RefState(cfg, prec)
    st = O.State(cfg, prec, build_ic=False)
    _random_moments(st, seed=sum(per) + 4 * per[0] + 2 * per[1])
    for s in range(cfg.ns):
        for w, n in enumerate(O.MOMENT_NAMES):
            ref.set_moment(s, w, st.moments[s][n].astype(np.float64))
        ref.apply_bc_ids(s)
        st.apply_bc_ids(s)
        for w, n in enumerate(O.MOMENT_NAMES):
            assert_bitwise(st.moments[s][n], ref.moment(s, w).astype(st.dt), f"s{s} {n}")
    ref.close()


@pytest.mark.parametrize("prec", PRECS)
def test_oracle_species_sum_matches_live_reference(prec):
    if not O.ref_available(prec):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    cfg = O.GemConfig(nxc=6, nyc=5, nzc=4, npcelx=[1] * 4, npcely=[1] * 4,
                      npcelz=[1] * 4)
    ref = O.RefState(cfg, prec)
    st = O.State(cfg, prec, build_ic=False)
    _random_moments(st, seed=11)
    for s in range(cfg.ns):
        for w, n in enumerate(O.MOMENT_NAMES):
            ref.set_moment(s, w, st.moments[s][n].astype(np.float64))
    got, want = st.sum_species(), ref.sum_species()
    for n in O.MOMENT_NAMES:
        assert_bitwise(got[n], want[n].astype(st.dt), n)
    ref.close()


def test_species_sum_exact_on_integers_and_identity():
    """Integer-valued moments sum exactly in any order: the net must equal
    the exact sum, which catches a species added twice or dropped and a
    moment or node paired with the wrong one; a net that starts nonzero
    keeps its value (+=); one nonzero species passes through unchanged."""
    cfg = O.GemConfig(nxc=6, nyc=5, nzc=4, npcelx=[1] * 4, npcely=[1] * 4,
                      npcelz=[1] * 4)
    st = O.State(cfg, "f64", build_ic=False)
    _random_moments(st, seed=3, integers=True)
    start = {n: np.arange(st.nn, dtype=np.float64) * (k + 1)
             for k, n in enumerate(O.MOMENT_NAMES)}
    net = st.sum_species(start)
    for k, n in enumerate(O.MOMENT_NAMES):
        exact = start[n] + sum(st.moments[s][n] for s in range(cfg.ns))
        assert_bitwise(net[n], exact, n)
    for s in range(1, cfg.ns):
        for n in O.MOMENT_NAMES:
            st.moments[s][n][:] = 0
    net = st.sum_species()
    for n in O.MOMENT_NAMES:
        assert_bitwise(net[n], st.moments[0][n], n)
