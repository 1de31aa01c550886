"""Snapshot file formats on the host (no GPU): the directory tree that
mirrors HDF5_Write_Particles' file (src/RW_IO.cpp:376-460) round-trips
its attributes and big-endian float32 datasets, and the raw dump names
follow save_ic_data (src/IC.cpp:259-313) for known cases."""
from __future__ import annotations

import os
import struct

import numpy as np

from sputnipic_amd import snapshot as SN


def test_layout_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    species = [{d: rng.normal(0, 1, n).astype(np.float64) for d in SN.DATASETS}
               for n in (5, 0, 17)]
    path = SN.write_layout(str(tmp_path / "tracked_particles_proc0_7"), 7, species)
    attrs, got = SN.read_layout(path)
    assert attrs == {"Cycle": 7, "Process": 0, "No. Processes": 1}
    assert len(got) == 3
    for s, arrs in enumerate(species):
        for d in SN.DATASETS:
            want = arrs[d].astype(">f4")
            assert got[s][d].dtype == np.dtype(">f4")
            assert np.array_equal(got[s][d].view(">u4"), want.view(">u4")), (s, d)
    # the bytes on disk are big-endian IEEE float32
    raw = open(os.path.join(path, "species_0", "x.f32be"), "rb").read()
    assert raw[:4] == struct.pack(">f", float(np.float32(species[0]["x"][0])))


def test_h5_and_ic_names():
    assert SN.h5_path("out", 10) == os.path.join("out", "tracked_particles_proc0_10.h5")
    assert SN.h5_path("d", 3, rank=2, name="tp") == os.path.join("d", "tp_proc2_3.h5")
    assert SN.ic_file_name("x", (131, 67, 67), 65536000, 2) == "x_131_67_67_2_65536000.bin"
    assert SN.ic_file_name("rhon", (131, 67, 67), 1000) == "rhon_131_67_67_1000.bin"
    assert SN.ic_file_name("rhoc", (130, 66, 66), 1000) == "rhoc_130_66_66_1000.bin"


def test_ic_particles_written_native(tmp_path):
    class G:
        nxn, nyn, nzn, nxc, nyc, nzc = 5, 4, 4, 4, 3, 3
    parts = {k: np.arange(6, dtype=np.float64) * (i + 1) for i, k in
             enumerate(("x", "y", "z", "u", "v", "w", "q"))}
    files = SN.write_ic_particles(str(tmp_path), G, 1, np.ones(80), np.ones(36), parts,
                                  np.float32, np.float32)
    names = sorted(os.path.basename(f) for f in files)
    assert "x_5_4_4_1_6.bin" in names and "rhoc_4_3_3_6.bin" in names
    x = np.fromfile(os.path.join(str(tmp_path), "x_5_4_4_1_6.bin"), dtype=np.float32)
    assert np.array_equal(x, parts["x"].astype(np.float32))


# --------------------------------------------------------------------------
# The whole IC dump against the reference's own save_ic_data / read_ic_data
# (oracle/_ref, built where the reference checkout exists).
# --------------------------------------------------------------------------
import oracle as O  # noqa: E402
import pytest  # noqa: E402


class _Grid:
    def __init__(self, ref):
        self.nxc, self.nyc, self.nzc = ref.grid_ints[:3]
        self.nxn, self.nyn, self.nzn = ref.grid_ints[3:6]


def _ref_state_arrays(ref, cfg):
    fields = {n: ref.field(w) for w, n in enumerate(O.FIELD_NAMES)}
    aux = {n: ref.aux(w) for w, n in enumerate(("Bxc", "Byc", "Bzc", "Exth", "Eyth", "Ezth"))}
    species = []
    for s in range(cfg.ns):
        sp = dict(ref.particles(s))
        sp["rhon"] = ref.moment(s, 0)
        sp["rhoc"] = ref.moment(s, 10)
        species.append(sp)
    return fields, aux, species


CFG_IC = dict(nxc=8, nyc=6, nzc=4, npcelx=[2, 1, 2, 1], npcely=[1, 2, 2, 2],
              npcelz=[1, 2, 2, 2])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ic_dump_equals_reference_save_ic_data(tmp_path, prec):
    """initGEM with a dump directory writes save_ic_data's files; the same
    arrays through write_ic_data -- the cell-centred B computed by
    interp_n2c_field, not taken from the reference -- give the same file
    names and the same bytes.  Species 1 and 3 share npmax, so their rhon and
    rhoc files are one file, holding the last species', as in the
    reference."""
    if not O.ref_available(prec):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    cfg = O.GemConfig(**CFG_IC)
    ref_dir = tmp_path / "ref"
    ref_dir.mkdir()
    ref = O.RefState(cfg, prec, io_dir=str(ref_dir))
    fields, aux, species = _ref_state_arrays(ref, cfg)
    grd = _Grid(ref)
    ref.close()
    dt = np.float64 if prec == "f64" else np.float32
    ours = tmp_path / "ours"
    got = SN.write_ic_data(str(ours), grd, fields, species, dt, dt, dt)
    names = sorted(os.listdir(ref_dir))
    assert sorted(os.path.basename(p) for p in set(got)) == names
    assert len(names) == 12 + 7 * cfg.ns + 2 * 3  # 3 distinct npmax
    for n in names:
        a = (ref_dir / n).read_bytes()
        b = (ours / n).read_bytes()
        assert a == b, n
    # and the cell-centred B from interp_n2c_field is the reference's
    for n in ("Bxc", "Byc", "Bzc"):
        want = aux[n].astype(dt)
        assert np.array_equal(SN.interp_n2c_field(grd, fields[n[:2] + "n"].astype(dt)), want)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ic_dump_read_back_by_the_reference(tmp_path, prec):
    """A dump written by write_ic_data from changed inputs (particles
    reversed and shifted, fields scaled) is what the reference's
    read_ic_data loads on restart, and read_ic_data returns it unchanged."""
    if not O.ref_available(prec):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    cfg = O.GemConfig(**CFG_IC)
    ref = O.RefState(cfg, prec)
    fields, aux, species = _ref_state_arrays(ref, cfg)
    grd = _Grid(ref)
    npmax = [ref.nop(s) for s in range(cfg.ns)]
    ref.close()
    dt = np.float64 if prec == "f64" else np.float32
    fields = {n: (1.5 * a).astype(dt) for n, a in fields.items()}
    aux = {n: (a + 0.25).astype(dt) for n, a in aux.items()}
    for sp in species:
        for k in SN.IC_SPECIES:
            sp[k] = (sp[k][::-1] * 0.5 + 0.125).astype(dt)
    d = tmp_path / "dump"
    # distinct npmax per species would be needed for distinct densities;
    # the reference reads the shared file for both species 1 and 3
    SN.write_ic_data(str(d), grd, fields, species, dt, dt, dt, aux=aux)
    f2, a2, s2 = SN.read_ic_data(str(d), grd, npmax, dt, dt, dt)
    for n in fields:
        assert np.array_equal(f2[n], fields[n])
    for n in aux:
        assert np.array_equal(a2[n], aux[n])
    last = {n: s for s, n in enumerate(npmax)}
    for s, sp in enumerate(species):
        for k in SN.IC_SPECIES:
            assert np.array_equal(s2[s][k], sp[k]), (s, k)
        for k in ("rhon", "rhoc"):
            assert np.array_equal(s2[s][k], species[last[npmax[s]]][k].astype(dt)), (s, k)
    ref = O.RefState(cfg, prec, io_dir=str(d), restart=True)
    try:
        for w, n in enumerate(O.FIELD_NAMES):
            assert np.array_equal(ref.field(w), fields[n].astype(np.float64)), n
        for w, n in enumerate(("Bxc", "Byc", "Bzc", "Exth", "Eyth", "Ezth")):
            assert np.array_equal(ref.aux(w), aux[n].astype(np.float64)), n
        for s, sp in enumerate(species):
            rp = ref.particles(s)
            for k in SN.IC_SPECIES:
                assert np.array_equal(rp[k], sp[k].astype(np.float64)), (s, k)
            assert np.array_equal(ref.moment(s, 0), s2[s]["rhon"].astype(np.float64))
            assert np.array_equal(ref.moment(s, 10), s2[s]["rhoc"].astype(np.float64))
    finally:
        ref.close()
