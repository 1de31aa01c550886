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
