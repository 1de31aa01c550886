"""Particle snapshots and raw dumps in the reference's formats.

* ``HDF5_Write_Particles`` (src/RW_IO.cpp:376-460): one file per rank and
  cycle, ``<dir>/<tracked_particles_filename>_proc<rank>_<cycle>.h5``, with
  the file attributes ``Cycle``, ``Process`` and ``No. Processes`` and, per
  species, a group ``/species_<k>`` holding the datasets ``x y z u v w``
  (no ``q``) as ``H5T_IEEE_F32BE``.  The device produces the dataset words
  (``spic_engine_snapshot_be32`` for the binned engine, in load order;
  ``spic_snapshot_be32`` for caller-order arrays); :class:`Snapshot` copies
  them to pinned host memory on a side stream, so the next particle phase
  can run while the copy is in flight.  No HDF5 library ships in this
  image: :func:`write_h5` needs h5py, and :func:`write_layout` writes the
  same attributes, groups and datasets as a directory tree instead (one
  ``<dataset>.f32be`` file of big-endian float32 per dataset, the
  attributes in ``attrs.json``).
* ``save_ic_data`` / ``read_ic_data`` (src/IC.cpp:190-313, 316-439): the
  initial-condition dump a restart reads back, raw native arrays, one file
  each: the node fields ``Bxn Byn Bzn``, the cell-centred ``Bxc Byc Bzc``,
  the node fields ``Exn Eyn Ezn`` and ``Exth Eyth Ezth``
  (``<name>_<n0>_<n1>_<n2>.bin``), then per species ``rhon`` and ``rhoc``
  (``<name>_<n0>_<n1>_<n2>_<npmax>.bin``, no species index: in the reference
  two species with the same npmax share these two files, the last one
  written wins) and ``x y z u v w q`` (``<name>_<nxn>_<nyn>_<nzn>_<is>_<npmax>.bin``).
  :func:`write_ic_data` writes the whole dump and :func:`read_ic_data`
  reads it; :func:`write_ic_particles` writes one species' files.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

from ._capi import check, lib, spic_grid, spic_species  # noqa: F401
from .core import DeviceArray, Event, Stream

DATASETS = ("x", "y", "z", "u", "v", "w")


def h5_path(save_dir: str, cycle: int, rank: int = 0,
            name: str = "tracked_particles") -> str:
    """HDF5_Write_Particles' file name (src/RW_IO.cpp:383)."""
    return os.path.join(save_dir, f"{name}_proc{rank}_{cycle}.h5")


class Snapshot:
    """Device staging and pinned host buffers for the snapshot words of a
    set of species; :meth:`capture_engine` / :meth:`capture_species`
    enqueue the conversion on the compute stream and the copy on a side
    stream; :meth:`result` waits for the copy only."""

    def __init__(self, counts: Sequence[int]):
        self.counts = [int(n) for n in counts]
        self.dev: List[List[DeviceArray]] = [
            [DeviceArray(max(n, 1), np.uint32) for _ in DATASETS] for n in self.counts]
        self.host: List[C.c_void_p] = []
        for n in self.counts:
            p = C.c_void_p()
            check(lib().spic_host_alloc(C.byref(p), 6 * 4 * max(n, 1)), "spic_host_alloc")
            self.host.append(p)
        self.side = Stream()
        self.ready = Event()

    def _dst(self, s: int):
        return (C.c_void_p * 6)(*[a.ptr for a in self.dev[s]])

    def _copy_out(self, stream: Optional[Stream]) -> None:
        self.ready.record(stream)
        check(lib().spic_stream_wait_event(self.side.ptr, self.ready.ptr), "stream wait")
        for s, n in enumerate(self.counts):
            for a in range(6):
                check(lib().spic_memcpy_d2h_async(self.host[s].value + a * 4 * n,
                                                  self.dev[s][a].ptr, 4 * n, self.side.ptr),
                      "snapshot copy")

    def capture_engine(self, eng, stream: Optional[Stream] = None) -> None:
        """Snapshot every species of a loaded engine, in load order."""
        for s in range(len(self.counts)):
            check(lib().spic_engine_snapshot_be32(eng.ptr, s, self._dst(s),
                                                  stream.ptr if stream else None),
                  "spic_engine_snapshot_be32")
        self._copy_out(stream)

    def capture_species(self, grd, parts, stream: Optional[Stream] = None) -> None:
        """Snapshot caller-order species (Particles objects)."""
        g = grd.c_struct()
        for s, p in enumerate(parts):
            sp = p.c_struct()
            check(lib().spic_snapshot_be32(C.byref(g), C.byref(sp), self._dst(s),
                                           stream.ptr if stream else None),
                  "spic_snapshot_be32")
        self._copy_out(stream)

    def result(self) -> List[Dict[str, np.ndarray]]:
        """Wait for the copy; per species, x..w as big-endian float32
        arrays (copies of the pinned buffers)."""
        self.side.synchronize()
        out = []
        for s, n in enumerate(self.counts):
            buf = (C.c_char * (24 * max(n, 1))).from_address(self.host[s].value)
            words = np.frombuffer(buf, dtype=">f4", count=6 * n)
            out.append({d: words[a * n:(a + 1) * n].copy() for a, d in enumerate(DATASETS)})
        return out

    def close(self) -> None:
        for p in self.host:
            if p.value:
                lib().spic_host_free(p)
        self.host = []
        for arrs in self.dev:
            for a in arrs:
                a.free()
        self.dev = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_layout(path: str, cycle: int, species: Sequence[Dict[str, np.ndarray]],
                 rank: int = 0, nranks: int = 1) -> str:
    """The HDF5 file's content as a directory tree at `path`: attrs.json
    with Cycle, Process and No. Processes, and species_<k>/<d>.f32be for
    the six datasets (big-endian float32)."""
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "attrs.json"), "w") as f:
        json.dump({"Cycle": int(cycle), "Process": int(rank), "No. Processes": int(nranks)}, f)
    for k, arrs in enumerate(species):
        g = os.path.join(path, f"species_{k}")
        os.makedirs(g, exist_ok=True)
        for d in DATASETS:
            np.ascontiguousarray(arrs[d], dtype=">f4").tofile(os.path.join(g, f"{d}.f32be"))
    return path


def read_layout(path: str):
    """(attributes, per-species datasets) of a tree written by write_layout."""
    with open(os.path.join(path, "attrs.json")) as f:
        attrs = json.load(f)
    species = []
    k = 0
    while os.path.isdir(os.path.join(path, f"species_{k}")):
        g = os.path.join(path, f"species_{k}")
        species.append({d: np.fromfile(os.path.join(g, f"{d}.f32be"), dtype=">f4")
                        for d in DATASETS})
        k += 1
    return attrs, species


def write_h5(path: str, cycle: int, species: Sequence[Dict[str, np.ndarray]],
             rank: int = 0, nranks: int = 1) -> str:
    """The HDF5 file itself, with h5py (not part of this image)."""
    try:
        import h5py  # noqa: F401
    except ImportError as e:
        raise RuntimeError("write_h5 needs h5py, which is not importable here; "
                           "write_layout writes the same content as files") from e
    with h5py.File(path, "w") as f:
        f.attrs.create("Cycle", cycle, dtype="<i4")
        f.attrs.create("Process", rank, dtype="<i4")
        f.attrs.create("No. Processes", nranks, dtype="<i4")
        for k, arrs in enumerate(species):
            g = f.create_group(f"/species_{k}")
            for d in DATASETS:
                g.create_dataset(d, data=np.asarray(arrs[d], dtype=">f4"))
    return path


def ic_file_name(kind: str, shape, npmax: Optional[int] = None,
                 species: Optional[int] = None) -> str:
    """save_ic_data's file names (src/IC.cpp:195-313): fields
    <kind>_<n0>_<n1>_<n2>.bin, densities <kind>_<n0>_<n1>_<n2>_<npmax>.bin,
    particle arrays <kind>_<nxn>_<nyn>_<nzn>_<is>_<npmax>.bin."""
    dims = "_".join(str(int(v)) for v in shape)
    if npmax is None:
        return f"{kind}_{dims}.bin"
    if species is None:
        return f"{kind}_{dims}_{int(npmax)}.bin"
    return f"{kind}_{dims}_{int(species)}_{int(npmax)}.bin"


# save_ic_data's field files in its order: (file kind, key in the caller's
# dicts, on cells)
IC_FIELDS = (("Bxn", "Bxn", False), ("Byn", "Byn", False), ("Bzn", "Bzn", False),
             ("Bxc", "Bxc", True), ("Byc", "Byc", True), ("Bzc", "Bzc", True),
             ("Exn", "Ex", False), ("Eyn", "Ey", False), ("Ezn", "Ez", False),
             ("Exth", "Exth", False), ("Eyth", "Eyth", False), ("Ezth", "Ezth", False))
IC_SPECIES = ("x", "y", "z", "u", "v", "w", "q")


def interp_n2c_field(grd, bn: np.ndarray) -> np.ndarray:
    """interpN2Cfield (src/Grid.cpp:86-120) for one component: the mean of
    the eight nodes around each interior cell, .125 times their sum in the
    reference's order; guard cells zero.  initGEM makes the cell-centred B
    this way (src/IC.cpp:99-100)."""
    b = np.asarray(bn).reshape(grd.nxn, grd.nyn, grd.nzn)
    out = np.zeros((grd.nxc, grd.nyc, grd.nzc), dtype=b.dtype)
    i = slice(1, grd.nxc - 1)
    j = slice(1, grd.nyc - 1)
    k = slice(1, grd.nzc - 1)
    i1 = slice(2, grd.nxc)
    j1 = slice(2, grd.nyc)
    k1 = slice(2, grd.nzc)
    s = b[i, j, k] + b[i1, j, k]
    s = s + b[i, j1, k]
    s = s + b[i, j, k1]
    s = s + b[i1, j1, k]
    s = s + b[i1, j, k1]
    s = s + b[i, j1, k1]
    s = s + b[i1, j1, k1]
    out[i, j, k] = (0.125 * s.astype(np.float64)).astype(b.dtype)
    return out.reshape(-1)


def write_ic_data(restart_dir: str, grd, fields: Dict[str, np.ndarray],
                  species: Sequence[Dict[str, np.ndarray]], field_dtype, part_dtype,
                  interp_dtype, aux: Optional[Dict[str, np.ndarray]] = None) -> List[str]:
    """save_ic_data (src/IC.cpp:190-313): every file of the dump into
    restart_dir, in the reference's order, names and native types.

    fields: Ex Ey Ez Bxn Byn Bzn on the nodes.  aux: Bxc Byc Bzc (cells)
    and Exth Eyth Ezth (nodes); when omitted, as initGEM leaves them: the
    cell-centred B from interp_n2c_field and E theta zero.  species: per
    species a dict with rhon (nodes), rhoc (cells) and x y z u v w q (q in
    the FPpart type, as the reference writes it)."""
    os.makedirs(restart_dir, exist_ok=True)
    nodes = (grd.nxn, grd.nyn, grd.nzn)
    cells = (grd.nxc, grd.nyc, grd.nzc)
    if aux is None:
        aux = {"Bxc": interp_n2c_field(grd, np.asarray(fields["Bxn"], field_dtype)),
               "Byc": interp_n2c_field(grd, np.asarray(fields["Byn"], field_dtype)),
               "Bzc": interp_n2c_field(grd, np.asarray(fields["Bzn"], field_dtype))}
        for n in ("Exth", "Eyth", "Ezth"):
            aux[n] = np.zeros(grd.nxn * grd.nyn * grd.nzn, field_dtype)
    src = dict(fields)
    src.update(aux)
    written = []
    for kind, key, on_cells in IC_FIELDS:
        p = os.path.join(restart_dir, ic_file_name(kind, cells if on_cells else nodes))
        np.ascontiguousarray(src[key], dtype=field_dtype).tofile(p)
        written.append(p)
    for s, sp in enumerate(species):
        written += write_ic_particles(restart_dir, grd, s, sp["rhon"], sp["rhoc"], sp,
                                      part_dtype, interp_dtype)
    return written


def read_ic_data(restart_dir: str, grd, npmax: Sequence[int], field_dtype, part_dtype,
                 interp_dtype):
    """read_ic_data (src/IC.cpp:316-439): the dump's files back, as
    (fields, aux, species) in the shapes write_ic_data takes them; npmax per
    species names the files (and sizes the particle arrays)."""
    nodes = (grd.nxn, grd.nyn, grd.nzn)
    cells = (grd.nxc, grd.nyc, grd.nzc)
    nn = grd.nxn * grd.nyn * grd.nzn
    nc = grd.nxc * grd.nyc * grd.nzc

    def load(name, dtype, count):
        a = np.fromfile(os.path.join(restart_dir, name), dtype=dtype)
        if a.size != count:
            raise ValueError(f"{name}: {a.size} values, {count} expected")
        return a

    fields, aux = {}, {}
    for kind, key, on_cells in IC_FIELDS:
        a = load(ic_file_name(kind, cells if on_cells else nodes), field_dtype,
                 nc if on_cells else nn)
        (aux if key in ("Bxc", "Byc", "Bzc", "Exth", "Eyth", "Ezth") else fields)[key] = a
    species = []
    for s, n in enumerate(npmax):
        sp = {"rhon": load(ic_file_name("rhon", nodes, n), interp_dtype, nn),
              "rhoc": load(ic_file_name("rhoc", cells, n), interp_dtype, nc)}
        for k in IC_SPECIES:
            sp[k] = load(ic_file_name(k, nodes, n, s), part_dtype, n)
        species.append(sp)
    return fields, aux, species


def write_ic_particles(restart_dir: str, grd, species: int, rhon: np.ndarray,
                       rhoc: np.ndarray, parts: Dict[str, np.ndarray],
                       part_dtype, interp_dtype) -> List[str]:
    """save_ic_data's per-species files: rhon (nodes), rhoc (cells) in the
    FPinterp type, x y z u v w q in the FPpart type, raw native arrays."""
    os.makedirs(restart_dir, exist_ok=True)
    npmax = len(parts["x"])
    nodes = (grd.nxn, grd.nyn, grd.nzn)
    cells = (grd.nxc, grd.nyc, grd.nzc)
    written = []
    for kind, arr, shape in (("rhon", rhon, nodes), ("rhoc", rhoc, cells)):
        p = os.path.join(restart_dir, ic_file_name(kind, shape, npmax))
        np.ascontiguousarray(arr, dtype=interp_dtype).tofile(p)
        written.append(p)
    for kind in IC_SPECIES:
        p = os.path.join(restart_dir, ic_file_name(kind, nodes, npmax, species))
        np.ascontiguousarray(parts[kind], dtype=part_dtype).tofile(p)
        written.append(p)
    return written
