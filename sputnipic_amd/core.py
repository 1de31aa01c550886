"""Host-side mirror of the reference's hot-path interface, over libspic.so.

The classes hold device (HBM) copies of the reference's structures:

=================  ==========================================================
``Parameters``     the members of ``struct parameters`` the path reads
                   (include/Parameters.h:10-235)
``Grid``           ``struct grid`` as ``setGrid`` builds it
                   (include/Grid.h:11-63, src/Grid.cpp:5-62)
``EMField``        ``struct EMfield`` node fields (include/EMfield.h:8-25)
``Particles``      ``struct particles`` SoA arrays (include/Particles.h:11-56)
``InterpDensSpecies``  the ten node moments of ``struct interpDensSpecies``
                   (include/InterpDensSpecies.h:9-40), FP64 on the device
=================  ==========================================================

and the module functions keep the reference's names and argument order:
``mover_PC(part, field, grd, param)``, ``interpP2G(part, ids, grd)``,
``mover_interp(part, field, ids, grd, param)``,
``particle_sort(param, part, grd)`` and ``setZeroSpeciesDensities``.
Failures raise :class:`SpicError` carrying the C status code.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, Optional, Sequence

import numpy as np

from ._capi import (SPIC_F32, SPIC_F64, SpicError, check, lib, spic_field, spic_grid,
                    spic_moments, spic_push, spic_species)

MOMENT_NAMES = ("rhon", "Jx", "Jy", "Jz", "pxx", "pxy", "pxz", "pyy", "pyz", "pzz")
FIELD_NAMES = ("Ex", "Ey", "Ez", "Bxn", "Byn", "Bzn")
PARTICLE_NAMES = ("x", "y", "z", "u", "v", "w", "q")


def _prec_code(prec: str) -> int:
    if prec not in ("f64", "f32"):
        raise ValueError("prec must be 'f64' or 'f32'")
    return SPIC_F64 if prec == "f64" else SPIC_F32


def _dtype(prec: str):
    return np.float64 if prec == "f64" else np.float32


class DeviceArray:
    """One device allocation with numpy upload/download."""

    def __init__(self, n: int, dtype, fill_zero: bool = False):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = max(1, self.n * self.dtype.itemsize)
        p = C.c_void_p()
        check(lib().spic_malloc(C.byref(p), self.nbytes), "spic_malloc")
        self.ptr = p.value
        if fill_zero:
            self.zero()

    @classmethod
    def from_numpy(cls, a: np.ndarray, dtype=None) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError(f"size {a.size} != {self.n}")
        if self.n:
            check(lib().spic_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes), "upload")

    def download(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if self.n:
            check(lib().spic_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes), "download")
        return out

    def zero(self) -> None:
        check(lib().spic_memset(self.ptr, 0, self.nbytes), "memset")

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib().spic_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class Parameters:
    """Members of ``struct parameters`` read on this path, with the
    defaults of src/RW_IO.cpp:37-225."""
    dt: float = 0.25
    c: float = 1.0
    NiterMover: int = 3
    n_sub_cycles: int = 1
    PERIODICX: bool = True
    PERIODICY: bool = False
    PERIODICZ: bool = True
    sort_cps: int = 4

    def push(self) -> spic_push:
        return spic_push(self.dt, self.c, int(self.PERIODICX), int(self.PERIODICY),
                         int(self.PERIODICZ))


def grid_constants(nxc: int, nyc: int, nzc: int, Lx: float, Ly: float, Lz: float,
                   prec: str = "f64") -> dict:
    """setGrid's spacings, FPfield inverses and node coordinates for nxc x
    nyc x nzc physical cells (src/Grid.cpp:5-62): dx = Lx / nxc in double;
    invdx = (FPfield)1.0 / dx and invVOL = (FPfield)1.0 / (dx * dy * dz),
    divided in double and narrowed on assignment; node i at
    xStart + (i - 1) * dx (double) narrowed to FPfield, i = 0 .. nxc + 2."""
    dt = _dtype(prec)
    dx, dy, dz = Lx / nxc, Ly / nyc, Lz / nzc
    return {
        "dx": dx, "dy": dy, "dz": dz,
        "invVOL": float(dt(1.0 / (dx * dy * dz))),
        "invdx": float(dt(1.0 / dx)),
        "invdy": float(dt(1.0 / dy)),
        "invdz": float(dt(1.0 / dz)),
        "xn": np.array([(i - 1) * dx for i in range(nxc + 3)], dtype=dt),
        "yn": np.array([(j - 1) * dy for j in range(nyc + 3)], dtype=dt),
        "zn": np.array([(k - 1) * dz for k in range(nzc + 3)], dtype=dt),
    }


class Grid:
    """``setGrid`` (src/Grid.cpp:5-62) for nxc x nyc x nzc physical cells on
    an Lx x Ly x Lz box: two guard cells per axis, nodes = cells + 1, node
    ``i`` at ``xStart + (i-1)*dx`` narrowed to FPfield, and the FPfield
    constants invdx = (FPfield)1/dx, invVOL = (FPfield)1/(dx*dy*dz)."""

    def __init__(self, nxc: int, nyc: int, nzc: int, Lx: float, Ly: float, Lz: float,
                 prec: str = "f64"):
        self.prec = prec
        self.nxc, self.nyc, self.nzc = nxc + 2, nyc + 2, nzc + 2
        self.nxn, self.nyn, self.nzn = self.nxc + 1, self.nyc + 1, self.nzc + 1
        k = grid_constants(nxc, nyc, nzc, Lx, Ly, Lz, prec)
        self.dx, self.dy, self.dz = k["dx"], k["dy"], k["dz"]
        self.invVOL = k["invVOL"]
        self.invdx, self.invdy, self.invdz = k["invdx"], k["invdy"], k["invdz"]
        self.xStart = self.yStart = self.zStart = 0.0
        self.Lx, self.Ly, self.Lz = Lx, Ly, Lz
        self.xn_host, self.yn_host, self.zn_host = k["xn"], k["yn"], k["zn"]
        self._xn = DeviceArray.from_numpy(self.xn_host)
        self._yn = DeviceArray.from_numpy(self.yn_host)
        self._zn = DeviceArray.from_numpy(self.zn_host)
        self.nn = self.nxn * self.nyn * self.nzn

    @property
    def shape(self):
        return (self.nxn, self.nyn, self.nzn)

    def c_struct(self) -> spic_grid:
        return spic_grid(self.nxc, self.nyc, self.nzc, self.nxn, self.nyn, self.nzn,
                         self.dx, self.dy, self.dz, self.invdx, self.invdy, self.invdz,
                         self.invVOL, self.xStart, self.yStart, self.zStart,
                         self.Lx, self.Ly, self.Lz, _prec_code(self.prec),
                         self._xn.ptr, self._yn.ptr, self._zn.ptr)


class EMField:
    """Six node-centred field arrays in the grid's precision."""

    def __init__(self, grd: Grid, arrays: Optional[Dict[str, np.ndarray]] = None):
        self.grd = grd
        dt = _dtype(grd.prec)
        self.d = {n: DeviceArray(grd.nn, dt, fill_zero=arrays is None) for n in FIELD_NAMES}
        if arrays is not None:
            self.upload(arrays)

    def upload(self, arrays: Dict[str, np.ndarray]) -> None:
        for n in FIELD_NAMES:
            self.d[n].upload(np.asarray(arrays[n]).reshape(-1))

    def c_struct(self) -> spic_field:
        return spic_field(*[self.d[n].ptr for n in FIELD_NAMES])


class Particles:
    """One species: SoA x, y, z, u, v, w, q of npmax elements, nop active."""

    def __init__(self, arrays: Dict[str, np.ndarray], qom: float, prec: str = "f64",
                 NiterMover: int = 3, n_sub_cycles: int = 1, nop: Optional[int] = None):
        dt = _dtype(prec)
        self.prec = prec
        self.npmax = int(len(arrays["x"]))
        self.nop = self.npmax if nop is None else int(nop)
        self.qom = float(qom)
        self.NiterMover = NiterMover
        self.n_sub_cycles = n_sub_cycles
        self.d = {k: DeviceArray.from_numpy(np.asarray(arrays[k]), dt) for k in PARTICLE_NAMES}

    def download(self) -> Dict[str, np.ndarray]:
        return {k: self.d[k].download() for k in PARTICLE_NAMES}

    def c_struct(self) -> spic_species:
        return spic_species(*[self.d[k].ptr for k in PARTICLE_NAMES], self.nop, self.npmax,
                            self.qom, self.NiterMover, self.n_sub_cycles)


class InterpDensSpecies:
    """The ten node moments of one species, FP64 on the device."""

    def __init__(self, grd: Grid):
        self.grd = grd
        self.d = {n: DeviceArray(grd.nn, np.float64, fill_zero=True) for n in MOMENT_NAMES}

    def download(self) -> Dict[str, np.ndarray]:
        return {n: self.d[n].download() for n in MOMENT_NAMES}

    def upload(self, arrays: Dict[str, np.ndarray]) -> None:
        for n in MOMENT_NAMES:
            self.d[n].upload(np.ascontiguousarray(arrays[n], np.float64))

    def c_struct(self) -> spic_moments:
        D = C.POINTER(C.c_double)
        return spic_moments(*[C.cast(self.d[n].ptr, D) for n in MOMENT_NAMES])


class InterpDensNet(InterpDensSpecies):
    """The net moments summed over species (include/InterpDensNet.h:9-40):
    the ten node moments and the cell-centred rhoc, FP64 on the device."""

    def __init__(self, grd: Grid):
        super().__init__(grd)
        ncell = grd.nxc * grd.nyc * grd.nzc
        self.rhoc = DeviceArray(ncell, np.float64, fill_zero=True)


class InterpDensAux:
    """struct interpDens_aux (include/InterpDens_aux.h:6-13): rho hat on the
    cells, J hat on the nodes, FP64 on the device."""

    def __init__(self, grd: Grid):
        self.grd = grd
        ncell = grd.nxc * grd.nyc * grd.nzc
        self.rhoh = DeviceArray(ncell, np.float64, fill_zero=True)
        self.J = {n: DeviceArray(grd.nn, np.float64, fill_zero=True)
                  for n in ("Jxh", "Jyh", "Jzh")}

    def download(self) -> Dict[str, np.ndarray]:
        out = {n: a.download() for n, a in self.J.items()}
        out["rhoh"] = self.rhoh.download()
        return out

    def c_struct(self):
        from ._capi import spic_hat
        D = C.POINTER(C.c_double)
        return spic_hat(C.cast(self.rhoh.ptr, D), *[C.cast(self.J[n].ptr, D)
                                                    for n in ("Jxh", "Jyh", "Jzh")])


@dataclass
class SolveParams:
    """The members of struct parameters the field-solve operators read
    (include/Parameters.h): dt, c, th, fourpi, periodicity, qom per species;
    defaults as src/RW_IO.cpp and inputfiles/GEM_3D.inp."""
    dt: float = 0.25
    c: float = 1.0
    th: float = 1.0
    fourpi: float = 12.5663706144
    PERIODICX: bool = True
    PERIODICY: bool = False
    PERIODICZ: bool = True
    qom: Sequence[float] = (-64.0, 1.0, -64.0, 1.0)

    def c_struct(self):
        from ._capi import spic_solve_params
        q = (C.c_double * 8)(*[float(v) for v in self.qom][:8])
        return spic_solve_params(self.dt, self.c, self.th, self.fourpi, int(self.PERIODICX),
                                 int(self.PERIODICY), int(self.PERIODICZ), q)


_WORK: Dict[tuple, "DeviceArray"] = {}


def _workspace(grd: Grid, ns: int) -> "DeviceArray":
    g = grd.c_struct()
    need = int(lib().spic_field_workspace_bytes(C.byref(g), ns))
    key = (grd.nxn, grd.nyn, grd.nzn)
    w = _WORK.get(key)
    if w is None or w.n * 8 < need:
        w = DeviceArray(max(need // 8, 1), np.float64)
        _WORK[key] = w
    return w


def applyBCscalar(a: "DeviceArray", grd: Grid, param: SolveParams, kind: int) -> None:
    """Ghost update in place: kind 0 applyBCscalarDensN, 1
    applyBCscalarFieldNzero, 2 applyBCscalarDensC, 3 applyBCscalarFieldC
    (src/BC.cpp)."""
    g, p = grd.c_struct(), param.c_struct()
    check(lib().spic_bc_scalar(C.byref(g), a.ptr, kind, C.byref(p), None), "spic_bc_scalar")


def interpN2Crho_net(idn: InterpDensNet, grd: Grid, param: SolveParams) -> None:
    """applyBCscalarDensN(idn.rhon) then interpN2Cinterp(idn.rhoc, idn.rhon),
    the driver's step before the solver (src/sputniPIC_CPU.cpp:276-277)."""
    g, p = grd.c_struct(), param.c_struct()
    check(lib().spic_rhoc_from_rhon(C.byref(g), idn.d["rhon"].ptr, idn.rhoc.ptr, C.byref(p),
                                    None), "spic_rhoc_from_rhon")


def calculateHatDensities(id_aux: InterpDensAux, idn: InterpDensNet,
                          ids: Sequence[InterpDensSpecies], field: "EMField", grd: Grid,
                          param: SolveParams, stream: Optional["Stream"] = None) -> None:
    """calculateHatDensities (src/MaxwellSolver.cpp:465-604), same argument
    order; ns = len(ids)."""
    ns = len(ids)
    g, f, p, h = grd.c_struct(), field.c_struct(), param.c_struct(), id_aux.c_struct()
    moms = (spic_moments * ns)(*[m.c_struct() for m in ids])
    w = _workspace(grd, ns)
    check(lib().spic_hat_densities(C.byref(g), C.byref(f), moms, ns, idn.rhoc.ptr, C.byref(p),
                                   C.byref(h), w.ptr, w.n * 8, stream.ptr if stream else None),
          "spic_hat_densities")


def MaxwellImage(im: "DeviceArray", vector: "DeviceArray", field: "EMField",
                 ids: Sequence[InterpDensSpecies], grd: Grid, param: SolveParams,
                 stream: Optional["Stream"] = None) -> None:
    """MaxwellImage (src/MaxwellSolver.cpp:95-257), same argument order:
    im = A vector over the 3 (nxn-2)(nyn-2)(nzn-2) interior unknowns."""
    ns = len(ids)
    g, f, p = grd.c_struct(), field.c_struct(), param.c_struct()
    moms = (spic_moments * max(ns, 1))(*[m.c_struct() for m in ids])
    w = _workspace(grd, ns)
    check(lib().spic_maxwell_image(C.byref(g), C.byref(f), moms, ns, C.byref(p), vector.ptr,
                                   im.ptr, w.ptr, w.n * 8, stream.ptr if stream else None),
          "spic_maxwell_image")


def _check_prec(*objs):
    precs = {o.prec for o in objs if hasattr(o, "prec")}
    if len(precs) > 1:
        raise ValueError(f"mixed precisions {precs}")


def set_strict_fp(strict: bool) -> None:
    """Reference rounding order (True) or production arithmetic (False)."""
    check(lib().spic_set_strict_fp(int(bool(strict))), "spic_set_strict_fp")


def set_subcycling(mode: int) -> None:
    """Sub-cycling of the mover for n_sub_cycles != 1 (spic_set_subcycling):
    SPIC_SUBCYCLE_REFUSE (default), SPIC_SUBCYCLE_LOOP (the reference GPU
    mover's loop, src/gpu/Particles_gpu.cu:336-437) or SPIC_SUBCYCLE_SINGLE
    (the reference CPU mover's single step of dt / n_sub_cycles,
    src/Mover.cpp:11)."""
    check(lib().spic_set_subcycling(int(mode)), "spic_set_subcycling")


_ASYNC = [False]


def set_async(async_mode: bool) -> None:
    check(lib().spic_set_async(int(bool(async_mode))), "spic_set_async")
    _ASYNC[0] = bool(async_mode)


def get_async() -> bool:
    """The mode last set through set_async (False at start)."""
    return _ASYNC[0]


def check_domain(stream: Optional["Stream"] = None) -> None:
    """Wait for the stream and raise if any call since the last check found
    particles outside the grid (the status an async-mode call leaves)."""
    check(lib().spic_check_domain(stream.ptr if stream else None), "spic_check_domain")


def mover_PC(part: Particles, field: EMField, grd: Grid, param: Parameters) -> int:
    """mover_PC (src/Mover.cpp:4-162); returns 0 like the reference."""
    _check_prec(part, grd)
    g, f, s, p = grd.c_struct(), field.c_struct(), part.c_struct(), param.push()
    check(lib().spic_mover_pc(C.byref(g), C.byref(f), C.byref(s), C.byref(p), None),
          "mover_PC")
    return 0


def interpP2G(part: Particles, ids: InterpDensSpecies, grd: Grid) -> None:
    """interpP2G (src/Mover.cpp:385-572): adds this species' moments."""
    _check_prec(part, grd)
    g, s, m = grd.c_struct(), part.c_struct(), ids.c_struct()
    check(lib().spic_interp_p2g(C.byref(g), C.byref(s), C.byref(m), None), "interpP2G")


def mover_interp(part: Particles, field: EMField, ids: InterpDensSpecies, grd: Grid,
                 param: Parameters) -> int:
    """mover_interp (src/Mover.cpp:575-862): fused push + deposit."""
    _check_prec(part, grd)
    g, f, s, m, p = (grd.c_struct(), field.c_struct(), part.c_struct(), ids.c_struct(),
                     param.push())
    check(lib().spic_mover_interp(C.byref(g), C.byref(f), C.byref(s), C.byref(m),
                                  C.byref(p), None), "mover_interp")
    return 0


def particle_sort(param: Parameters, part: Particles, grd: Grid) -> None:
    """particle_sort (src/Particles.cpp:72-234), reference output order."""
    _check_prec(part, grd)
    g, s = grd.c_struct(), part.c_struct()
    check(lib().spic_particle_sort(C.byref(g), int(param.sort_cps), C.byref(s), None),
          "particle_sort")


def setZeroSpeciesDensities(ids: Sequence[InterpDensSpecies], grd: Grid, ns: int,
                            stream: Optional["Stream"] = None) -> None:
    """setZeroSpeciesDensities (src/InterpDensNet.cpp:40-75), node moments."""
    g = grd.c_struct()
    for is_ in range(ns):
        m = ids[is_].c_struct()
        check(lib().spic_zero_moments(C.byref(g), C.byref(m), stream.ptr if stream else None),
              "zero moments")


def applyBCids(ids: InterpDensSpecies, grd: Grid, param: Parameters) -> None:
    """applyBCids (src/BC.cpp:362-619): fold one species' ghost-node
    moments in place on the device."""
    g, m, p = grd.c_struct(), ids.c_struct(), param.push()
    check(lib().spic_apply_bc_ids(C.byref(g), C.byref(m), 1, C.byref(p), None),
          "applyBCids")


def applyBCids_all(ids: Sequence[InterpDensSpecies], grd: Grid, param: Parameters) -> None:
    """applyBCids for every species at once (one launch per fold step)."""
    g, p = grd.c_struct(), param.push()
    moms = (spic_moments * len(ids))(*[m.c_struct() for m in ids])
    check(lib().spic_apply_bc_ids(C.byref(g), moms, len(ids), C.byref(p), None),
          "applyBCids")


def sumOverSpecies(idn: InterpDensNet, ids: Sequence[InterpDensSpecies], grd: Grid,
                   ns: int) -> None:
    """sumOverSpecies (src/InterpDensNet.cpp:120-144): idn += ids[0..ns-1]."""
    g, net = grd.c_struct(), idn.c_struct()
    moms = (spic_moments * ns)(*[ids[i].c_struct() for i in range(ns)])
    check(lib().spic_sum_species(C.byref(g), C.byref(net), moms, ns, None),
          "sumOverSpecies")


def device_count() -> int:
    n = C.c_int(0)
    rc = lib().spic_device_count(C.byref(n))
    return n.value if rc == 0 else 0


# --------------------------------------------------------------------------
# Synthetic Harris-sheet input on the device (spic_gem_fields /
# spic_gem_particles): initGEM's formulas, counter-based Maxwellian draws.
# --------------------------------------------------------------------------
@dataclass
class GemParams:
    """Harris-sheet constants (inputfiles/GEM_3D.inp; pertX of src/IC.cpp:11)."""
    B0x: float = 0.0195
    B0y: float = 0.0
    B0z: float = 0.0
    delta: float = 0.5
    pertX: float = 0.4
    fourpi: float = 12.5663706144

    def c_struct(self):
        from ._capi import spic_gem_params
        return spic_gem_params(self.B0x, self.B0y, self.B0z, self.delta, self.pertX,
                               self.fourpi)


@dataclass
class GemSpecies:
    """One species of the generator; ``sheet`` marks the current-sheet
    density profile (species 0 and 1 of the reference IC)."""
    qom: float
    rhoINIT: float
    npcelx: int = 3
    npcely: int = 3
    npcelz: int = 3
    uth: float = 0.045
    vth: float = 0.045
    wth: float = 0.045
    u0: float = 0.0
    v0: float = 0.0
    w0: float = 0.0
    sheet: bool = True

    def count(self, grd: Grid) -> int:
        # the reference's lattice loop runs its z index to npcely
        # (src/IC.cpp:128), so a cell holds npcelx * npcely * npcely
        cells = (grd.nxc - 2) * (grd.nyc - 2) * (grd.nzc - 2)
        return cells * self.npcelx * self.npcely * self.npcely


def harris_species(ppc=(3, 3, 3)):
    """The four species of inputfiles/GEM_3D.inp."""
    px, py, pz = ppc
    return [
        GemSpecies(-64.0, 1.0, px, py, pz, 0.045, 0.045, 0.045, 0, 0, 0.0065, True),
        GemSpecies(1.0, 1.0, px, py, pz, 0.0126, 0.0126, 0.0126, 0, 0, -0.0325, True),
        GemSpecies(-64.0, 0.1, px, py, pz, 0.045, 0.045, 0.045, 0, 0, 0.0, False),
        GemSpecies(1.0, 0.1, px, py, pz, 0.0126, 0.0126, 0.0126, 0, 0, 0.0, False),
    ]


def beam_species(ppc=(1, 5, 5), drift: float = 0.1, uth: float = 0.02,
                 rho: float = 0.1):
    """Four uniform beams along x: counter-streaming electrons (qom -64,
    u0 = +drift / -drift) and ions (qom 1, likewise), thermal speed uth."""
    px, py, pz = ppc
    return [
        GemSpecies(-64.0, rho, px, py, pz, uth, uth, uth, +drift, 0, 0, False),
        GemSpecies(-64.0, rho, px, py, pz, uth, uth, uth, -drift, 0, 0, False),
        GemSpecies(1.0, rho, px, py, pz, uth, uth, uth, +drift, 0, 0, False),
        GemSpecies(1.0, rho, px, py, pz, uth, uth, uth, -drift, 0, 0, False),
    ]


def gem_fields(grd: Grid, gem: GemParams) -> EMField:
    from ._capi import spic_gem_params  # noqa: F401
    field = EMField(grd)
    g, f, p = grd.c_struct(), field.c_struct(), gem.c_struct()
    check(lib().spic_gem_fields(C.byref(g), C.byref(p), C.byref(f), None), "spic_gem_fields")
    return field


class DeviceParticles(Particles):
    """Particles allocated on the device without a host copy."""

    def __init__(self, n: int, qom: float, prec: str = "f64", NiterMover: int = 3,
                 n_sub_cycles: int = 1):
        dt = _dtype(prec)
        self.prec = prec
        self.npmax = self.nop = int(n)
        self.qom = float(qom)
        self.NiterMover = NiterMover
        self.n_sub_cycles = n_sub_cycles
        self.d = {k: DeviceArray(self.npmax, dt) for k in PARTICLE_NAMES}


def gem_particles(grd: Grid, gem: GemParams, sp: GemSpecies, species: int, seed: int = 0,
                  NiterMover: int = 3) -> DeviceParticles:
    from ._capi import spic_gem_species
    part = DeviceParticles(sp.count(grd), sp.qom, grd.prec, NiterMover)
    c = spic_gem_species(species, seed, int(sp.sheet), sp.npcelx, sp.npcely, sp.npcelz,
                         sp.qom, sp.rhoINIT, sp.u0, sp.v0, sp.w0, sp.uth, sp.vth, sp.wth)
    g, p, s = grd.c_struct(), gem.c_struct(), part.c_struct()
    check(lib().spic_gem_particles(C.byref(g), C.byref(p), C.byref(c), C.byref(s), None),
          "spic_gem_particles")
    return part


class Stream:
    def __init__(self):
        p = C.c_void_p()
        check(lib().spic_stream_create(C.byref(p)), "stream create")
        self.ptr = p.value

    def synchronize(self):
        check(lib().spic_stream_synchronize(self.ptr), "stream sync")


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib().spic_event_create(C.byref(p)), "event create")
        self.ptr = p.value

    def record(self, stream: Optional[Stream] = None):
        check(lib().spic_event_record(self.ptr, stream.ptr if stream else None), "event record")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib().spic_event_elapsed_ms(C.byref(ms), self.ptr, end.ptr), "event elapsed")
        return ms.value


class Graph:
    """A captured run of calls on one stream (spic_graph_*), replayed with
    one launch.  Use as ``with Graph(stream) as g: ...`` (capture; async mode
    is switched on for the capture), then ``g.launch()``."""

    def __init__(self, stream: Stream):
        self.stream = stream
        self.ptr = None

    def __enter__(self):
        set_async(True)
        check(lib().spic_graph_begin(self.stream.ptr), "spic_graph_begin")
        return self

    def __exit__(self, exc_type, exc, tb):
        g = C.c_void_p()
        rc = lib().spic_graph_end(self.stream.ptr, C.byref(g))
        if exc_type is None:
            check(rc, "spic_graph_end")
            self.ptr = g.value
        return False

    def launch(self) -> None:
        check(lib().spic_graph_launch(self.ptr, self.stream.ptr), "spic_graph_launch")

    def close(self) -> None:
        if self.ptr:
            lib().spic_graph_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """Tile-resident engine (spic_engine_*): every species' mover_PC +
    interpP2G in one launch over HBM-resident particles binned by tile."""

    def __init__(self, grd: Grid, ns: int, tile=(4, 4, 4), sort_every: int = 0,
                 lds_partials: int = 0, slack_pct: int = 0, set_moments: bool = False,
                 sort_mode: int = 0, positions: int = 0, order: int = 0,
                 overflow_pct: int = 0, particles_per_thread: int = 0):
        """positions: 0 (default: fixed point for an FP32 engine created in
        production arithmetic, else float), SPIC_POS_FLOAT or
        SPIC_POS_FIXED.  order: 0 / SPIC_ORDER_LOAD (store in load order) or
        SPIC_ORDER_ENGINE (no per-particle index; store in the engine's
        order).  overflow_pct: 0 (default, n overflow slots) or 1..100.
        particles_per_thread: 0 / 1 (default) or 2 (FP32 fixed point): the
        particles each thread pushes and bins per batch of the phase."""
        from ._capi import spic_engine_opts
        self.grd = grd
        self.ns = ns
        p = C.c_void_p()
        opts = spic_engine_opts(tile[0], tile[1], tile[2], sort_every, lds_partials,
                                slack_pct, int(bool(set_moments)), int(sort_mode),
                                int(positions), int(order), int(overflow_pct),
                                int(particles_per_thread))
        g = grd.c_struct()
        check(lib().spic_engine_create(C.byref(p), C.byref(g), ns, C.byref(opts)),
              "spic_engine_create")
        self.ptr = p.value

    def load(self, is_: int, part: Particles, stream: Optional["Stream"] = None) -> None:
        """Load species is_ from `part`, whose precision may be FP64 for an
        FP32 engine (spic_engine_load_as)."""
        s = part.c_struct()
        check(lib().spic_engine_load_as(self.ptr, is_, C.byref(s), _prec_code(part.prec),
                                        stream.ptr if stream else None), "spic_engine_load")

    def store(self, is_: int, part: Particles, stream: Optional["Stream"] = None) -> None:
        """Store species is_ into `part` in load order, in part's precision
        (spic_engine_store_as)."""
        s = part.c_struct()
        check(lib().spic_engine_store_as(self.ptr, is_, C.byref(s), _prec_code(part.prec),
                                         stream.ptr if stream else None), "spic_engine_store")

    def phase(self, field: EMField, ids: Sequence[InterpDensSpecies], param: Parameters,
              stream: Optional["Stream"] = None) -> None:
        moms = (spic_moments * self.ns)(*[ids[i].c_struct() for i in range(self.ns)])
        f, p = field.c_struct(), param.push()
        check(lib().spic_engine_phase(self.ptr, C.byref(f), moms, C.byref(p),
                                      stream.ptr if stream else None), "spic_engine_phase")

    def deposit(self, ids: Sequence[InterpDensSpecies],
                stream: Optional["Stream"] = None) -> None:
        """interpP2G of every loaded species, no push (spic_engine_deposit)."""
        moms = (spic_moments * self.ns)(*[ids[i].c_struct() for i in range(self.ns)])
        check(lib().spic_engine_deposit(self.ptr, moms, stream.ptr if stream else None),
              "spic_engine_deposit")

    def sort(self, stream: Optional["Stream"] = None) -> None:
        check(lib().spic_engine_sort(self.ptr, stream.ptr if stream else None),
              "spic_engine_sort")

    def stats(self) -> dict:
        from ._capi import spic_engine_stats
        st = spic_engine_stats()
        check(lib().spic_engine_get_stats(self.ptr, C.byref(st)), "stats")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self) -> None:
        if getattr(self, "ptr", None):
            lib().spic_engine_destroy(self.ptr)
            self.ptr = None


class Batch:
    """Independent PIC instances advanced together (spic_batch_*): one
    launch per kernel and phase over every instance's tiles.

    members: (engine, field, moment sets, parameters) per instance, the
    engines' species already loaded.  The descriptors are bound here; the
    arrays behind them must outlive the batch.  set_moments: each phase
    writes the moments instead of adding to them (no zeroing needed)."""

    def __init__(self, members, sort_every: int = 0, set_moments: bool = False):
        from ._capi import spic_batch_member
        n = len(members)
        self._keep = []
        arr = (spic_batch_member * n)()
        for i, (eng, field, ids, param) in enumerate(members):
            f = field.c_struct()
            moms = (spic_moments * eng.ns)(*[ids[j].c_struct() for j in range(eng.ns)])
            p = param.push()
            self._keep += [f, moms, p]
            arr[i].engine = eng.ptr
            arr[i].field = C.pointer(f)
            arr[i].moments = C.cast(moms, C.POINTER(spic_moments))
            arr[i].push = C.pointer(p)
        self.engines = [m[0] for m in members]
        self.set_moments = bool(set_moments)
        ptr = C.c_void_p()
        check(lib().spic_batch_create(C.byref(ptr), arr, n, sort_every, int(bool(set_moments))),
              "spic_batch_create")
        self.ptr = ptr.value

    def phase(self, stream: Optional["Stream"] = None) -> None:
        check(lib().spic_batch_phase(self.ptr, stream.ptr if stream else None),
              "spic_batch_phase")

    def close(self) -> None:
        if getattr(self, "ptr", None):
            lib().spic_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
