"""CPU oracle for the sputniPIC particle hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only to check against or to time
as the CPU baseline: the product (``sputnipic_amd``) never imports, links or
calls it.

Two libraries are wrapped:

* ``liboracle.so`` -- the plain-C restatement in ``spic_oracle.c`` of the
  reference's CPU hot path, instantiated for FP64 and FP32 (the reference's
  default precision).  Functions and the reference lines they follow:

  ===================  ===================================================
  mover_pc             src/Mover.cpp:4-162 (mover_PC)
  interp_p2g           src/Mover.cpp:385-572 (interpP2G)
  mover_interp         src/Mover.cpp:575-862 (mover_interp)
  particle_sort        src/Particles.cpp:67-234 (particle_sort, binned)
  set_grid             src/Grid.cpp:5-62 (setGrid)
  gem_fields           src/IC.cpp:38-105 + src/InterpDensSpecies.cpp:64-77
  gem_particles        src/IC.cpp:117-177
  apply_bc_ids         src/BC.cpp:362-619 (applyBCids)
  sum_species          src/InterpDensNet.cpp:120-144 (sumOverSpecies)
  bc_scalar            src/BC.cpp:11-123, 242-352, 623-735, 737-850
                       (applyBCscalarDensN / FieldNzero / DensC / FieldC)
  rhoc_from_rhon       src/sputniPIC_CPU.cpp:276-277 (applyBCscalarDensN,
                       interpN2Cinterp src/Grid.cpp:164-185)
  hat_densities        src/MaxwellSolver.cpp:465-604 (calculateHatDensities)
  maxwell_image        src/MaxwellSolver.cpp:95-257 (MaxwellImage)
  ===================  ===================================================

  How each is pinned (tests/test_oracle.py): every function is compared bit
  for bit with the reference's own translation units (``_ref/`` below) on the
  small Harris-sheet case, in both precisions, and against the committed
  fixtures in ``tests/golden/``; mover_pc is also pinned by closed forms
  (free streaming, uniform-E kick, uniform-B rotation, exact trilinear
  reproduction of linear fields); interp_p2g by charge / current / dipole
  conservation and a long-double brute force; particle_sort by permutation
  and key-order invariants; hat_densities and maxwell_image by the
  reference's own MaxwellSolver.cpp in both precisions, for every
  periodicity.

* ``_ref/libref_{f64,f32}.so`` -- the reference's own ``src/Mover.cpp``,
  ``Grid.cpp``, ``Particles.cpp``, ``IC.cpp``, ``InterpDensSpecies.cpp``,
  ``InterpDensNet.cpp``, ``EMfield.cpp``, ``BC.cpp``, ``MaxwellSolver.cpp``,
  ``Basic.cpp`` and ``Solvers.cpp`` compiled unmodified
  from the reference checkout (``oracle/Makefile``) with the test harness in
  ``ref_harness.cpp``.  Present only where the reference checkout exists at
  build time; never where the GPU tests run without it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field as dc_field
from typing import Dict, List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

MOMENT_NAMES = ("rhon", "Jx", "Jy", "Jz", "pxx", "pxy", "pxz", "pyy", "pyz", "pzz")
FIELD_NAMES = ("Ex", "Ey", "Ez", "Bxn", "Byn", "Bzn")
PARTICLE_NAMES = ("x", "y", "z", "u", "v", "w", "q")

_P = C.c_void_p


class OGrid(C.Structure):
    _fields_ = [
        ("nxc", C.c_int), ("nyc", C.c_int), ("nzc", C.c_int),
        ("nxn", C.c_int), ("nyn", C.c_int), ("nzn", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("invdx", C.c_double), ("invdy", C.c_double), ("invdz", C.c_double),
        ("invVOL", C.c_double),
        ("xStart", C.c_double), ("yStart", C.c_double), ("zStart", C.c_double),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
    ]


class OFields(C.Structure):
    _fields_ = [(n, _P) for n in FIELD_NAMES]


class ONodes(C.Structure):
    _fields_ = [("XN", _P), ("YN", _P), ("ZN", _P)]


class OParticles(C.Structure):
    _fields_ = [(n, _P) for n in PARTICLE_NAMES] + [
        ("nop", C.c_long), ("npmax", C.c_long), ("qom", C.c_double),
        ("NiterMover", C.c_int), ("n_sub_cycles", C.c_int),
    ]


class OMoments(C.Structure):
    _fields_ = [(n, _P) for n in MOMENT_NAMES]


class OPush(C.Structure):
    _fields_ = [("dt", C.c_double), ("c", C.c_double),
                ("PERIODICX", C.c_int), ("PERIODICY", C.c_int), ("PERIODICZ", C.c_int)]


class OGem(C.Structure):
    _fields_ = [
        ("ns", C.c_int),
        ("B0x", C.c_double), ("B0y", C.c_double), ("B0z", C.c_double),
        ("delta", C.c_double), ("fourpi", C.c_double), ("c", C.c_double),
        ("rhoINIT", C.c_double * 16),
        ("qom", C.c_double * 16), ("u0", C.c_double * 16), ("v0", C.c_double * 16),
        ("w0", C.c_double * 16), ("uth", C.c_double * 16), ("vth", C.c_double * 16),
        ("wth", C.c_double * 16),
        ("npcelx", C.c_int * 16), ("npcely", C.c_int * 16), ("npcelz", C.c_int * 16),
        ("seed_base", C.c_uint), ("pertX", C.c_double),
    ]


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


_lib = None


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH}; run make -C oracle")
        L = C.CDLL(LIB_PATH)
        for sfx in ("f64", "f32"):
            getattr(L, f"oracle_set_grid_{sfx}").argtypes = [
                C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                C.POINTER(OGrid), _P, _P, _P]
            f = getattr(L, f"oracle_mover_pc_{sfx}")
            f.argtypes = [C.POINTER(OGrid), C.POINTER(ONodes), C.POINTER(OFields),
                          C.POINTER(OParticles), C.POINTER(OPush)]
            f.restype = C.c_int
            f = getattr(L, f"oracle_interp_p2g_{sfx}")
            f.argtypes = [C.POINTER(OGrid), C.POINTER(ONodes), C.POINTER(OParticles),
                          C.POINTER(OMoments)]
            f = getattr(L, f"oracle_mover_interp_{sfx}")
            f.argtypes = [C.POINTER(OGrid), C.POINTER(ONodes), C.POINTER(OFields),
                          C.POINTER(OParticles), C.POINTER(OMoments), C.POINTER(OPush)]
            f.restype = C.c_int
            f = getattr(L, f"oracle_particle_sort_{sfx}")
            f.argtypes = [C.POINTER(OGrid), C.c_long, C.POINTER(OParticles)]
            f.restype = C.c_int
            f = getattr(L, f"oracle_gem_fields_{sfx}")
            f.argtypes = [C.POINTER(OGrid), _P, _P, C.POINTER(OGem)] + [_P] * 6 + [
                C.POINTER(_P), C.POINTER(_P)]
            f = getattr(L, f"oracle_gem_particles_{sfx}")
            f.argtypes = [C.POINTER(OGrid), _P, _P, _P, C.POINTER(OGem), C.c_int, _P,
                          C.c_long] + [_P] * 7
            f.restype = C.c_long
            f = getattr(L, f"oracle_apply_bc_ids_{sfx}")
            f.argtypes = [C.POINTER(OGrid), C.POINTER(OMoments), C.c_int, C.c_int, C.c_int]
            f = getattr(L, f"oracle_sum_species_{sfx}")
            f.argtypes = [C.POINTER(OGrid), C.POINTER(OMoments), C.POINTER(OMoments), C.c_int]
            f = getattr(L, f"oracle_bc_scalar_{sfx}")
            f.argtypes = [C.POINTER(OGrid), _P, C.c_int, C.c_int, C.c_int, C.c_int]
            f = getattr(L, f"oracle_rhoc_from_rhon_{sfx}")
            f.argtypes = [C.POINTER(OGrid), _P, _P, C.c_int, C.c_int, C.c_int]
            f = getattr(L, f"oracle_hat_densities_{sfx}")
            f.argtypes = [C.POINTER(OGrid), C.POINTER(OMoments), C.c_int, _P, _P, _P, _P, _P,
                          C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                          _P, _P, _P, _P]
            f = getattr(L, f"oracle_maxwell_image_{sfx}")
            f.argtypes = [C.POINTER(OGrid), _P, _P, _P, C.POINTER(OMoments), C.c_int, _P,
                          C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                          C.c_int, _P, _P]
            f = getattr(L, f"oracle_zero_moments_{sfx}")
            f.argtypes = [C.POINTER(OGrid), C.POINTER(OMoments)]
            f = getattr(L, f"oracle_time_cycles_{sfx}")
            f.argtypes = [C.POINTER(OGrid), C.POINTER(ONodes), C.POINTER(OFields),
                          C.POINTER(OParticles), C.POINTER(OMoments), C.c_int,
                          C.POINTER(OPush), C.c_int, C.POINTER(C.c_double),
                          C.POINTER(C.c_double)]
            f.restype = C.c_double
            getattr(L, f"oracle_wtime_{sfx}").restype = C.c_double
            getattr(L, f"oracle_max_threads_{sfx}").restype = C.c_int
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# Problem description (the subset of struct parameters the path reads)
# --------------------------------------------------------------------------
@dataclass
class GemConfig:
    """Harris-sheet (GEM) problem; defaults are inputfiles/GEM_3D.inp with
    the src/RW_IO.cpp defaults for keys that file leaves out."""
    nxc: int = 32
    nyc: int = 16
    nzc: int = 16
    Lx: float = 20.0
    Ly: float = 10.0
    Lz: float = 10.0
    dt: float = 0.25
    c: float = 1.0
    NiterMover: int = 3
    n_sub_cycles: int = 1
    ns: int = 4
    npcelx: List[int] = dc_field(default_factory=lambda: [3, 3, 3, 3])
    npcely: List[int] = dc_field(default_factory=lambda: [3, 3, 3, 3])
    npcelz: List[int] = dc_field(default_factory=lambda: [3, 3, 3, 3])
    qom: List[float] = dc_field(default_factory=lambda: [-64.0, 1.0, -64.0, 1.0])
    rhoINIT: List[float] = dc_field(default_factory=lambda: [1.0, 1.0, 0.1, 0.1])
    uth: List[float] = dc_field(default_factory=lambda: [0.045, 0.0126, 0.045, 0.0126])
    vth: List[float] = dc_field(default_factory=lambda: [0.045, 0.0126, 0.045, 0.0126])
    wth: List[float] = dc_field(default_factory=lambda: [0.045, 0.0126, 0.045, 0.0126])
    u0: List[float] = dc_field(default_factory=lambda: [0.0, 0.0, 0.0, 0.0])
    v0: List[float] = dc_field(default_factory=lambda: [0.0, 0.0, 0.0, 0.0])
    w0: List[float] = dc_field(default_factory=lambda: [0.0065, -0.0325, 0.0, 0.0])
    B0x: float = 0.0195
    B0y: float = 0.0
    B0z: float = 0.0
    delta: float = 0.5
    fourpi: float = 12.5663706144
    pertX: float = 0.4
    th: float = 1.0
    PERIODICX: bool = True
    PERIODICY: bool = False
    PERIODICZ: bool = True
    sort_cps: int = 2
    seed_base: int = 0

    def np_species(self, i: int) -> int:
        return self.npcelx[i] * self.npcely[i] * self.npcelz[i] * self.nxc * self.nyc * self.nzc


class State:
    """Grid, fields and per-species particles and moments held in numpy
    arrays of the chosen precision ("f64" or "f32")."""

    def __init__(self, cfg: GemConfig, prec: str = "f64", max_per_species: int = 0,
                 build_ic: bool = True):
        assert prec in ("f64", "f32")
        self.cfg = cfg
        self.prec = prec
        self.dt = np.float64 if prec == "f64" else np.float32
        L = lib()
        self.grid = OGrid()
        g = self.grid
        getattr(L, f"oracle_set_grid_{prec}")(cfg.nxc, cfg.nyc, cfg.nzc, cfg.Lx, cfg.Ly,
                                             cfg.Lz, C.byref(g), None, None, None)
        self.nn = g.nxn * g.nyn * g.nzn
        self.ncell = g.nxc * g.nyc * g.nzc
        self.XN = np.zeros(self.nn, self.dt)
        self.YN = np.zeros(self.nn, self.dt)
        self.ZN = np.zeros(self.nn, self.dt)
        getattr(L, f"oracle_set_grid_{prec}")(cfg.nxc, cfg.nyc, cfg.nzc, cfg.Lx, cfg.Ly,
                                             cfg.Lz, C.byref(g), _ptr(self.XN),
                                             _ptr(self.YN), _ptr(self.ZN))
        self.fields: Dict[str, np.ndarray] = {n: np.zeros(self.nn, self.dt) for n in FIELD_NAMES}
        self.rhon_ic = [np.zeros(self.nn, self.dt) for _ in range(cfg.ns)]
        self.rhoc = [np.zeros(self.ncell, self.dt) for _ in range(cfg.ns)]
        self.parts: List[Dict[str, np.ndarray]] = []
        self.moments: List[Dict[str, np.ndarray]] = [
            {n: np.zeros(self.nn, self.dt) for n in MOMENT_NAMES} for _ in range(cfg.ns)]
        if build_ic:
            self._gem(max_per_species)

    # -- structures ---------------------------------------------------------
    def gem_struct(self) -> OGem:
        c = self.cfg
        p = OGem()
        p.ns = c.ns
        p.B0x, p.B0y, p.B0z, p.delta, p.fourpi, p.c = c.B0x, c.B0y, c.B0z, c.delta, c.fourpi, c.c
        for i in range(c.ns):
            p.rhoINIT[i] = c.rhoINIT[i]
            p.qom[i] = c.qom[i]
            p.u0[i], p.v0[i], p.w0[i] = c.u0[i], c.v0[i], c.w0[i]
            p.uth[i], p.vth[i], p.wth[i] = c.uth[i], c.vth[i], c.wth[i]
            p.npcelx[i], p.npcely[i], p.npcelz[i] = c.npcelx[i], c.npcely[i], c.npcelz[i]
        p.seed_base = c.seed_base
        p.pertX = c.pertX
        return p

    def nodes(self) -> ONodes:
        return ONodes(_ptr(self.XN), _ptr(self.YN), _ptr(self.ZN))

    def field_struct(self) -> OFields:
        return OFields(*[_ptr(self.fields[n]) for n in FIELD_NAMES])

    def part_struct(self, is_: int) -> OParticles:
        p = self.parts[is_]
        n = len(p["x"])
        return OParticles(*[_ptr(p[k]) for k in PARTICLE_NAMES], n, n, self.cfg.qom[is_],
                          self.cfg.NiterMover, self.cfg.n_sub_cycles)

    def mom_struct(self, is_: int) -> OMoments:
        return OMoments(*[_ptr(self.moments[is_][n]) for n in MOMENT_NAMES])

    def push_struct(self) -> OPush:
        c = self.cfg
        return OPush(c.dt, c.c, int(c.PERIODICX), int(c.PERIODICY), int(c.PERIODICZ))

    # -- IC -----------------------------------------------------------------
    def _gem(self, max_per_species: int):
        L = lib()
        gem = self.gem_struct()
        rhon_ptrs = (_P * self.cfg.ns)(*[_ptr(a) for a in self.rhon_ic])
        rhoc_ptrs = (_P * self.cfg.ns)(*[_ptr(a) for a in self.rhoc])
        f = self.fields
        getattr(L, f"oracle_gem_fields_{self.prec}")(
            C.byref(self.grid), _ptr(self.XN), _ptr(self.YN), C.byref(gem),
            *[_ptr(f[n]) for n in FIELD_NAMES], rhon_ptrs, rhoc_ptrs)
        for is_ in range(self.cfg.ns):
            n = self.cfg.np_species(is_)
            if max_per_species > 0:
                n = min(n, max_per_species)
            arrs = {k: np.zeros(n, self.dt) for k in PARTICLE_NAMES}
            got = getattr(L, f"oracle_gem_particles_{self.prec}")(
                C.byref(self.grid), _ptr(self.XN), _ptr(self.YN), _ptr(self.ZN),
                C.byref(gem), is_, _ptr(self.rhoc[is_]), n,
                *[_ptr(arrs[k]) for k in PARTICLE_NAMES])
            if got != n:
                raise RuntimeError(f"oracle IC for species {is_}: {got} != {n}")
            self.parts.append(arrs)

    # -- the path -----------------------------------------------------------
    def mover_pc(self, is_: int) -> int:
        ps = self.part_struct(is_)
        return getattr(lib(), f"oracle_mover_pc_{self.prec}")(
            C.byref(self.grid), C.byref(self.nodes()), C.byref(self.field_struct()),
            C.byref(ps), C.byref(self.push_struct()))

    def interp_p2g(self, is_: int) -> None:
        ps = self.part_struct(is_)
        ms = self.mom_struct(is_)
        getattr(lib(), f"oracle_interp_p2g_{self.prec}")(
            C.byref(self.grid), C.byref(self.nodes()), C.byref(ps), C.byref(ms))

    def mover_interp(self, is_: int) -> int:
        ps = self.part_struct(is_)
        ms = self.mom_struct(is_)
        return getattr(lib(), f"oracle_mover_interp_{self.prec}")(
            C.byref(self.grid), C.byref(self.nodes()), C.byref(self.field_struct()),
            C.byref(ps), C.byref(ms), C.byref(self.push_struct()))

    def particle_sort(self, is_: int, sort_cps: Optional[int] = None) -> int:
        ps = self.part_struct(is_)
        return getattr(lib(), f"oracle_particle_sort_{self.prec}")(
            C.byref(self.grid), int(sort_cps or self.cfg.sort_cps), C.byref(ps))

    def zero_moments(self, is_: int) -> None:
        for a in self.moments[is_].values():
            a[:] = 0

    def apply_bc_ids(self, is_: int) -> None:
        c = self.cfg
        ms = self.mom_struct(is_)
        getattr(lib(), f"oracle_apply_bc_ids_{self.prec}")(
            C.byref(self.grid), C.byref(ms), int(c.PERIODICX), int(c.PERIODICY),
            int(c.PERIODICZ))

    def sum_species(self, net: Optional[Dict[str, np.ndarray]] = None) -> Dict[str, np.ndarray]:
        """sumOverSpecies: net += every species' moments, in species order;
        net starts at zero unless given."""
        dt = self.moments[0]["rhon"].dtype
        if net is None:
            net = {n: np.zeros(self.nn, dt) for n in MOMENT_NAMES}
        net = {n: np.ascontiguousarray(net[n], dt).copy() for n in MOMENT_NAMES}
        ns_ = self.cfg.ns
        ids = (OMoments * ns_)(*[self.mom_struct(i) for i in range(ns_)])
        out = OMoments(*[_ptr(net[n]) for n in MOMENT_NAMES])
        getattr(lib(), f"oracle_sum_species_{self.prec}")(C.byref(self.grid), C.byref(out),
                                                           ids, ns_)
        return net

    # -- field-solve side ----------------------------------------------------
    def bc_scalar(self, a: np.ndarray, kind: int) -> np.ndarray:
        """Ghost update of a copy of `a` (kind 0 DensN, 1 FieldNzero, 2
        DensC, 3 FieldC)."""
        a = np.ascontiguousarray(a, self.dt).copy()
        c = self.cfg
        getattr(lib(), f"oracle_bc_scalar_{self.prec}")(
            C.byref(self.grid), _ptr(a), kind, int(c.PERIODICX), int(c.PERIODICY),
            int(c.PERIODICZ))
        return a

    def rhoc_from_rhon(self, rhon: np.ndarray):
        """(rhon after applyBCscalarDensN, rhoc = interpN2Cinterp(rhon)); the
        ghost cells of rhoc are zero."""
        rhon = np.ascontiguousarray(rhon, self.dt).copy()
        rhoc = np.zeros(self.ncell, self.dt)
        c = self.cfg
        getattr(lib(), f"oracle_rhoc_from_rhon_{self.prec}")(
            C.byref(self.grid), _ptr(rhoc), _ptr(rhon), int(c.PERIODICX), int(c.PERIODICY),
            int(c.PERIODICZ))
        return rhon, rhoc

    def hat_densities(self, rhoc: np.ndarray) -> Dict[str, np.ndarray]:
        """calculateHatDensities from this state's species moments and B and
        the given net rhoc: rhoh (cells), Jxh Jyh Jzh (nodes)."""
        c = self.cfg
        ns_ = c.ns
        ids = (OMoments * ns_)(*[self.mom_struct(i) for i in range(ns_)])
        qom = np.array(c.qom[:ns_], np.float64)
        rhoc = np.ascontiguousarray(rhoc, self.dt)
        out = {"rhoh": np.zeros(self.ncell, self.dt)}
        for n in ("Jxh", "Jyh", "Jzh"):
            out[n] = np.zeros(self.nn, self.dt)
        getattr(lib(), f"oracle_hat_densities_{self.prec}")(
            C.byref(self.grid), ids, ns_, _ptr(qom), _ptr(self.fields["Bxn"]),
            _ptr(self.fields["Byn"]), _ptr(self.fields["Bzn"]), _ptr(rhoc), c.dt, c.c, c.th,
            int(c.PERIODICX), int(c.PERIODICY), int(c.PERIODICZ), _ptr(out["rhoh"]),
            _ptr(out["Jxh"]), _ptr(out["Jyh"]), _ptr(out["Jzh"]))
        return out

    def maxwell_image(self, vec: np.ndarray) -> np.ndarray:
        """MaxwellImage of a Krylov vector (3 (nxn-2)(nyn-2)(nzn-2) values)
        with this state's species rhon and B."""
        c = self.cfg
        ns_ = c.ns
        ids = (OMoments * ns_)(*[self.mom_struct(i) for i in range(ns_)])
        qom = np.array(c.qom[:ns_], np.float64)
        vec = np.ascontiguousarray(vec, self.dt)
        im = np.zeros_like(vec)
        getattr(lib(), f"oracle_maxwell_image_{self.prec}")(
            C.byref(self.grid), _ptr(self.fields["Bxn"]), _ptr(self.fields["Byn"]),
            _ptr(self.fields["Bzn"]), ids, ns_, _ptr(qom), c.dt, c.c, c.th, c.fourpi,
            int(c.PERIODICX), int(c.PERIODICY), int(c.PERIODICZ), _ptr(vec), _ptr(im))
        return im

    def cycle(self) -> None:
        """One cycle of the hot path as src/sputniPIC_CPU.cpp:232-256 runs
        it: mover_PC for every species, zero the moments, interpP2G for every
        species (no BC fold: that step lies after the path)."""
        for is_ in range(self.cfg.ns):
            self.mover_pc(is_)
        for is_ in range(self.cfg.ns):
            self.zero_moments(is_)
            self.interp_p2g(is_)

    def time_cycles(self, ncycles: int = 1):
        """CPU-baseline timing: the reference driver's particle phase
        (src/sputniPIC_CPU.cpp:232-256) for all species, ncycles times.
        Returns (total, mover, deposit) seconds."""
        ns = self.cfg.ns
        parts = (OParticles * ns)(*[self.part_struct(s) for s in range(ns)])
        moms = (OMoments * ns)(*[self.mom_struct(s) for s in range(ns)])
        tm, td = C.c_double(), C.c_double()
        t = getattr(lib(), f"oracle_time_cycles_{self.prec}")(
            C.byref(self.grid), C.byref(self.nodes()), C.byref(self.field_struct()), parts,
            moms, ns, C.byref(self.push_struct()), ncycles, C.byref(tm), C.byref(td))
        return t, tm.value, td.value

    @staticmethod
    def max_threads() -> int:
        return lib().oracle_max_threads_f64()

    def copy_from(self, other: "State") -> None:
        """Take particles and fields from another state (cast to this
        state's precision)."""
        for n in FIELD_NAMES:
            self.fields[n][:] = other.fields[n]
        self.parts = [{k: v.astype(self.dt) for k, v in p.items()} for p in other.parts]


# --------------------------------------------------------------------------
# The reference's own translation units (this container only)
# --------------------------------------------------------------------------
class RefCfg(C.Structure):
    _fields_ = [
        ("nxc", C.c_int), ("nyc", C.c_int), ("nzc", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("dt", C.c_double), ("c", C.c_double),
        ("NiterMover", C.c_int), ("n_sub_cycles", C.c_int), ("ns", C.c_int),
        ("npcelx", C.c_int * 6), ("npcely", C.c_int * 6), ("npcelz", C.c_int * 6),
        ("qom", C.c_double * 6), ("rhoINIT", C.c_double * 6), ("uth", C.c_double * 6),
        ("vth", C.c_double * 6), ("wth", C.c_double * 6), ("u0", C.c_double * 6),
        ("v0", C.c_double * 6), ("w0", C.c_double * 6),
        ("B0x", C.c_double), ("B0y", C.c_double), ("B0z", C.c_double), ("delta", C.c_double),
        ("PERIODICX", C.c_int), ("PERIODICY", C.c_int), ("PERIODICZ", C.c_int),
        ("sort_cps", C.c_long),
    ]


def ref_available(prec: str = "f64") -> bool:
    """prec: "f64", "f32", or "f64_O3" (the -O3 timing build)."""
    return os.path.exists(os.path.join(REF_DIR, f"libref_{prec}.so"))


class RefState:
    """A reference simulation state built by the reference's own setGrid,
    allocators and initGEM (seed = species index, as src/IC.cpp:121-122)."""

    _libs: Dict[str, C.CDLL] = {}

    def __init__(self, cfg: GemConfig, prec: str = "f64", ic: bool = True,
                 io_dir: Optional[str] = None, restart: bool = False):
        """ic False: skip initGEM; fields and particles (cfg's counts) are
        then set with set_field / set_particles before use.  io_dir:
        initGEM's IC dump directory -- it writes its save_ic_data files
        there, or with restart reads read_ic_data's files from there
        instead of computing the IC."""
        if cfg.ns > 6:
            raise ValueError("the reference holds at most NS_MAX = 6 species")
        if ic and (cfg.seed_base != 0 or cfg.pertX != 0.4 or cfg.fourpi != 12.5663706144
                   or cfg.th != 1.0):
            raise ValueError("the reference IC is fixed to seed 0, pertX 0.4, default fourpi "
                             "and th")
        self.cfg, self.prec = cfg, prec
        L = self._load(prec)
        rc = RefCfg()
        for k in ("nxc", "nyc", "nzc", "Lx", "Ly", "Lz", "dt", "c", "NiterMover",
                  "n_sub_cycles", "ns", "B0x", "B0y", "B0z", "delta", "sort_cps"):
            setattr(rc, k, getattr(cfg, k))
        rc.PERIODICX, rc.PERIODICY, rc.PERIODICZ = (int(cfg.PERIODICX), int(cfg.PERIODICY),
                                                    int(cfg.PERIODICZ))
        for i in range(cfg.ns):
            for k in ("npcelx", "npcely", "npcelz", "qom", "rhoINIT", "uth", "vth", "wth",
                      "u0", "v0", "w0"):
                getattr(rc, k)[i] = getattr(cfg, k)[i]
        self.L = L
        if io_dir is not None:
            self.h = L.ref_create_io(C.byref(rc), io_dir.encode(), int(bool(restart)))
        else:
            self.h = (L.ref_create if ic else L.ref_create_empty)(C.byref(rc))
        ints = (C.c_int * 6)()
        dbl = (C.c_double * 13)()
        L.ref_grid(self.h, ints, dbl)
        self.nxn, self.nyn, self.nzn = ints[3], ints[4], ints[5]
        self.nn = self.nxn * self.nyn * self.nzn
        self.ncell = ints[0] * ints[1] * ints[2]
        self.grid_ints = list(ints)
        self.grid_dbl = list(dbl)

    def time_cycles(self, ncycles: int = 1):
        """The reference driver's particle phase timed with omp_get_wtime
        (ref_time_cycles); returns (total, mover, deposit) seconds."""
        tm, td = C.c_double(), C.c_double()
        t = self.L.ref_time_cycles(self.h, ncycles, C.byref(tm), C.byref(td))
        return t, tm.value, td.value

    def max_threads(self) -> int:
        return self.L.ref_max_threads()

    @classmethod
    def _load(cls, prec):
        if prec not in cls._libs:
            path = os.path.join(REF_DIR, f"libref_{prec}.so")
            if not os.path.exists(path):
                raise RuntimeError(f"reference build missing: {path}")
            L = C.CDLL(path)
            for f in ("ref_create", "ref_create_empty"):
                getattr(L, f).restype = _P
                getattr(L, f).argtypes = [C.POINTER(RefCfg)]
            L.ref_create_io.restype = _P
            L.ref_create_io.argtypes = [C.POINTER(RefCfg), C.c_char_p, C.c_int]
            L.ref_get_aux.argtypes = [_P, C.c_int, _P]
            L.ref_destroy.argtypes = [_P]
            L.ref_grid.argtypes = [_P, _P, _P]
            L.ref_nop.restype = C.c_long
            L.ref_nop.argtypes = [_P, C.c_int]
            for f in ("ref_get_particles", "ref_set_particles", "ref_get_moment",
                      "ref_set_moment"):
                getattr(L, f).argtypes = [_P, C.c_int, C.c_int, _P]
            for f in ("ref_get_field", "ref_set_field"):
                getattr(L, f).argtypes = [_P, C.c_int, _P]
            for f in ("ref_mover_pc", "ref_mover_interp"):
                getattr(L, f).argtypes = [_P, C.c_int]
                getattr(L, f).restype = C.c_int
            for f in ("ref_interp_p2g", "ref_particle_sort", "ref_apply_bc_ids"):
                getattr(L, f).argtypes = [_P, C.c_int]
            L.ref_zero_species_moments.argtypes = [_P]
            L.ref_sum_species.argtypes = [_P, C.POINTER(C.c_double)]
            L.ref_hat_densities.argtypes = [_P, _P]
            L.ref_time_field.argtypes = [_P, C.c_int, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]
            L.ref_maxwell_image.argtypes = [_P, _P, _P]
            L.ref_time_cycles.argtypes = [_P, C.c_int, C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]
            L.ref_time_cycles.restype = C.c_double
            L.ref_max_threads.restype = C.c_int
            cls._libs[prec] = L
        return cls._libs[prec]

    def close(self):
        if self.h:
            self.L.ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nop(self, is_):
        return self.L.ref_nop(self.h, is_)

    def particles(self, is_) -> Dict[str, np.ndarray]:
        n = self.nop(is_)
        out = {}
        for w, k in enumerate(PARTICLE_NAMES):
            a = np.zeros(n, np.float64)
            self.L.ref_get_particles(self.h, is_, w, _ptr(a))
            out[k] = a
        return out

    def set_particles(self, is_, parts: Dict[str, np.ndarray]):
        for w, k in enumerate(PARTICLE_NAMES):
            a = np.ascontiguousarray(parts[k], dtype=np.float64)
            assert len(a) == self.nop(is_)
            self.L.ref_set_particles(self.h, is_, w, _ptr(a))

    def field(self, which: int) -> np.ndarray:
        a = np.zeros(self.nn, np.float64)
        self.L.ref_get_field(self.h, which, _ptr(a))
        return a

    def aux(self, which: int) -> np.ndarray:
        """0..2 Bxc Byc Bzc (cells), 3..5 Exth Eyth Ezth (nodes)."""
        a = np.zeros(self.ncell if which < 3 else self.nn, np.float64)
        self.L.ref_get_aux(self.h, which, _ptr(a))
        return a

    def set_field(self, which: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self.L.ref_set_field(self.h, which, _ptr(a))

    def moment(self, is_, which: int) -> np.ndarray:
        a = np.zeros(self.ncell if which == 10 else self.nn, np.float64)
        self.L.ref_get_moment(self.h, is_, which, _ptr(a))
        return a

    def set_moment(self, is_, which: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self.L.ref_set_moment(self.h, is_, which, _ptr(a))

    def zero_moments(self):
        self.L.ref_zero_species_moments(self.h)

    def mover_pc(self, is_):
        return self.L.ref_mover_pc(self.h, is_)

    def interp_p2g(self, is_):
        self.L.ref_interp_p2g(self.h, is_)

    def mover_interp(self, is_):
        return self.L.ref_mover_interp(self.h, is_)

    def particle_sort(self, is_):
        self.L.ref_particle_sort(self.h, is_)

    def apply_bc_ids(self, is_):
        self.L.ref_apply_bc_ids(self.h, is_)

    def sum_species(self) -> Dict[str, np.ndarray]:
        """sumOverSpecies into a zeroed net; the ten node moments (float64)."""
        nn = self.nn
        out = np.zeros(10 * nn, np.float64)
        self.L.ref_sum_species(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return {n: out[i * nn:(i + 1) * nn].copy() for i, n in enumerate(MOMENT_NAMES)}

    def hat_densities(self) -> Dict[str, np.ndarray]:
        """The reference's sumOverSpecies, applyBCscalarDensN(net rhon),
        interpN2Cinterp(net rhoc) and calculateHatDensities: rhon (after the
        BC), rhoc, rhoh, Jxh, Jyh, Jzh (float64)."""
        nn, nc = self.nn, self.ncell
        out = np.zeros(4 * nn + 2 * nc, np.float64)
        self.L.ref_hat_densities(self.h, _ptr(out))
        sizes = [("rhon", nn), ("rhoc", nc), ("rhoh", nc), ("Jxh", nn), ("Jyh", nn),
                 ("Jzh", nn)]
        res, o = {}, 0
        for name, n in sizes:
            res[name] = out[o:o + n].copy()
            o += n
        return res

    def maxwell_image(self, vec: np.ndarray) -> np.ndarray:
        vec = np.ascontiguousarray(vec, np.float64)
        im = np.zeros_like(vec)
        self.L.ref_maxwell_image(self.h, _ptr(vec), _ptr(im))
        return im

    def time_field(self, reps: int = 1):
        """Seconds for `reps` hat-density steps and `reps` MaxwellImage
        calls (ref_time_field)."""
        th, ti = C.c_double(), C.c_double()
        self.L.ref_time_field(self.h, reps, C.byref(th), C.byref(ti))
        return th.value, ti.value
