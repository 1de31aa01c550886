"""ctypes binding of libspic.so (the C ABI declared in include/spic.h).

The library is built in-tree (``sputnipic_amd/lib/libspic.so``) by
``__graft_entry__.build()`` / ``make -C sputnipic_amd/csrc``.  There is no
fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libspic.so")

SPIC_OK = 0
SPIC_EINVAL = 1
SPIC_EHIP = 2
SPIC_EDOMAIN = 3
SPIC_EUNSUPPORTED = 4
SPIC_ENOMEM = 5
SPIC_F64 = 0
SPIC_F32 = 1
SPIC_SORT_TILE = 0
SPIC_SORT_RADIX = 1
SPIC_POS_FLOAT = 1
SPIC_POS_FIXED = 2
SPIC_ORDER_LOAD = 1
SPIC_ORDER_ENGINE = 2
SPIC_SUBCYCLE_REFUSE = 0
SPIC_SUBCYCLE_LOOP = 1
SPIC_SUBCYCLE_SINGLE = 2

STATUS_NAMES = {
    SPIC_OK: "SPIC_OK", SPIC_EINVAL: "SPIC_EINVAL", SPIC_EHIP: "SPIC_EHIP",
    SPIC_EDOMAIN: "SPIC_EDOMAIN", SPIC_EUNSUPPORTED: "SPIC_EUNSUPPORTED",
    SPIC_ENOMEM: "SPIC_ENOMEM",
}

_P = C.c_void_p


class spic_grid(C.Structure):
    _fields_ = [
        ("nxc", C.c_int), ("nyc", C.c_int), ("nzc", C.c_int),
        ("nxn", C.c_int), ("nyn", C.c_int), ("nzn", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("invdx", C.c_double), ("invdy", C.c_double), ("invdz", C.c_double),
        ("invVOL", C.c_double),
        ("xStart", C.c_double), ("yStart", C.c_double), ("zStart", C.c_double),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("prec", C.c_int),
        ("xn", _P), ("yn", _P), ("zn", _P),
    ]


class spic_field(C.Structure):
    _fields_ = [("Ex", _P), ("Ey", _P), ("Ez", _P), ("Bxn", _P), ("Byn", _P), ("Bzn", _P)]


class spic_species(C.Structure):
    _fields_ = [
        ("x", _P), ("y", _P), ("z", _P), ("u", _P), ("v", _P), ("w", _P), ("q", _P),
        ("nop", C.c_int64), ("npmax", C.c_int64), ("qom", C.c_double),
        ("NiterMover", C.c_int), ("n_sub_cycles", C.c_int),
    ]


class spic_moments(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in
                ("rhon", "Jx", "Jy", "Jz", "pxx", "pxy", "pxz", "pyy", "pyz", "pzz")]


class spic_push(C.Structure):
    _fields_ = [("dt", C.c_double), ("c", C.c_double),
                ("PERIODICX", C.c_int), ("PERIODICY", C.c_int), ("PERIODICZ", C.c_int)]


class spic_solve_params(C.Structure):
    _fields_ = [("dt", C.c_double), ("c", C.c_double), ("th", C.c_double),
                ("fourpi", C.c_double), ("PERIODICX", C.c_int), ("PERIODICY", C.c_int),
                ("PERIODICZ", C.c_int), ("qom", C.c_double * 8)]


class spic_hat(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("rhoh", "Jxh", "Jyh", "Jzh")]


class spic_gem_params(C.Structure):
    _fields_ = [("B0x", C.c_double), ("B0y", C.c_double), ("B0z", C.c_double),
                ("delta", C.c_double), ("pertX", C.c_double), ("fourpi", C.c_double)]


class spic_gem_species(C.Structure):
    _fields_ = [("species", C.c_int), ("seed", C.c_uint32), ("sheet", C.c_int),
                ("npcelx", C.c_int), ("npcely", C.c_int), ("npcelz", C.c_int),
                ("qom", C.c_double), ("rhoINIT", C.c_double),
                ("u0", C.c_double), ("v0", C.c_double), ("w0", C.c_double),
                ("uth", C.c_double), ("vth", C.c_double), ("wth", C.c_double)]


class spic_engine_opts(C.Structure):
    _fields_ = [("tile_x", C.c_int), ("tile_y", C.c_int), ("tile_z", C.c_int),
                ("sort_every", C.c_int), ("lds_partials", C.c_int),
                ("slack_pct", C.c_int), ("set_moments", C.c_int), ("sort_mode", C.c_int),
                ("positions", C.c_int), ("order", C.c_int), ("overflow_pct", C.c_int),
                ("particles_per_thread", C.c_int)]


class spic_engine_stats(C.Structure):
    _fields_ = [("phases", C.c_int64), ("sorts", C.c_int64), ("ntiles", C.c_int),
                ("tile_x", C.c_int), ("tile_y", C.c_int), ("tile_z", C.c_int),
                ("device_bytes", C.c_size_t),
                ("lds_gathers", C.c_int64), ("global_gathers", C.c_int64),
                ("lds_deposits", C.c_int64), ("global_deposits", C.c_int64),
                ("overflow", C.c_int64), ("positions", C.c_int),
                ("workspace_bytes", C.c_size_t)]


class spic_batch_member(C.Structure):
    _fields_ = [("engine", _P), ("field", C.POINTER(spic_field)),
                ("moments", C.POINTER(spic_moments)), ("push", C.POINTER(spic_push))]


# name -> (restype, argtypes); every entry point of include/spic.h
SIGNATURES = {
    "spic_mover_pc": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_field),
                                C.POINTER(spic_species), C.POINTER(spic_push), _P]),
    "spic_interp_p2g": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_species),
                                  C.POINTER(spic_moments), _P]),
    "spic_mover_interp": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_field),
                                    C.POINTER(spic_species), C.POINTER(spic_moments),
                                    C.POINTER(spic_push), _P]),
    "spic_particle_sort": (C.c_int, [C.POINTER(spic_grid), C.c_int64,
                                     C.POINTER(spic_species), _P]),
    "spic_particle_sort_noalloc": (C.c_int, [C.POINTER(spic_grid), C.c_int64,
                                             C.POINTER(spic_species), _P]),
    "spic_zero_moments": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_moments), _P]),
    "spic_apply_bc_ids": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_moments), C.c_int,
                                    C.POINTER(spic_push), _P]),
    "spic_sum_species": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_moments),
                                   C.POINTER(spic_moments), C.c_int, _P]),
    "spic_field_workspace_bytes": (C.c_size_t, [C.POINTER(spic_grid), C.c_int]),
    "spic_bc_scalar": (C.c_int, [C.POINTER(spic_grid), _P, C.c_int,
                                 C.POINTER(spic_solve_params), _P]),
    "spic_rhoc_from_rhon": (C.c_int, [C.POINTER(spic_grid), _P, _P,
                                      C.POINTER(spic_solve_params), _P]),
    "spic_hat_densities": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_field),
                                     C.POINTER(spic_moments), C.c_int, _P,
                                     C.POINTER(spic_solve_params), C.POINTER(spic_hat), _P,
                                     C.c_size_t, _P]),
    "spic_maxwell_image": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_field),
                                     C.POINTER(spic_moments), C.c_int,
                                     C.POINTER(spic_solve_params), _P, _P, _P, C.c_size_t,
                                     _P]),
    "spic_engine_create": (C.c_int, [C.POINTER(_P), C.POINTER(spic_grid), C.c_int,
                                     C.POINTER(spic_engine_opts)]),
    "spic_engine_destroy": (C.c_int, [_P]),
    "spic_engine_load": (C.c_int, [_P, C.c_int, C.POINTER(spic_species), _P]),
    "spic_engine_store": (C.c_int, [_P, C.c_int, C.POINTER(spic_species), _P]),
    "spic_engine_load_as": (C.c_int, [_P, C.c_int, C.POINTER(spic_species), C.c_int, _P]),
    "spic_engine_store_as": (C.c_int, [_P, C.c_int, C.POINTER(spic_species), C.c_int, _P]),
    "spic_engine_deposit": (C.c_int, [_P, C.POINTER(spic_moments), _P]),
    "spic_engine_phase": (C.c_int, [_P, C.POINTER(spic_field), C.POINTER(spic_moments),
                                    C.POINTER(spic_push), _P]),
    "spic_engine_sort": (C.c_int, [_P, _P]),
    "spic_engine_snapshot_be32": (C.c_int, [_P, C.c_int, C.POINTER(_P), _P]),
    "spic_snapshot_be32": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_species),
                                     C.POINTER(_P), _P]),
    "spic_permute_species": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_species),
                                       C.POINTER(spic_species), C.c_uint64, C.c_uint64, _P]),
    "spic_host_alloc": (C.c_int, [C.POINTER(_P), C.c_size_t]),
    "spic_host_free": (C.c_int, [_P]),
    "spic_memcpy_d2h_async": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "spic_stream_wait_event": (C.c_int, [_P, _P]),
    "spic_engine_get_stats": (C.c_int, [_P, C.POINTER(spic_engine_stats)]),
    "spic_batch_create": (C.c_int, [C.POINTER(_P), C.POINTER(spic_batch_member), C.c_int,
                                    C.c_int, C.c_int]),
    "spic_batch_phase": (C.c_int, [_P, _P]),
    "spic_batch_destroy": (C.c_int, [_P]),
    "spic_gem_fields": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_gem_params),
                                  C.POINTER(spic_field), _P]),
    "spic_gem_particles": (C.c_int, [C.POINTER(spic_grid), C.POINTER(spic_gem_params),
                                     C.POINTER(spic_gem_species), C.POINTER(spic_species),
                                     _P]),
    "spic_event_create": (C.c_int, [C.POINTER(_P)]),
    "spic_event_destroy": (C.c_int, [_P]),
    "spic_event_record": (C.c_int, [_P, _P]),
    "spic_event_elapsed_ms": (C.c_int, [C.POINTER(C.c_float), _P, _P]),
    "spic_graph_begin": (C.c_int, [_P]),
    "spic_graph_end": (C.c_int, [_P, C.POINTER(_P)]),
    "spic_graph_launch": (C.c_int, [_P, _P]),
    "spic_graph_destroy": (C.c_int, [_P]),
    "spic_last_error": (C.c_char_p, []),
    "spic_abi_version": (C.c_int, []),
    "spic_set_async": (C.c_int, [C.c_int]),
    "spic_set_strict_fp": (C.c_int, [C.c_int]),
    "spic_set_subcycling": (C.c_int, [C.c_int]),
    "spic_check_domain": (C.c_int, [_P]),
    "spic_test_fail_engine_alloc": (C.c_int, [C.c_int]),
    "spic_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "spic_malloc": (C.c_int, [C.POINTER(_P), C.c_size_t]),
    "spic_free": (C.c_int, [_P]),
    "spic_memcpy_h2d": (C.c_int, [_P, _P, C.c_size_t]),
    "spic_memcpy_d2h": (C.c_int, [_P, _P, C.c_size_t]),
    "spic_memcpy_d2d": (C.c_int, [_P, _P, C.c_size_t]),
    "spic_memset": (C.c_int, [_P, C.c_int, C.c_size_t]),
    "spic_synchronize": (C.c_int, []),
    "spic_stream_create": (C.c_int, [C.POINTER(_P)]),
    "spic_stream_destroy": (C.c_int, [_P]),
    "spic_stream_synchronize": (C.c_int, [_P]),
}

_lib = None


class SpicError(RuntimeError):
    """A nonzero status from libspic; ``.code`` is the SPIC_* value."""

    def __init__(self, code: int, what: str, msg: str):
        self.code = code
        super().__init__(f"{what}: {STATUS_NAMES.get(code, code)}: {msg}")


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libspic.so not built ({LIB_PATH}); run __graft_entry__.build() "
                "or make -C sputnipic_amd/csrc")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != SPIC_OK:
        raise SpicError(rc, what, lib().spic_last_error().decode(errors="replace"))
