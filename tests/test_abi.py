"""The C ABI of libspic.so, checked without a GPU: every function that
include/spic.h declares is exported, the ctypes mirror of each struct has
the layout a C compiler gives the header, and argument validation returns
the documented status codes before any device call is made."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from sputnipic_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spic.h")


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(spic_\w+)\s*\(", text,
                                 flags=re.M)))


def test_library_exports_every_declared_function():
    L = _capi.lib()
    names = declared_functions(HEADER)
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_capi.SIGNATURES), set(names) ^ set(_capi.SIGNATURES)
    assert L.spic_abi_version() == 1


STRUCTS = {
    "spic_grid": _capi.spic_grid,
    "spic_field": _capi.spic_field,
    "spic_species": _capi.spic_species,
    "spic_moments": _capi.spic_moments,
    "spic_push": _capi.spic_push,
    "spic_gem_params": _capi.spic_gem_params,
    "spic_gem_species": _capi.spic_gem_species,
    "spic_engine_opts": _capi.spic_engine_opts,
    "spic_engine_stats": _capi.spic_engine_stats,
}


def test_ctypes_layout_matches_c_header():
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "spic.h"',
             "int main(void) {"]
    for name, cls in STRUCTS.items():
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{name} {fname} %zu\\n", offsetof({name}, {fname}));')
    lines += ["  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), src, "-o",
                        exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        s, f, v = line.split()
        got[(s, f)] = int(v)
    for name, cls in STRUCTS.items():
        assert got[(name, "size")] == C.sizeof(cls), name
        for fname, _ in cls._fields_:
            assert got[(name, fname)] == getattr(cls, fname).offset, (name, fname)


def _fake_args(prec=_capi.SPIC_F64, n_sub_cycles=1, niter=3):
    fake = 0x1000  # never dereferenced: validation fails first
    g = _capi.spic_grid(34, 18, 18, 35, 19, 19, 0.625, 0.625, 0.625, 1.6, 1.6, 1.6, 4.096,
                        0, 0, 0, 20, 10, 10, prec, fake, fake, fake)
    f = _capi.spic_field(fake, fake, fake, fake, fake, fake)
    s = _capi.spic_species(fake, fake, fake, fake, fake, fake, fake, 10, 10, -64.0, niter,
                           n_sub_cycles)
    p = _capi.spic_push(0.25, 1.0, 1, 0, 1)
    return g, f, s, p


def test_null_arguments_are_rejected():
    L = _capi.lib()
    g, f, s, p = _fake_args()
    assert L.spic_mover_pc(None, C.byref(f), C.byref(s), C.byref(p), None) == _capi.SPIC_EINVAL
    assert b"grid" in L.spic_last_error()
    assert L.spic_mover_pc(C.byref(g), None, C.byref(s), C.byref(p), None) == _capi.SPIC_EINVAL
    assert L.spic_interp_p2g(C.byref(g), C.byref(s), None, None) == _capi.SPIC_EINVAL
    assert L.spic_particle_sort(C.byref(g), 0, C.byref(s), None) == _capi.SPIC_EINVAL


def test_sub_cycling_is_refused():
    """The CPU mover has no sub-cycle loop (src/Mover.cpp:11) while the GPU
    reference has one (src/gpu/Particles_gpu.cu:336): by default
    n_sub_cycles != 1 is refused rather than silently following either."""
    L = _capi.lib()
    g, f, s, p = _fake_args(n_sub_cycles=2)
    assert L.spic_mover_pc(C.byref(g), C.byref(f), C.byref(s), C.byref(p),
                           None) == _capi.SPIC_EUNSUPPORTED
    assert b"n_sub_cycles" in L.spic_last_error()


def test_sub_cycling_modes_pass_the_species_check():
    """With a sub-cycling mode chosen, a species with n_sub_cycles = 2 passes
    the species check: the call stops at the next check (a NULL push
    descriptor, so nothing reaches the device); n_sub_cycles = 0 and an
    unknown mode are refused."""
    L = _capi.lib()
    g, f, s, p = _fake_args(n_sub_cycles=2)
    try:
        assert L.spic_mover_pc(C.byref(g), C.byref(f), C.byref(s), None,
                               None) == _capi.SPIC_EUNSUPPORTED
        for mode in (_capi.SPIC_SUBCYCLE_LOOP, _capi.SPIC_SUBCYCLE_SINGLE):
            assert L.spic_set_subcycling(mode) == _capi.SPIC_OK
            assert L.spic_mover_pc(C.byref(g), C.byref(f), C.byref(s), None,
                                   None) == _capi.SPIC_EINVAL
            assert b"n_sub_cycles" not in L.spic_last_error()
        g, f, s, p = _fake_args(n_sub_cycles=0)
        assert L.spic_mover_pc(C.byref(g), C.byref(f), C.byref(s), None,
                               None) == _capi.SPIC_EINVAL
        assert b"n_sub_cycles" in L.spic_last_error()
        assert L.spic_set_subcycling(7) == _capi.SPIC_EINVAL
    finally:
        L.spic_set_subcycling(_capi.SPIC_SUBCYCLE_REFUSE)


@pytest.mark.parametrize("field,value", [("prec", 7), ("nxn", 10), ("Lx", -1.0)])
def test_bad_grid_is_rejected(field, value):
    L = _capi.lib()
    g, f, s, p = _fake_args()
    setattr(g, field, value)
    assert L.spic_mover_pc(C.byref(g), C.byref(f), C.byref(s), C.byref(p),
                           None) == _capi.SPIC_EINVAL


def test_bad_species_is_rejected():
    L = _capi.lib()
    g, f, s, p = _fake_args()
    s.nop = s.npmax + 1
    assert L.spic_mover_pc(C.byref(g), C.byref(f), C.byref(s), C.byref(p),
                           None) == _capi.SPIC_EINVAL
    g, f, s, p = _fake_args(niter=0)
    assert L.spic_mover_pc(C.byref(g), C.byref(f), C.byref(s), C.byref(p),
                           None) == _capi.SPIC_EINVAL


def test_missing_library_fails_loudly(monkeypatch):
    """Without libspic.so the package raises instead of computing anything
    another way: there is no fallback path."""
    import sputnipic_amd as S
    monkeypatch.setattr(_capi, "_lib", None)
    monkeypatch.setattr(_capi, "LIB_PATH", os.path.join(ROOT, "no-such-dir", "libspic.so"))
    with pytest.raises(RuntimeError, match="libspic.so not built"):
        _capi.lib()
    with pytest.raises(RuntimeError, match="libspic.so not built"):
        S.device_count()


def test_engine_particles_per_thread_is_validated_before_any_device_work():
    """particles_per_thread: 0, 1 or 2, and 2 only for FP32 engines with
    fixed-point positions; both refusals come before any device call (the
    grid's node tables are fake pointers here)."""
    L = _capi.lib()
    g, f, s, p = _fake_args()
    eng = C.c_void_p()
    o = _capi.spic_engine_opts()
    o.particles_per_thread = 3
    assert L.spic_engine_create(C.byref(eng), C.byref(g), 1, C.byref(o)) == _capi.SPIC_EINVAL
    assert b"particles_per_thread" in L.spic_last_error()
    o.particles_per_thread = 2  # an FP64 grid stores float positions
    assert L.spic_engine_create(C.byref(eng), C.byref(g), 1,
                                C.byref(o)) == _capi.SPIC_EUNSUPPORTED
    assert b"fixed-point" in L.spic_last_error()
    assert not eng.value
