"""The C++ drop-in (include/sputnipic/sputnipic.h, libspic_dropin_*.so),
checked without a GPU:

* each library exports the reference's four particle entry points and the
  field-solve side's calculateHatDensities and MaxwellImage under their C++
  mangled names, so a reference driver links against it in place of
  src/Mover.cpp, src/Particles.cpp and those two routines;
* the seven structs have the reference's layout (offset and size of every
  member), compared with a program compiled against the reference's own
  headers where that checkout exists.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "sputnipic_amd", "lib")
REF_INCLUDE = "/root/reference/include"

MANGLED = {
    "mover_PC": "_Z8mover_PCP9particlesP7EMfieldP4gridP10parameters",
    "interpP2G": "_Z9interpP2GP9particlesP17interpDensSpeciesP4grid",
    "mover_interp": "_Z12mover_interpP9particlesP7EMfieldP17interpDensSpeciesP4gridP10parameters",
    "particle_sort": "_Z13particle_sortP10parametersP9particlesP4grid",
}


# the field-solve entry points (include/MaxwellSolver.h:12, 37); MaxwellImage
# takes FPfield pointers, so its name depends on the precision
MANGLED_FIELD = {
    "f32": {"calculateHatDensities": "_Z21calculateHatDensitiesP14interpDens_auxP13interpDensNetP17interpDensSpeciesP7EMfieldP4gridP10parameters",
            "MaxwellImage": "_Z12MaxwellImagePfS_P7EMfieldP17interpDensSpeciesP4gridP10parameters"},
    "f64": {"calculateHatDensities": "_Z21calculateHatDensitiesP14interpDens_auxP13interpDensNetP17interpDensSpeciesP7EMfieldP4gridP10parameters",
            "MaxwellImage": "_Z12MaxwellImagePdS_P7EMfieldP17interpDensSpeciesP4gridP10parameters"},
}


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_dropin_exports_reference_symbols(prec):
    path = os.path.join(LIB, f"libspic_dropin_{prec}.so")
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True,
                         capture_output=True, text=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name, mangled in list(MANGLED.items()) + list(MANGLED_FIELD[prec].items()):
        assert mangled in syms, name
    for ext in ("spic_dropin_set_resident", "spic_dropin_sync_to_host",
                "spic_dropin_host_modified", "spic_dropin_last_status",
                "spic_dropin_release_all"):
        assert ext in syms, ext


FIELDS = {
    "particles": ["species_ID", "npmax", "nop", "NiterMover", "n_sub_cycles", "npcel",
                  "npcelx", "npcely", "npcelz", "qom", "u0", "v0", "w0", "uth", "vth",
                  "wth", "x", "y", "z", "u", "v", "w", "q", "track_particle"],
    "EMfield": ["Ex", "Ex_flat", "Ey", "Ey_flat", "Ez", "Ez_flat", "Bxn", "Bxn_flat",
                "Byn", "Byn_flat", "Bzn", "Bzn_flat"],
    "grid": ["nxc", "nxn", "nyc", "nyn", "nzc", "nzn", "dx", "dy", "dz", "invdx",
             "invdy", "invdz", "invVOL", "xStart", "xEnd", "yStart", "yEnd", "zStart",
             "zEnd", "Lx", "Ly", "Lz", "PERIODICX", "PERIODICY", "PERIODICZ", "XN_flat",
             "XN", "YN_flat", "YN", "ZN_flat", "ZN"],
    "interpDensSpecies": ["species_ID", "rhon", "rhon_flat", "rhoc", "rhoc_flat", "Jx",
                          "Jx_flat", "Jy", "Jy_flat", "Jz", "Jz_flat", "pxx", "pxx_flat",
                          "pxy", "pxy_flat", "pxz", "pxz_flat", "pyy", "pyy_flat", "pyz",
                          "pyz_flat", "pzz", "pzz_flat"],
    "interpDensNet": ["rhon", "rhon_flat", "rhoc", "rhoc_flat", "Jx", "Jx_flat", "Jy",
                      "Jy_flat", "Jz", "Jz_flat", "pxx", "pxx_flat", "pxy", "pxy_flat",
                      "pxz", "pxz_flat", "pyy", "pyy_flat", "pyz", "pyz_flat", "pzz",
                      "pzz_flat"],
    "interpDens_aux": ["rhoh", "Jxh", "Jyh", "Jzh"],
    "parameters": ["c", "fourpi", "dt", "th", "ncycles", "NiterMover", "n_sub_cycles",
                   "sort", "sort_every_n", "sort_cps", "number_of_batches",
                   "threads_per_block", "track_particles", "n_tracked_particles",
                   "tracking_start_cycle", "tracking_end_cycle", "tracking_Lx",
                   "tracking_Ly", "tracking_Lz", "tracked_particles_filename", "Lx", "Ly",
                   "Lz", "nxc", "nyc", "nzc", "x_center", "y_center", "z_center",
                   "L_square", "ns", "npcelx", "npcely", "npcelz", "np", "npMax",
                   "NpMaxNpRatio", "qom", "rhoINIT", "uth", "vth", "wth", "u0", "v0", "w0",
                   "PERIODICX", "PERIODICY", "PERIODICZ", "PERIODICX_P", "PERIODICY_P",
                   "PERIODICZ_P", "bcPfaceXright", "bcPfaceXleft", "bcPfaceYright",
                   "bcPfaceYleft", "bcPfaceZright", "bcPfaceZleft", "bcPHIfaceXright",
                   "bcPHIfaceXleft", "bcPHIfaceYright", "bcPHIfaceYleft", "bcPHIfaceZright",
                   "bcPHIfaceZleft", "bcEMfaceXright", "bcEMfaceXleft", "bcEMfaceYright",
                   "bcEMfaceYleft", "bcEMfaceZright", "bcEMfaceZleft", "Vinj", "delta",
                   "B0x", "B0y", "B0z", "Nwaves", "dBoB0", "pitch_angle", "energy",
                   "SmoothON", "SmoothValue", "SmoothTimes", "verbose", "RESTART",
                   "PoissonCorrection", "CGtol", "GMREStol", "first_cycle_n",
                   "FieldOutputCycle", "ParticlesOutputCycle", "RestartOutputCycle",
                   "DiagnosticsOutputCycle", "inputfile", "SaveDirName", "RestartDirName",
                   "WaveFile"],
}


def layout(include_lines, flags):
    body = ["#include <cstdio>", "#include <cstddef>"] + include_lines + ["int main() {"]
    for s, fields in FIELDS.items():
        body.append(f'  std::printf("{s} size %zu\\n", sizeof(struct {s}));')
        for f in fields:
            body.append(f'  std::printf("{s} {f} %zu\\n", offsetof(struct {s}, {f}));')
    body += ["  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "l.cpp")
        exe = os.path.join(d, "l")
        open(src, "w").write("\n".join(body))
        subprocess.run(["g++", "-std=c++11", "-Wno-invalid-offsetof"] + flags +
                       [src, "-o", exe], check=True)
        return subprocess.run([exe], check=True, capture_output=True, text=True).stdout


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_struct_layout_matches_reference_headers(prec):
    if not os.path.isdir(REF_INCLUDE):
        pytest.skip("reference headers not present (build container only)")
    ours = layout(['#include "sputnipic/sputnipic.h"'],
                  ["-I", os.path.join(ROOT, "include")] +
                  (["-DSPIC_REF_FP64"] if prec == "f64" else []))
    ref_flags = ["-I", REF_INCLUDE]
    if prec == "f64":
        ref_flags += ["-include", os.path.join(ROOT, "oracle", "fp64_precision.h")]
    ref = layout(['#include "Particles.h"', '#include "EMfield.h"', '#include "Grid.h"',
                  '#include "InterpDensSpecies.h"', '#include "Parameters.h"',
                  '#include "InterpDensNet.h"', '#include "InterpDens_aux.h"'], ref_flags)
    assert ours == ref
