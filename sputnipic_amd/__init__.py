"""sputnipic_amd -- MI355X (gfx950) implementation of sputniPIC's particle
hot path: the implicit predictor-corrector mover (mover_PC) and the
ten-moment CIC deposit (interpP2G), behind the reference's interface.

The compute lives in ``lib/libspic.so`` (hand-written HIP, C ABI in
``include/spic.h``); this package is the host-side mirror of the
reference's interface over that ABI.  There is no CPU fallback.
"""
from ._capi import SpicError, SPIC_EDOMAIN, SPIC_EINVAL, SPIC_EUNSUPPORTED  # noqa: F401
from ._capi import SPIC_POS_FIXED, SPIC_POS_FLOAT, SPIC_SORT_RADIX, SPIC_SORT_TILE  # noqa: F401
from ._capi import SPIC_SUBCYCLE_LOOP, SPIC_SUBCYCLE_REFUSE, SPIC_SUBCYCLE_SINGLE  # noqa: F401
from ._capi import SPIC_ORDER_ENGINE, SPIC_ORDER_LOAD  # noqa: F401
from .core import (Batch, DeviceArray, DeviceParticles, EMField, Engine, Event, GemParams, Graph,  # noqa: F401
                   InterpDensAux, MaxwellImage, SolveParams, applyBCscalar, calculateHatDensities,
                   interpN2Crho_net,
                   GemSpecies, Grid, InterpDensNet, InterpDensSpecies, Parameters,
                   Particles, Stream, FIELD_NAMES, MOMENT_NAMES, PARTICLE_NAMES, applyBCids,
                   applyBCids_all, beam_species, check_domain, device_count, gem_fields, gem_particles, get_async, harris_species,
                   interpP2G, mover_interp, mover_PC, particle_sort, set_async,
                   set_strict_fp, set_subcycling, setZeroSpeciesDensities, sumOverSpecies)

__all__ = [
    "SpicError", "DeviceArray", "EMField", "Grid", "InterpDensSpecies", "Parameters",
    "Particles", "mover_PC", "interpP2G", "mover_interp", "particle_sort",
    "setZeroSpeciesDensities", "applyBCids", "sumOverSpecies", "InterpDensNet",
    "set_strict_fp", "set_async", "device_count",
]
