"""clmsoil — MI355X (gfx950) per-column implicit soil solves for CLM.jl.

Host-side mirror of CLM.jl's interfaces for the soil-water (moisture-form
Richards) and snow/soil heat steps; the compute runs in the HIP library
`libclmsoil.so` behind the C ABI declared in include/clmsoil.h.
"""
from ._abi import (BC_AQUIFER, BC_FLUX, BC_HEAD, BC_WATERTABLE, BC_ZERO_FLUX,  # noqa: F401
                   MOISTURE_FORM, ZENGDECKER_2009, LibraryMissing, load_library)

__all__ = ["load_library", "LibraryMissing"]
