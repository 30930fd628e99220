"""MI355X short-range engine (SRE): Python binding over libsre.so's C ABI.

The hot path (rc assignment, cell binning, Verlet list, LJ force/energy/
virial, tail correction, adjoint), the MD loop around it and the adaptive
rc-lattice producer run in hand-written gfx950 HIP kernels behind the C ABI
declared in include/sre.h; this package only marshals arguments.
``reference_api`` mirrors the reference's class names.
"""
from ._lib import LIB_PATH, lib  # noqa: F401
from .engine import (  # noqa: F401
    BoxTooSmall, Engine, ForceRange, InvalidArgument, MDExcptCuda, NotReady,
    SREError, ShortCellList, ShortNeighborList, SkinExceeded, check, gen_fcc, gen_slab)
from .arc import Arc  # noqa: F401,E402
