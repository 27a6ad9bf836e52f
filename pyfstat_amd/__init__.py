"""MI355X-native transient-CW F-statistic maps for PyFstat.

The package provides a ``tCWFstatMapVersion="hip"`` backend: hand-written
gfx950 HIP kernels (``csrc/tcw_map.hip``, built into ``_lib/libtcw_hip.so``)
behind a C ABI (``include/tcw_hip.h``), loaded with ctypes, with PyTorch-ROCm
holding device memory and streams.

Modules:
  tcw_fstat_map_funcs  drop-in counterpart of pyfstat.tcw_fstat_map_funcs
  batched              many templates in one device pass
  atoms                atom and window-range containers (lalpulsar stand-ins)
  windows              window ranges from ComputeFstat-style options
  synth                seeded synthetic atoms for the benchmark workloads
  build                in-tree hipcc build of the library
"""

from . import atoms, synth, windows
from .atoms import (
    TRANSIENT_EXPONENTIAL,
    TRANSIENT_LAST,
    TRANSIENT_NONE,
    TRANSIENT_RECTANGULAR,
    CreateFstatAtomVector,
    CreateMultiFstatAtomVector,
    FstatAtomVector,
    MultiFstatAtomVector,
    transientWindowRange_t,
)

__version__ = "0.1.0"

__all__ = [
    "atoms",
    "synth",
    "windows",
    "TRANSIENT_NONE",
    "TRANSIENT_RECTANGULAR",
    "TRANSIENT_EXPONENTIAL",
    "TRANSIENT_LAST",
    "CreateFstatAtomVector",
    "CreateMultiFstatAtomVector",
    "FstatAtomVector",
    "MultiFstatAtomVector",
    "transientWindowRange_t",
]
