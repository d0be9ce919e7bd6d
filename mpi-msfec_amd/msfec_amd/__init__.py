"""MI355X multiscale basis construction (MsFEC Q1, and the 2-D Q1 x Nedelec and
Raviart-Thomas x DQ0 pairings), Python face.

The compute path is libmsfec_hip.so (hand-written HIP for gfx950) behind the
C-ABI in include/msfec.h; this package only binds it (ctypes) and generates
seeded synthetic inputs.
"""
from . import _lib
from ._lib import MsfecError
from .device import DeviceBuffer
from .ned import NedBasisBatch
from .qbasis import QBasisBatch
from .rt import RTDQBasisBatch
from .uneven import UnevenBatch

__all__ = ["DeviceBuffer", "QBasisBatch", "NedBasisBatch", "RTDQBasisBatch", "UnevenBatch", "MsfecError", "_lib"]
