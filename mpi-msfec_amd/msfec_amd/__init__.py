"""MI355X multiscale basis construction (MsFEC Q1, the 2-D Q1 x Nedelec
pairing, the Raviart-Thomas x DQ0 pairing in 2-D and 3-D and the 3-D
Nedelec x Raviart-Thomas pairing), Python face.

The compute path is libmsfec_hip.so (hand-written HIP for gfx950) behind the
C-ABI in include/msfec.h; this package only binds it (ctypes) and generates
seeded synthetic inputs.
"""
from . import _lib
from ._lib import MsfecError
from .device import DeviceBuffer
from .ned import NedBasisBatch
from .nrt import NedRTBasisBatch
from .qbasis import QBasisBatch
from .rt import RTDQBasisBatch
from .rt3 import RTDQ3BasisBatch
from .uneven import UnevenBatch

__all__ = ["DeviceBuffer", "QBasisBatch", "NedBasisBatch", "RTDQBasisBatch", "RTDQ3BasisBatch",
           "NedRTBasisBatch", "UnevenBatch", "MsfecError", "_lib"]
