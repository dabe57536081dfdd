"""tnhip — MI355X-native MPS contraction hot path of TensorNetworks.jl.

Host-side mirror (Python) of the reference's Julia API for the DMRG/TEBD path,
calling the HIP kernels of libtnhip.so through its C-ABI (include/tnhip.h).
The library is required: there is no CPU fallback for any hot-path operator.
"""
from ._lib import TnhError, version  # noqa: F401  (raises ImportError if the .so is missing)
from .core import Context, Tensor  # noqa: F401
from . import ops  # noqa: F401
from .models import GMPS, MPO, OpList, Sitetypes, productMPS, spinhalf  # noqa: F401
from .mps import DMRGState, dmrg, movecenter, normalize, randomMPS, replacesites  # noqa: F401
