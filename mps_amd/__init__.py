"""mps_amd: MI355X-native engine for the U(1) block-sparse two-site DMRG hot path.

Host mirror of amiragha/mps ``SymTensors`` / ``TensorNetAlgs`` over the
libmpsh C ABI (include/mpsh.h); every data operation is a HIP kernel.
"""
from . import _lib  # noqa: F401
from .symtensors import (SymTensor, U1Space, allsectorsandsizes, array, axpby_, axpy_, contract,  # noqa: F401
                         dot, dual, fill, fill_linearindex, fuselegs, get_struct, issimilar, matmul, mul_, norm,
                         permutelegs, randn, rmul_, scale, splitleg, symmatrix, zeros)
from .tensornetalgs import _applymps2site, _mpsupdateleft, _mpsupdateright, heff_plan  # noqa: F401

__version__ = "0.1.0"
