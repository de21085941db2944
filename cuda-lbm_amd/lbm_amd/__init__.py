"""lbm_amd — MI355X (gfx950) D2Q9 and D3Q27 lattice-Boltzmann library, Python side.

The compute path is liblbm_hip.so (hand-written HIP kernels behind the C ABI of
include/lbm_hip.h and include/lbm3d_hip.h); this package only binds it.
"""
from . import _abi as abi  # noqa: F401
from ._abi import load_library  # noqa: F401
from .solver import Config, LbmError, Solver  # noqa: F401
from .solver3d import Config3D, Solver3D  # noqa: F401
