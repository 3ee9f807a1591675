"""lb_amd -- MI355X lattice-Boltzmann collide-stream core (D3Q27, A-A in-place, HIP/gfx950).

The compute path is the HIP shared library lb_amd/liblbm_core_hip.so behind the
C ABI in include/lbm_core.h; this package is the Python-side handle on it.
"""
from . import abi
from .abi import load_core
from .core import Lattice, LbmError, make_desc

__all__ = ["abi", "load_core", "Lattice", "LbmError", "make_desc"]
