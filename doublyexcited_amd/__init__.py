"""MI355X implementation of the helium two-electron complex-scaled sinc-DVR
matvec and the GPLMR / harmonic-Ritz block eigensolver (djhaxton/doublyexcited
hot path), as a native HIP library with a C-ABI and this Python mirror."""
from ._lib import DexError  # noqa: F401
from .hamiltonian import HeliumGrid, MatvecGraph, SolveResult  # noqa: F401

__all__ = ["HeliumGrid", "MatvecGraph", "SolveResult", "DexError"]
