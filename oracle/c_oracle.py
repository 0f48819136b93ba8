"""ctypes wrapper of the C oracle (oracle/hammult.c).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .hamiltonian import Hamiltonian

_LIB = Path(__file__).resolve().parent / "_build" / "liboracle.so"
_lib = None


def load():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            from .build_oracle import build

            build(verbose=False)
        _lib = C.CDLL(str(_LIB))
        dp = C.POINTER(C.c_double)
        _lib.oracle_hammult.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, dp, C.c_int]
        _lib.oracle_hammult.restype = C.c_int
        _lib.oracle_max_threads.restype = C.c_int
        _lib.oracle_potential2.argtypes = [C.c_int, C.c_int, dp, C.c_double, C.c_double, C.c_double,
                                           C.c_double, dp]
        _lib.oracle_potential2.restype = C.c_int
    return _lib


def potential2(h: Hamiltonian) -> np.ndarray:
    """POTENTIAL2 / SCALE through the C restatement: V(P,Q) flat with P fastest."""
    load()
    dp = C.POINTER(C.c_double)
    tinv = np.ascontiguousarray(h.tinv.transpose(2, 1, 0))  # [k][j][i]: i fastest in memory
    v = np.empty(h.d, dtype=np.complex128)
    _lib.oracle_potential2(h.nn, h.nout, tinv.ctypes.data_as(dp), float(h.z), h.delta, h.scale.real,
                           h.scale.imag, v.ctypes.data_as(dp))
    return v


class CHammult:
    """HAMMULT through the C restatement; holds T, V and the KEMULT2 scratch."""

    def __init__(self, h: Hamiltonian, nthreads: int = 0, c_potential: bool = False):
        self.h = h
        self.t = np.asfortranarray(h.t, dtype=np.complex128)
        if c_potential:  # large grids: the C restatement of POTENTIAL2 (no (M, M, 3) index array)
            self.v = potential2(h)
        else:
            self.v = np.ascontiguousarray(h.potential2().reshape(-1, order="F"), dtype=np.complex128)
        self.work = np.empty(2 * h.d, dtype=np.complex128)
        self.nthreads = nthreads
        load()

    def __call__(self, x: np.ndarray) -> np.ndarray:
        """x: (D,) or (D, k) Fortran-order complex128."""
        x = np.asfortranarray(x, dtype=np.complex128)
        k = 1 if x.ndim == 1 else x.shape[1]
        y = np.empty_like(x, order="F")
        dp = C.POINTER(C.c_double)
        _lib.oracle_hammult(self.h.nn, self.t.ctypes.data_as(dp), self.v.ctypes.data_as(dp),
                            x.ctypes.data_as(dp), y.ctypes.data_as(dp), k,
                            self.work.ctypes.data_as(dp), self.nthreads)
        return y

    @staticmethod
    def max_threads() -> int:
        return int(load().oracle_max_threads())
