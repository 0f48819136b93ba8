"""Eigenvector files in the reference's format (READWRITE_EIGENVECTORS,
utils.F90:82-181), for warm starts across theta / grid continuations and
for checkpointing solves.  Thin wrappers over the library's C-ABI.

* ``write_eigenvectors(path, psi)``: psi is (NUM, M*M) complex128 (one vector per row,
  electron 1 fastest within a vector, as the reference's ALLPSI(M, M)).
* ``read_eigenvectors(path, m=None, num=None)``: (NUM, M*M); with ``m`` different from
  the file's, each vector passes through SIXD_FIX (utils.F90:184-198).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check

_dp = C.POINTER(C.c_double)


def write_eigenvectors(path: str | os.PathLike, psi: np.ndarray) -> None:
    psi = np.ascontiguousarray(np.atleast_2d(psi), dtype=np.complex128)
    num, mm = psi.shape
    m = int(round(mm ** 0.5))
    if m * m != mm:
        raise ValueError("each vector must hold M*M entries")
    lib = _lib.load()
    check(lib.dex_write_eigvecs(os.fsencode(path), psi.view(np.float64).ctypes.data_as(_dp), m, num))


def eigenvector_file_info(path: str | os.PathLike) -> tuple[int, int]:
    """(M, NUM) from the header record."""
    lib = _lib.load()
    m, num = C.c_int32(), C.c_int32()
    check(lib.dex_read_eigvecs_info(os.fsencode(path), C.byref(m), C.byref(num)))
    return m.value, num.value


def read_eigenvectors(path: str | os.PathLike, m: int | None = None, num: int | None = None) -> np.ndarray:
    m_file, num_file = eigenvector_file_info(path)
    m = m_file if m is None else m
    num = num_file if num is None else num
    out = np.zeros((num, m * m), dtype=np.complex128)
    lib = _lib.load()
    check(lib.dex_read_eigvecs(os.fsencode(path), m, num, out.view(np.float64).ctypes.data_as(_dp)))
    return out


def sixd_fix(psi: np.ndarray, n_in: int, n_out: int) -> np.ndarray:
    """SIXD_FIX (utils.F90:184-198) of one vector or of rows of vectors: the
    central block common to the (2 n_in + 1)^6 and (2 n_out + 1)^6 grids."""
    psi = np.ascontiguousarray(np.atleast_2d(psi), dtype=np.complex128)
    d_out = (2 * n_out + 1) ** 6
    out = np.zeros((psi.shape[0], d_out), dtype=np.complex128)
    lib = _lib.load()
    for v in range(psi.shape[0]):
        row = np.ascontiguousarray(psi[v])
        dst = np.zeros(d_out, dtype=np.complex128)
        check(lib.dex_host_sixd_fix(row.view(np.float64).ctypes.data_as(_dp), n_in,
                                    dst.view(np.float64).ctypes.data_as(_dp), n_out))
        out[v] = dst
    return out
