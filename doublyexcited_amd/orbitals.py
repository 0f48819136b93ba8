"""One-electron (He+) orbitals for the orbital-product start vectors.

The reference obtains them with a block Lanczos on the one-electron operator
(lanmod.F90:18-219, KEMULT1/POTENTIAL1, hammod.F90:71-134,432-498).  Here the
one-electron Hamiltonian H1 = T3D / SCALE^2 + Vn / SCALE (SCALE = e^{i theta};
initmod.F90:204-224) is built on the host from the library's own tables
(``dex_host_kematrix``, ``dex_host_getinverse``) and diagonalised densely with
LAPACK, which is practical up to NN = 13 (M = 2197).  The lowest ``numorbs``
eigenpairs by real part are returned, as the Lanczos delivers them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DexError, check

_dp = C.POINTER(C.c_double)
MAX_DENSE_M = 2197


def one_electron_hamiltonian(n: int, delta: float, theta: float, charge: int = 2, nsmall: int = 22,
                             tfactor: int = 4) -> np.ndarray:
    """Dense complex-scaled He+ Hamiltonian on the (2n+1)^3 grid, index i fastest."""
    lib = _lib.load()
    nn = 2 * n + 1
    t = np.zeros((nn, nn), order="F")
    check(lib.dex_host_kematrix(nn, float(delta), t.ctypes.data_as(_dp)))
    L = 2 * nn + 1  # Tinv on [-NOUT, NOUT]^3 with NOUT = NN (initmod.F90:39)
    tinv = np.zeros(L ** 3)
    check(lib.dex_host_getinverse(nn, float(delta), nsmall, tfactor, tinv.ctypes.data_as(_dp)))
    tinv = tinv.reshape(L, L, L, order="F")
    idx = np.arange(nn) - n
    kk, jj, ii = np.meshgrid(idx, idx, idx, indexing="ij")
    vn = (-charge / delta ** 3) * tinv[ii.ravel() + nn, jj.ravel() + nn, kk.ravel() + nn]
    eye = np.eye(nn)
    t3 = np.kron(np.kron(eye, eye), t) + np.kron(np.kron(eye, t), eye) + np.kron(np.kron(t, eye), eye)
    scale = complex(np.cos(theta), np.sin(theta))
    h = t3.astype(np.complex128) / scale ** 2
    h[np.diag_indices_from(h)] += vn / scale
    return h


def heplus_orbitals(n: int, delta: float, theta: float, numorbs: int, charge: int = 2, nsmall: int = 22,
                    tfactor: int = 4):
    """(eigenvalues (numorbs), orbitals M x numorbs), lowest real parts first."""
    m = (2 * n + 1) ** 3
    if m > MAX_DENSE_M:
        raise DexError(_lib.DEX_EUNSUPPORTED,
                       f"dense He+ orbitals are built up to NN=13 (M={MAX_DENSE_M}); start from an "
                       "eigenvector file (Read / Infile=) on larger grids")
    w, v = np.linalg.eig(one_electron_hamiltonian(n, delta, theta, charge, nsmall, tfactor))
    o = np.argsort(w.real, kind="stable")[:numorbs]
    return w[o], v[:, o]
