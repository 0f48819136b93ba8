"""One-electron (He+) orbitals for the orbital-product start vectors.

The reference obtains them with a block Lanczos on the one-electron operator
(lanmod.F90:18-219, KEMULT1/POTENTIAL1, hammod.F90:71-134,432-498), which
``heplus_orbitals`` runs on the device (``dex_heplus_lanczos_device``) for any
grid.  ``method="dense"`` instead builds H1 = T3D / SCALE^2 + Vn / SCALE
(SCALE = e^{i theta}; initmod.F90:204-224) on the host from the library's own
tables (``dex_host_kematrix``, ``dex_host_getinverse``) and diagonalises it
with LAPACK, practical up to NN = 13 (M = 2197).  Either way the ``numorbs``
eigenpairs of smallest real part are returned, as the Lanczos delivers them.
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
                    tfactor: int = 4, method: str = "lanczos", blocksize: int = 3, kdim: int = 501,
                    check: int = 10, lantol: float = 1e-8, seed: int = 1):
    """(eigenvalues (numorbs), orbitals M x numorbs), lowest real parts first.

    ``method="lanczos"``: BLOCKLANCZOS on the device with LANBLOCKSIZE
    ``blocksize``, KDIM ``kdim``, CHECK ``check`` and LANTOL ``lantol``;
    ``"dense"``: the host eigendecomposition (NN <= 13)."""
    if method == "lanczos":
        from .hamiltonian import HeliumGrid

        with HeliumGrid(n, delta, theta, charge=charge, nsmall=nsmall, tfactor=tfactor) as g:
            e, orb, _, _ = g.heplus_lanczos(numorbs, blocksize=blocksize, kdim=kdim, check=check, tol=lantol,
                                            seed=seed)
        return e, orb
    if method != "dense":
        raise ValueError(f"unknown method {method!r}")
    m = (2 * n + 1) ** 3
    if m > MAX_DENSE_M:
        raise DexError(_lib.DEX_EUNSUPPORTED,
                       f"dense He+ orbitals are built up to NN=13 (M={MAX_DENSE_M}); use the device "
                       "block Lanczos (method='lanczos') on larger grids")
    w, v = np.linalg.eig(one_electron_hamiltonian(n, delta, theta, charge, nsmall, tfactor))
    o = np.argsort(w.real, kind="stable")[:numorbs]
    return w[o], v[:, o]
