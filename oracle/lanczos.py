"""CPU oracle (test infrastructure only) for the one-electron block Lanczos.

Imported only by ``tests/``; the product never calls it.  A numpy restatement
of ``BLOCKLANCZOS`` (complexscale-djhnow/code/lanmod.F90:18-219) with its
``GRAMSCHMIDT`` (:228-244) and ``EIGSORT`` (:262-284), on the one-electron
operator ``H1 = T3D / SCALE^2 + Vn / SCALE`` (KEMULT1 + POTENTIAL1,
hammod.F90:71-134,432-498), step by step: column Gram-Schmidt in the
bilinear form (MPIDOT is sum(x * y), no conjugation) against the stored basis,
the previous block and the earlier new columns, normalisation by the complex
square root of x^T x, twice; the block tridiagonal matrix diagonalised by
LAPACK ZGEEV every CHECK steps; Ritz vectors normalised and residuals measured
in the Hermitian norm (MPIHERMDOT).

Pinning (tests/test_oracle.py): the eigenvalues equal those of the dense H1
(numpy eig of the Kronecker-built matrix) at small grids, and at NN=11,
delta=2, theta=0.2 with the logged run's block size 3 and NUMORBS 14 they
reproduce the He+ eigenvalues the reference printed (CSCALE-WORK/Outs/
Out.S.N5:51-64) to the accuracy of its LANTOL = 1e-2 convergence.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import scipy.linalg as sla

from .hamiltonian import Hamiltonian


@dataclasses.dataclass
class LanczosResult:
    eigenvalues: np.ndarray
    vectors: np.ndarray   # M x numeigs, Hermitian-normalised
    residuals: np.ndarray
    iterations: int
    converged: bool


def h1_apply(h: Hamiltonian, x: np.ndarray) -> np.ndarray:
    """KEMULT1 (T / SCALE^2 on the three axes) plus POTENTIAL1 / SCALE on a
    (M,) or (M, k) block, index i fastest."""
    nn = h.nn
    xb = x.reshape(nn, nn, nn, -1, order="F")  # [i, j, k, col]
    t = h.t  # T(l, i) / SCALE^2 (initmod.F90:207-208)
    y = np.einsum("li,ljkc->ijkc", t, xb, optimize=True)
    y += np.einsum("lj,ilkc->ijkc", t, xb, optimize=True)
    y += np.einsum("lk,ijlc->ijkc", t, xb, optimize=True)
    y = y.reshape(x.shape, order="F") if x.ndim == 2 else y.reshape(-1, order="F")
    v = h.potential1()
    return y + (v[:, None] * x if x.ndim == 2 else v * x)


def _gramschmidt(u: np.ndarray, v: np.ndarray, n: int) -> np.ndarray:
    """GRAMSCHMIDT(U, V, N): V -= (U_k . V) U_k for k = 1..N-1, in order
    (lanmod.F90:228-244; the dot is the bilinear MPIDOT)."""
    for k in range(n - 1):
        d = np.sum(u[:, k] * v)
        v = v - d * u[:, k]
    return v


def block_lanczos(h: Hamiltonian, bb: int, kdim: int, check: int, tol: float, numeigs: int,
                  v0: np.ndarray) -> LanczosResult:
    """BLOCKLANCZOS with GSORTH and the start block ``v0`` (M x bb) in place of
    RANDOM_NUMBER (lanmod.F90:76-79)."""
    m = h.m
    v = np.array(v0, dtype=np.complex128, copy=True)
    for _ in range(2):  # lanmod.F90:81-86
        for ib in range(bb):
            v[:, ib] = _gramschmidt(v, v[:, ib], ib + 1)
            v[:, ib] = v[:, ib] / np.sqrt(np.sum(v[:, ib] * v[:, ib]))
    vmat = np.zeros((m, bb * kdim), np.complex128)
    tdm = np.zeros((bb * kdim, bb * kdim), np.complex128)
    bprev = None
    res = None
    for j in range(1, kdim):
        vmat[:, (j - 1) * bb:j * bb] = v
        w = h1_apply(h, v)
        a = v.T @ w  # A(JB, IB, J) = V_JB . W_IB (lanmod.F90:89-96)
        vold = v.copy()
        vnew = w.copy()
        for _ in range(2):  # lanmod.F90:101-116
            for ib in range(bb):
                x = _gramschmidt(vmat, vnew[:, ib], j * bb + 1)
                x = _gramschmidt(vold, x, bb + 1)
                vnew[:, ib] = x
                vnew[:, ib] = _gramschmidt(vnew, vnew[:, ib], ib + 1)
                vnew[:, ib] = vnew[:, ib] / np.sqrt(np.sum(vnew[:, ib] * vnew[:, ib]))
        v = vnew
        b = v.T @ w  # B(IB, JB, J+1) = V_IB . W_JB (lanmod.F90:120-124)
        o = (j - 1) * bb
        tdm[o:o + bb, o:o + bb] = a
        if j < kdim:
            tdm[o:o + bb, o + bb:o + 2 * bb] = b.T
        if j > 1:
            tdm[o:o + bb, o - bb:o] = bprev
        bprev = b
        if j % check:
            continue
        n = j * bb
        lam, z = sla.eig(tdm[:n, :n])  # ZGEEV (lanmod.F90:139-147)
        ind = np.argsort(lam.real, kind="stable")[:numeigs]  # EIGSORT on the real part
        u = vmat[:, :n] @ z[:, ind]
        u = u / np.sqrt(np.sum(np.abs(u) ** 2, axis=0))[None, :]
        r = h1_apply(h, u) - u * lam[ind][None, :]
        resid = np.sqrt(np.sum(np.abs(r) ** 2, axis=0))
        res = LanczosResult(lam[ind], u, resid, j, bool(resid.max() < tol))
        if res.converged:
            return res
    if res is None:
        raise RuntimeError("LANCZOS FAILED TO CONVERGE.")
    return res
