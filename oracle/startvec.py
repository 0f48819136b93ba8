"""CPU oracle (test infrastructure only) for the physically targeted start
vectors: He+ orbitals and the orbital-product subspace.

Imported only by ``tests/``; the product never routes through it.

* ``heplus_orbitals``: the lowest one-electron eigenpairs (by real part) of the
  dense complex-scaled He+ Hamiltonian, standing in for the reference's block
  Lanczos (lanmod.F90:18-219), whose converged vectors span the same levels.
* ``orbital_gather``: HELIUM_ORBITALGATHER's Gram-Schmidt (hammod.F90:1080-1113).
* ``diagsubspace``: HELIUM_ORBITAL (hammod.F90:1041-1075) and DIAGSUBSPACE
  (hammod.F90:1125-1256) with mysort (hammod.F90:1658-1690), plus GET_EXPECTS
  (hammod.F90:1261-1301) with the Hermitian product.

Pinning: the subspace eigenvalues and the first expectation value are checked
against the reference's logged values (CSCALE-WORK/Outs/Out.S.N5:97-112) at
the level the log's Lanczos orbitals allow (they were converged to LANTOL
1e-2: the degenerate members printed there differ by up to 6e-6).
"""
from __future__ import annotations

import numpy as np


def heplus_orbitals(h, numorbs: int):
    """(eigenvalues, eigenvectors M x numorbs) of the dense one-electron H, lowest real parts first."""
    w, v = np.linalg.eig(h.dense_one_electron())
    o = np.argsort(w.real, kind="stable")[:numorbs]
    return w[o], v[:, o]


def orbital_gather(x, hermopt: bool):
    """HELIUM_ORBITALGATHER (hammod.F90:1096-1112): one Gram-Schmidt pass with
    the bilinear product (hermopt False) or the Hermitian one."""
    x = np.array(x, dtype=np.complex128, copy=True)
    dot = (lambda a, b: np.vdot(a, b)) if hermopt else (lambda a, b: np.sum(a * b))
    for i in range(x.shape[1]):
        for j in range(i):
            x[:, i] -= x[:, j] * dot(x[:, j], x[:, i])
        x[:, i] /= np.sqrt(dot(x[:, i], x[:, i]))
    return x


def helium_orbital(x, i: int, j: int, sign: int):
    """HELIUM_ORBITAL (hammod.F90:1041-1075): PSI(p, q) = x_i(p) x_j(q) + sign x_j(p) x_i(q),
    halved for i = j, else divided by sqrt 2; flattened with p (electron 1) fastest."""
    psi = np.outer(x[:, i], x[:, j]) + sign * np.outer(x[:, j], x[:, i])
    psi /= 2.0 if i == j else np.sqrt(2.0)
    return psi.reshape(-1, order="F")


def mysort_order(values, etarget):
    """mysort (hammod.F90:1658-1690): repeatedly the untaken index of smallest
    |value - etarget|, the last such index on ties."""
    n = len(values)
    taken = np.zeros(n, bool)
    order = []
    for _ in range(n):
        best, low = -1, 0.0
        for i in range(n):
            if taken[i]:
                continue
            dv = abs(values[i] - etarget)
            if best < 0 or dv <= low:
                best, low = i, dv
        taken[best] = True
        order.append(best)
    return order


def diagsubspace(amul, orbitals, singlet: bool, hermopt: bool, etarget: complex, numout: int):
    """DIAGSUBSPACE + GET_EXPECTS.  ``amul(block)`` applies H to a (D, k) block.
    Returns (e0 (numout), psi (D, numout) Hermitian-normalised, expects (numout))."""
    x = orbital_gather(orbitals, hermopt)
    k = x.shape[1]
    sign = 1 if singlet else -1
    pairs = [(a, b) for a in range(k) for b in range(a + (0 if singlet else 1), k)]
    basis = np.stack([helium_orbital(x, a, b, sign) for a, b in pairs], axis=1)
    hb = amul(basis)
    amat = (basis.conj().T if hermopt else basis.T) @ hb  # A(I, J) (:1196-1220)
    e, vr = np.linalg.eig(amat)  # ZGEEV (:1226)
    order = mysort_order(e, etarget)[:numout]
    psi = basis @ vr[:, order]
    psi /= np.sqrt(np.einsum("ij,ij->j", psi.conj(), psi).real)[None, :]
    hp = amul(psi)
    expects = np.einsum("ij,ij->j", psi.conj(), hp) / np.einsum("ij,ij->j", psi.conj(), psi)
    return e[order], psi, expects
