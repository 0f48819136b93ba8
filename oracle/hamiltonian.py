"""CPU oracle (test infrastructure only) for the discretised helium Hamiltonian.

This module is the checker that the HIP path is compared against.  It is
imported only by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``; the product library never calls it.

It restates, in numpy, the reference's discretisation of the two-electron
complex-scaled sinc-DVR Hamiltonian

    H(theta) = e^{-2 i theta} (T1 + T2) + e^{-i theta} V(r1, r2)

Every function cites the reference file:line it follows (paths relative to
``complexscale-djhnow/code/``).

Pinning (each is a test under ``tests/``):
  * ``kematrix``: closed form equal to ``sineDVR`` (SineDVR.F90:29-37) and the
    plane-wave symbol k^2/2 of the infinite sinc-DVR.
  * ``getinverse``: the defining identity T3D (R + Q) = 2 pi delta_0 on the small
    grid, exact delta^2 scaling, the far field delta^3/|n|, and (through the
    dense one-electron Hamiltonian) the reference's logged He+ eigenvalues
    (CSCALE-WORK/Outs/Out.S.N5:51-64) to ~1e-14.
  * ``potential2`` / ``hammult``: a dense Kronecker-product Hamiltonian built
    independently at NN=3 and NN=5, the exchange symmetry P H P = H and
    complex symmetry H^T = H; and, against the reference itself, the
    orbital-product subspace eigenvalues and the chosen shift of its logged
    singlet run (Out.S.N5:97-113, through ``startvec.diagsubspace``) to the
    accuracy of the log's Lanczos orbitals, which involves the whole
    two-electron operator including the e-e term.
"""
from __future__ import annotations

import numpy as np

NSMALL_DEFAULT = 22  # input.F90:31
TFACTOR_DEFAULT = 4  # input.F90:33


def kematrix(nn: int, delta: float) -> np.ndarray:
    """1-D sinc-DVR kinetic matrix T(i,j) (real), hammod.F90:235-260.

    T_ii = (pi/delta)^2 / 6,  T_ij = (-1)^(i-j) / (delta^2 (i-j)^2).
    """
    i = np.arange(nn)
    d = i[:, None] - i[None, :]
    dd = np.where(d == 0, 1, d).astype(np.float64)
    off = np.where(np.abs(d) % 2 == 0, 1.0, -1.0) / (delta * delta * dd * dd)
    return np.where(d == 0, (np.pi / delta) ** 2 / 6.0, off)


def _kinetic_rect(rows: np.ndarray, cols: np.ndarray, delta: float) -> np.ndarray:
    """Kinetic elements between integer grid labels (tooth.f90:152-160)."""
    d = rows[:, None] - cols[None, :]
    dd = np.where(d == 0, 1, d).astype(np.float64)
    off = np.where(np.abs(d) % 2 == 0, 1.0, -1.0) / (delta * delta * dd * dd)
    return np.where(d == 0, np.pi ** 2 / (6.0 * delta ** 2), off)


def getinverse(nout: int, delta: float, nsmall: int = NSMALL_DEFAULT,
               tfactor: int = TFACTOR_DEFAULT) -> np.ndarray:
    """Regularised Coulomb table Tinv(-nout:nout)^3, tooth.f90:14-101.

    R(n) = delta^3/|n delta| on the big grid [-nbig, nbig]^3 (RVECTOR,
    tooth.f90:118-139), TR = -(T_big R)|small + 2 pi delta_0
    (tooth.f90:43-44,143-169), Q = T3D^{-1} TR on the small grid by the 1-D
    eigendecomposition of T (tooth.f90:48-85), Tinv = R + Q on the small grid
    and R outside it (tooth.f90:87-99).  Returned as an array indexed
    [i+nout, j+nout, k+nout].
    """
    nbig = nsmall * tfactor
    if nbig < nout:
        raise ValueError("nbig < nout")  # tooth.f90:38-40
    g = np.arange(-nbig, nbig + 1, dtype=np.float64) * delta
    rad = np.sqrt(g[:, None, None] ** 2 + g[None, :, None] ** 2 + g[None, None, :] ** 2)
    with np.errstate(divide="ignore"):
        R = np.where(rad == 0.0, 0.0, delta ** 3 / rad)
    small = np.arange(-nsmall, nsmall + 1)
    big = np.arange(-nbig, nbig + 1)
    tsb = _kinetic_rect(small, big, delta)  # (ns, nb)
    s = slice(nbig - nsmall, nbig + nsmall + 1)
    tr = -(np.einsum("ip,pjk->ijk", tsb, R[:, s, s], optimize=True)
           + np.einsum("jp,ipk->ijk", tsb, R[s, :, s], optimize=True)
           + np.einsum("kp,ijp->ijk", tsb, R[s, s, :], optimize=True))
    tr[nsmall, nsmall, nsmall] += 2.0 * np.pi
    ts = kematrix(2 * nsmall + 1, delta)
    e, w = np.linalg.eigh(ts)  # DSYEV('V','U') at tooth.f90:70
    qq = np.einsum("ai,bj,ck,abc->ijk", w, w, w, tr, optimize=True)
    qq /= e[:, None, None] + e[None, :, None] + e[None, None, :]
    q = np.einsum("ia,jb,kc,abc->ijk", w, w, w, qq, optimize=True)
    tinv = R.copy()
    tinv[s, s, s] += q
    c = nbig
    return np.ascontiguousarray(tinv[c - nout:c + nout + 1, c - nout:c + nout + 1,
                                     c - nout:c + nout + 1])


class Hamiltonian:
    """Grid parameters and the pieces of H, initmod.F90:15-50,174-228.

    ``n`` is the reference's half-width N (NN = 2N+1 points per axis,
    initmod.F90:31); the nucleus of charge ``z`` sits at the origin grid point
    (input.F90:244-250 sets RNUC = 0).
    """

    def __init__(self, n: int, delta: float, theta: float, z: int = 2,
                 nsmall: int = NSMALL_DEFAULT, tfactor: int = TFACTOR_DEFAULT,
                 tinv: np.ndarray | None = None):
        self.n = n
        self.nn = 2 * n + 1
        self.m = self.nn ** 3
        self.d = self.m * self.m
        self.delta = float(delta)
        self.theta = float(theta)
        self.z = z
        self.scale = complex(np.cos(theta), np.sin(theta))  # initmod.F90:48
        self.nout = 2 * n + 1  # initmod.F90:39 (PARAMS%NBIG)
        self.tinv = getinverse(self.nout, delta, nsmall, tfactor) if tinv is None else tinv
        self.t_real = kematrix(self.nn, delta)
        self.t = self.t_real / self.scale ** 2  # initmod.F90:207-208

    # --- potentials ----------------------------------------------------
    def grid_labels(self) -> np.ndarray:
        """(M, 3) integer labels (i, j, k) - N with i fastest (hammod.F90:560-562)."""
        idx = np.arange(self.nn) - self.n
        k, j, i = np.meshgrid(idx, idx, idx, indexing="ij")
        return np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)

    def vnuc_real(self) -> np.ndarray:
        """A*Tinv(r) per one-electron grid point, A = -Z/delta^3 (hammod.F90:554,571)."""
        lab = self.grid_labels() + self.nout
        return (-self.z / self.delta ** 3) * self.tinv[lab[:, 0], lab[:, 1], lab[:, 2]]

    def potential1(self) -> np.ndarray:
        """One-electron V(P)/SCALE, hammod.F90:432-498 with initmod.F90:224."""
        return self.vnuc_real() / self.scale

    def potential2_real(self) -> np.ndarray:
        """Real V(P,Q) of POTENTIAL2 before the 1/SCALE, hammod.F90:502-582.

        Shape (M, M) in Fortran order V(P, Q): P is electron 1, Q electron 2.
        """
        vn = self.vnuc_real()
        lab = self.grid_labels()
        diff = lab[:, None, :] - lab[None, :, :] + self.nout
        vee = self.tinv[diff[..., 0], diff[..., 1], diff[..., 2]] / self.delta ** 3
        # V(P,Q) = A Tinv(r_Q) + A Tinv(r_P) + Tinv(r_Q - r_P)/delta^3
        return vn[None, :] + vn[:, None] + vee.T

    def potential2(self) -> np.ndarray:
        return self.potential2_real() / self.scale  # initmod.F90:224

    # --- operators -------------------------------------------------------
    def kemults(self, x: np.ndarray) -> np.ndarray:
        """KEMULTS on cubes, hammod.F90:264-286.  x has shape (..., k, j, i)."""
        t = self.t
        y = np.einsum("li,...kjl->...kji", t, x, optimize=True)
        y += np.einsum("lj,...kli->...kji", t, x, optimize=True)
        y += np.einsum("lk,...lji->...kji", t, x, optimize=True)
        return y

    def hammult(self, x: np.ndarray, v: np.ndarray | None = None) -> np.ndarray:
        """Y = H X for one two-electron vector, HAMMULT/KEMULT2 (hammod.F90:19-67,138-187).

        ``x`` is the flat Fortran-order vector X(i1,j1,k1,i2,j2,k2) of length NN^6.
        """
        nn = self.nn
        xm = x.reshape(self.m, self.m)  # row = electron-2 index Q, col = electron-1 index P
        x6 = xm.reshape(self.m, nn, nn, nn)  # [Q, k1, j1, i1]
        y = self.kemults(x6).reshape(self.m, self.m)  # electron 1 (hammod.F90:164-168)
        w = np.ascontiguousarray(xm.T).reshape(self.m, nn, nn, nn)  # MPITRANSPOSE, NP=1
        z = self.kemults(w).reshape(self.m, self.m)  # electron 2 (hammod.F90:175-179)
        y = y + z.T  # hammod.F90:182-185
        if v is None:
            v = self.potential2()
        # V(P,Q) is stored with P fastest: in the [Q, P] view it is V.T
        y = y + v.T * xm  # hammod.F90:59
        return y.reshape(-1)

    def dense(self) -> np.ndarray:
        """Dense H by Kronecker products (tiny grids only): independent of hammult."""
        nn = self.nn
        eye = np.eye(nn)
        t = self.t.T  # Y(i) = sum_l T(l,i) X(l): matrix element H[i,l] = T(l,i)
        # flat index p = i + nn*j + nn^2*k  ->  kron(K, J, I) ordering
        t3 = (np.kron(np.kron(eye, eye), t) + np.kron(np.kron(eye, t), eye)
              + np.kron(np.kron(t, eye), eye))
        em = np.eye(self.m)
        h = np.kron(em, t3) + np.kron(t3, em)
        vflat = self.potential2().reshape(-1, order="F")  # V(P,Q) flat with P fastest
        h[np.diag_indices_from(h)] += vflat
        return h

    def dense_one_electron(self) -> np.ndarray:
        nn = self.nn
        eye = np.eye(nn)
        t = self.t.T
        t3 = (np.kron(np.kron(eye, eye), t) + np.kron(np.kron(eye, t), eye)
              + np.kron(np.kron(t, eye), eye))
        t3 = t3.astype(np.complex128)
        t3[np.diag_indices_from(t3)] += self.potential1()
        return t3


def exchange(x: np.ndarray, m: int) -> np.ndarray:
    """Swap electrons: X(P,Q) -> X(Q,P) on a flat vector of length m*m."""
    return np.ascontiguousarray(x.reshape(m, m).T).reshape(-1)


def singlet_project(x: np.ndarray, m: int) -> np.ndarray:
    """(X + X^T)/2 in the M x M view (the start-block projection both sides use)."""
    return 0.5 * (x + exchange(x, m))
