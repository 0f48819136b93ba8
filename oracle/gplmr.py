"""CPU oracle (test infrastructure only) for the GPLMR / harmonic-Ritz block eigensolver.

Imported only by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``; the product never routes through it.

A numpy restatement of ``GPLMR_EIG`` (complexscale-djhnow/code/gplmr.F90:9-386)
and its dense linear-algebra helpers (lamod.f90:21-180), with the projected
pencil solved by LAPACK ZGGEV through scipy exactly as ``SOLVE_GEVP`` does
(gplmr.F90:395-439).  The operator is a callback ``amul(block) -> block`` on
(D, k) complex128 blocks, the ``MATBLOCK`` convention (gplmr.F90:443-462).

Pinning: the fixed point is checked against an independent dense
eigendecomposition (numpy ``eig``) at tiny grids, the lamod identities
(Q^H Q = I, R upper triangular) are checked directly, and the solver is driven to
the reference's logged singlet resonance at NN=11, delta=2, theta=0.2.
"""
from __future__ import annotations

import dataclasses
from typing import Callable

import numpy as np
import scipy.linalg as sla


class GplmrError(RuntimeError):
    """Raised where the reference prints and STOPs."""


def normalize(x):
    """lamod.f90:21-48: unit 2-norm columns."""
    nrm = np.sqrt(np.einsum("ij,ij->j", x.real, x.real) + np.einsum("ij,ij->j", x.imag, x.imag))
    x /= nrm[None, :]
    return x


def crossmult_diag(x, y):
    """lamod.f90:51-76: D = diag(X^H Y)."""
    return np.einsum("ij,ij->j", x.conj(), y)


def colnorms(x):
    """lamod.f90:79-104: sqrt(diag(X^H X))."""
    return np.sqrt(np.einsum("ij,ij->j", x.real, x.real) + np.einsum("ij,ij->j", x.imag, x.imag))


def crossmult(x, y):
    """lamod.f90:108-125: Z = X^H Y (ZGEMM 'C','N')."""
    return x.conj().T @ y


STRICT_CHOLESKY = False  # True: fail where the reference STOPs (lamod.f90:169-175)
PIVOT_TOL = 1e-12  # smallest accepted pivot of the unit-diagonal Gram


def _upper_chol(g):
    return np.linalg.cholesky(g).conj().T  # upper factor: G = R^H R


def orth_block(x, drop, strict=None):
    """Orthonormalise the columns of ``x``: cholQR (lamod.f90:128-180) plus a fallback.

    Returns ``(xq, t)`` with ``xq = x @ t``.  Where ZPOTRF succeeds and the
    unit-diagonal Gram has all pivots above ``PIVOT_TOL``, ``t = R^{-1}`` with
    ``R`` the reference's upper Cholesky factor, exactly as cholQR.  Where the
    reference would STOP ("Cholesky failed") or continue with a numerically
    singular R, this restatement (unless ``strict``) takes the eigendecomposition
    of the unit-diagonal Gram (SVQB, Stathopoulos & Wu, SIAM J. Sci. Comput. 23
    (2002) 2165) and, when ``drop``, keeps only the independent directions
    (fewer columns), otherwise clamps the small eigenvalues (same column count);
    one more cholQR pass then restores orthonormality.
    """
    strict = STRICT_CHOLESKY if strict is None else strict
    g = crossmult(x, x)
    if np.sum(np.abs(g)) == 0.0:
        raise GplmrError("zero block in cholQR")  # lamod.f90:157-166
    dg = np.sqrt(np.maximum(np.real(np.diag(g)), 0.0))
    ok = bool(np.all(dg > 0))
    if ok:
        try:
            r = _upper_chol(g)
            piv = (np.real(np.diag(r)) / dg) ** 2
            ok = bool(np.all(piv > PIVOT_TOL))
        except np.linalg.LinAlgError:
            ok = False
    if ok:
        t = sla.solve_triangular(r, np.eye(len(dg), dtype=np.complex128), lower=False)
        return x @ t, t
    if strict:
        raise GplmrError("Cholesky failed")  # lamod.f90:169-175
    k = g.shape[0]
    dinv = np.where(dg > 0, 1.0 / np.where(dg > 0, dg, 1.0), 0.0)
    gs = dinv[:, None] * g * dinv[None, :]
    lam, u = np.linalg.eigh(gs)
    lmax = lam.max()
    if drop:
        keep = lam > PIVOT_TOL * lmax
        t = dinv[:, None] * (u[:, keep] / np.sqrt(lam[keep])[None, :])
    else:
        lamc = np.maximum(lam, PIVOT_TOL * lmax)
        t = dinv[:, None] * (u / np.sqrt(lamc)[None, :])
    xq = x @ t
    g2 = crossmult(xq, xq)
    r2 = _upper_chol(g2)
    t2 = sla.solve_triangular(r2, np.eye(g2.shape[0], dtype=np.complex128), lower=False)
    return xq @ t2, t @ t2


def solve_gevp(kp, mp, k, nev):
    """gplmr.F90:395-439: ZGGEV of the leading k x k pencil, NEV smallest |alpha/beta|.

    EIGSORT (gplmr.F90:640-660) picks the first minimum on ties (MINLOC).
    Returns (vectors (k, nev), mu (nev)).
    """
    a = kp[:k, :k]
    b = mp[:k, :k]
    try:
        w, vr = sla.eig(a, b, right=True, homogeneous_eigvals=True)
    except (np.linalg.LinAlgError, ValueError) as exc:
        raise GplmrError("SOLVE OF THE PROJECTED EIGENVALUE PROBLEM FAILED") from exc
    alpha, beta = w
    with np.errstate(divide="ignore", invalid="ignore"):
        lam = np.abs(alpha / beta)
    lam = np.where(np.isnan(lam), np.inf, lam)
    idx = np.argsort(lam, kind="stable")[:nev]
    return vr[:, idx], (alpha / beta)[idx]


def _herm_sqrt_pair(g):
    """hermsqrtmatsmooth with tol = 0 (hr_eigen.f90:214-291): G^{-1/2} and G^{1/2}
    from the SVD of the Hermitian PSD G; zero singular values stay zero."""
    u, s, vh = np.linalg.svd(g)
    v = vh.conj().T
    sinv = np.where(s != 0.0, 1.0 / np.where(s != 0.0, s, 1.0), 0.0)
    ginvhalf = (v * np.sqrt(sinv)[None, :]) @ u.conj().T
    ghalf = (v * np.sqrt(s)[None, :]) @ u.conj().T
    return ginvhalf, ghalf


def basic_block_lan_solve(amul, b, sigma, order):
    """PRECBLOCK_BASIC -> basicblocklansolve (gplmr.F90:674-770;
    harmonic_ritz_small.F90:258-324): a fixed-order, unrestarted block
    minimal-residual approximation to (H - sigma)^{-1} b.

    Followed step by step, with every column of the block multiplied (the
    reference's ``allhmultsub`` multiplies only the first column,
    gplmr.F90:704-716, which is the NEV > 1 defect it is fixed here):

    * ``harm_orthog_oneblock`` (harmonic_ritz_small.F90:158-245): B0 = (H - sigma) b,
      G = B0^H B0, V1 = b G^{-1/2} so that (H - sigma) V1 is orthonormal, and
      ``lantrans`` = G^{1/2};
    * ``fast_harmritz_setup`` (:100-155): blocks V_k = H V_{k-1}, B_k = (H - sigma) V_k,
      each column Gram-Schmidt'ed twice (``myhgramschmidt_many``, :28-81, with
      factors (1, 0): the inner product is that of the B part) against every
      earlier column, the same combination applied to V;
    * the solve (:309-318): y = (B^H V) [G^{1/2}; 0], x = V y.

    Returns x with the shape of b.
    """
    if order <= 0:
        raise GplmrError("ORDERERROR")  # harmonic_ritz_small.F90:284-286
    k = b.shape[1]
    hb = amul(b) - sigma * b
    ginvhalf, ghalf = _herm_sqrt_pair(crossmult(hb, hb))
    v1 = b @ ginvhalf.conj().T  # ZGEMM('N','C') (:204-205)
    dim = b.shape[0]
    vall = np.empty((dim, order * k), np.complex128)
    ball = np.empty((dim, order * k), np.complex128)
    for io in range(order):
        v = v1.copy() if io == 0 else amul(vall[:, (io - 1) * k:io * k])
        bb = amul(v) - sigma * v
        for i in range(k):
            col = io * k + i
            vi, bi = v[:, i], bb[:, i]
            for _ in range(2):
                if col:
                    c = ball[:, :col].conj().T @ bi
                    bi = bi - ball[:, :col] @ c
                    vi = vi - vall[:, :col] @ c
                nrm = np.sqrt(np.vdot(bi, bi).real)
                bi = bi / nrm
                vi = vi / nrm
            ball[:, col] = bi
            vall[:, col] = vi
    lanham = ball.conj().T @ vall
    rhs = np.zeros((order * k, k), np.complex128)
    rhs[:k] = ghalf
    return vall @ (lanham @ rhs)


@dataclasses.dataclass
class GplmrResult:
    x: np.ndarray
    e: np.ndarray
    resnrm: np.ndarray
    iters: int
    converged: bool
    matvecs: int
    history: list


def gplmr_eig(amul: Callable[[np.ndarray], np.ndarray], x0: np.ndarray, sigma: complex,
              nev: int, ns: int = 1, maxit: int = 100, tol: float = 1e-8,
              inner_iter: int = 0, callback=None, lock: bool = False, nconv: int | None = None) -> GplmrResult:
    """GPLMR_EIG, gplmr.F90:9-386.

    ``nconv`` (not in the reference; default ``nev``): convergence is tested
    on the first ``nconv`` columns only, the others being guard columns that
    the harmonic selection keeps behind the wanted pairs (it orders them
    nearest first).

    ``inner_iter == 1`` applies PRECBLOCK's ``W <- H W`` (gplmr.F90:491-493);
    ``0`` skips the preconditioner (guards at gplmr.F90:92,118,151); ``>= 2``
    runs PRECBLOCK_BASIC's block solve (``basic_block_lan_solve``).

    ``lock`` (not in the reference): search directions W, S and P are built
    only for columns whose residual is still above ``tol``; every column stays
    in X and in the projected pencil.  With ``lock=False`` every column
    contributes directions, as in the reference.
    """
    if ns < 1:
        raise GplmrError("NS must be >= 1")  # R is (NS*NEV, NEV), gplmr.F90:36
    if inner_iter < 0:
        raise GplmrError("GM_MAX_INNER_ITER must be >= 0")
    x = np.array(x0, dtype=np.complex128, copy=True)
    dim = x.shape[0]
    ldr = ns * nev
    nproj = (3 + ns) * nev
    nmat = 0

    def matblock(b):
        nonlocal nmat
        nmat += b.shape[1]
        return amul(b)

    def prec(b):
        # PRECBLOCK (gplmr.F90:466-502): 0 never called, 1 -> MATBLOCK, else BASIC
        if inner_iter == 0:
            return b
        if inner_iter == 1:
            return matblock(b)
        return basic_block_lan_solve(matblock, b, sigma, inner_iter)

    normalize(x)
    ax = matblock(x)
    e = crossmult_diag(x, ax)
    w = ax - x * e[None, :]
    nconv = nev if nconv is None else nconv
    resnrm = colnorms(w)
    maxres = resnrm[:nconv].max()
    p = np.zeros((dim, 0), np.complex128)
    ap = np.zeros_like(p)
    it = 1
    history = []
    while it <= maxit and maxres > tol:
        active = np.flatnonzero(resnrm > tol) if lock else np.arange(nev)
        w = w[:, active]
        x, t = orth_block(x, drop=False)  # :88
        ax = ax @ t  # :90
        w = prec(w)
        w = w - x @ crossmult(x, w)  # :101-103
        w, _ = orth_block(w, drop=True)  # :105
        aw = matblock(w)  # :108
        # S_1 = AW - W E  (:114-116); E pairs with the leading active columns
        eact = e[active]
        ew = eact[:w.shape[1]]
        sblocks, asblocks = [], []
        snew = aw - w * ew[None, :]
        for l in range(1, ns + 1):  # :114-172
            if l > 1:
                prev, aprev = sblocks[-1], asblocks[-1]
                snew = aprev - prev * eact[:prev.shape[1]][None, :]
            snew = prec(snew)
            snew = snew - x @ crossmult(x, snew)  # :127-128
            snew = snew - w @ crossmult(w, snew)  # :129-130
            if sblocks:
                sall = np.concatenate(sblocks, axis=1)
                snew = snew - sall @ crossmult(sall, snew)
            snew, _ = orth_block(snew, drop=True)  # :131
            sblocks.append(snew)
            asblocks.append(matblock(snew))  # :134
        s = np.concatenate(sblocks, axis=1)
        as_ = np.concatenate(asblocks, axis=1)
        if it > 1 and p.shape[1] > 0:  # :177-196
            r = crossmult(x, p)
            p = p - x @ r
            ap = ap - ax @ r
            r = crossmult(w, p)
            p = p - w @ r
            ap = ap - aw @ r
            r = crossmult(s, p)
            p = p - s @ r
            ap = ap - as_ @ r
            p, t = orth_block(p, drop=True)
            ap = ap @ t
        # harmonic projection (:201-270)
        q = np.concatenate([x, w, s, p], axis=1)
        aq = np.concatenate([ax, aw, as_, ap], axis=1)
        asig = aq - sigma * q
        kp = asig.conj().T @ asig
        mp = asig.conj().T @ q
        k = q.shape[1]  # iteration 1 has no P block (:286-290)
        coords, mu = solve_gevp(kp, mp, k, nev)
        nw, nsb = w.shape[1], s.shape[1]
        cx = coords[:nev]
        cw = coords[nev:nev + nw]
        cs = coords[nev + nw:nev + nw + nsb]
        cp = coords[nev + nw + nsb:]
        p_new = s @ cs + w @ cw  # :307-312
        ap_new = as_ @ cs + aw @ cw
        if p.shape[1] > 0:  # :298-305
            p_new += p @ cp
            ap_new += ap @ cp
        x = x @ cx + p_new  # :316-318
        ax = ax @ cx + ap_new
        p, ap = p_new[:, active], ap_new[:, active]
        nrm = colnorms(x)  # :324
        x /= nrm[None, :]
        ax /= nrm[None, :]
        e = crossmult_diag(x, ax)  # :332
        w = ax - x * e[None, :]  # :334-336
        resnrm = colnorms(w)
        maxres = resnrm[:nconv].max()
        history.append((it, e.copy(), resnrm.copy()))
        if callback is not None:
            callback(it, e, resnrm)
        it += 1
    return GplmrResult(x=x, e=e, resnrm=resnrm, iters=it - 1, converged=bool(maxres <= tol),
                       matvecs=nmat, history=history)
