"""Theta / grid continuation scans: a sequence of resonance solves, each
started from the previous solve's eigenvectors (mapped by SIXD_FIX onto the
next grid when N changes), as the reference's scripts chain runs through
its eigenvector files (CSCALE-WORK/BAK/runsometheta:22-29, runsomen:43-52:
each run reads the previous run's Bin/ vectors, READWRITE_EIGENVECTORS with
READFLAG, utils.F90:82-181).  Here the vectors stay on the GPU between
solves of one grid; ``checkpoint_dir`` also writes each point's vectors in
the reference's file format.
"""
from __future__ import annotations

import dataclasses
import os
import time

import numpy as np


@dataclasses.dataclass
class ScanPoint:
    n: int
    delta: float
    theta: float
    e: np.ndarray
    residuals: np.ndarray
    iterations: int
    converged_pairs: int
    seconds: float
    warm: bool


def theta_scan(points, sigma: complex, nev: int, tol: float, ns: int = 1, inner_iter: int = 0,
               guard: int = 0, maxit: int = 1000, warm: bool = True, seed: int = 1004,
               precision: str = "fp64", checkpoint_dir: str | os.PathLike | None = None,
               device: int = 0, on_point=None) -> list[ScanPoint]:
    """Solves the singlet resonances nearest ``sigma`` at each (n, delta, theta)
    of ``points`` in order.  Cold points start from seeded complex-Gaussian
    singlet blocks (seed + index); warm points from the previous point's
    returned vectors.  ``on_point(point)``, if given, is called with each
    ScanPoint as soon as it is solved (progress reporting)."""
    import torch

    from .eigvecs import sixd_fix, write_eigenvectors
    from .hamiltonian import HeliumGrid

    out: list[ScanPoint] = []
    prev, prev_n = None, None
    for i, (n, delta, theta) in enumerate(points):
        with HeliumGrid(n, delta, theta, precision=precision, device=device) as g:
            x = torch.empty(g.d * nev, dtype=torch.complex128, device=f"cuda:{device}")
            use_warm = warm and prev is not None
            if use_warm and prev_n == n:
                x.copy_(prev)
            elif use_warm:
                host = prev.cpu().numpy().reshape(nev, -1)
                x.copy_(torch.from_numpy(sixd_fix(host, prev_n, n).reshape(-1)).to(x.device))
            else:
                g.fill_gaussian_device(x, nev, seed=seed + i, symmetry=1)
            t0 = time.time()
            r = g.gplmr_eig_device(x, sigma, nev, ns=ns, maxit=maxit, tol=tol, symmetry=1,
                                   inner_iter=inner_iter, guard=guard, fp32_tol=max(tol, 1e-5))
            res = g.residuals_device(x, r.e)
            dt = time.time() - t0
            out.append(ScanPoint(n, delta, theta, r.e, res, r.iterations, int(np.sum(res <= tol)), dt, use_warm))
            if on_point is not None:
                on_point(out[-1])
            if checkpoint_dir is not None:
                write_eigenvectors(os.path.join(checkpoint_dir, f"eigvecs.S.N{n}.T{theta:.4f}"),
                                   x.cpu().numpy().reshape(nev, -1))
            prev, prev_n = x, n
    return out
