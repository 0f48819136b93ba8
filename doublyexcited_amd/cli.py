"""Command-line driver with the reference's input deck and arguments.

``python -m doublyexcited_amd.cli [Inp=Input.Inp] [N=5] [Theta=0.2] [Delta=2] ...``

* The deck is the reference's namelist file (input.F90:144-244: groups
  &main, &grid, &molecule, &lanczos, &gmres, &orbitals, &gplmr; a key given
  twice takes its last value, as a Fortran namelist read does).
* Command-line arguments override it as ``get_args`` does (input.F90:264-362):
  N=, Shift=, Outfile=, Infile=, NoPetsc, Petsc, Theta=, R= (AUTODELTA on),
  Delta= (AUTODELTA off), Nev=, Norbs=, Gptol=, Skip=, Gmtol=, Iter=, Read,
  Takereal, Notakereal, Sing, Trip.
* The run follows the reference driver's hot path (doublyexcited.F90:120-246):
  build H; start vectors from He+ orbital products (DIAGSUBSPACE) or from an
  eigenvector file (Read, with SIXD_FIX); the GPLMR shift is the mean Hermitian
  expectation value of the start vectors (real part with TAKEREALPART); then
  GPLMR_EIG, printing "ITERATION" / "EIG j = ... ERROR = ..." every OUTMODULUS
  iterations and "FINAL EIGVALS:"; WRITEVECS writes Bin/<OUTVECFILE>.

Deliberately not supported (an error says so): several nuclei or a nucleus
off the grid (NUCLEI != 1, NUC_ON_GRID false), STRETCH / MODBASIS / VQUAD, the
PETSc inner solve (gm_usepetscflag 0 runs the in-house solve instead, and the
run says so), and HERMOPT-dependent c-norm shifts (CSHIFTOPT true).
"""
from __future__ import annotations

import re
import sys
from pathlib import Path

import numpy as np

# input.F90:6-92 defaults of the hot path's parameters
DEFAULTS = {
    "THETA": 0.0, "WRITEVECS": False, "READVECS": False, "INVECFILE": "eigvecs", "OUTVECFILE": "eigvecs",
    "STRETCH": False, "MODBASIS": False, "VQUAD": False,
    "N": 2, "DELTA": 0.5, "NSMALL": 22, "TFACTOR": 4, "RADIUS": 7.0, "AUTODELTA": True,
    "NUCLEI": 1, "CHARGE": 2, "NUC_ON_GRID": True,
    "LANBLOCKSIZE": 3, "KDIM": 2001, "CHECK": 10, "LANTOL": 1e-6,
    "GM_USEPETSCFLAG": 1, "GM_MAX_INNER_ITER": 5, "GM_TOL": 1e-3,
    "NUMORBS": 14, "SINGLET": False, "NUMSKIP": 0,
    "OUTMODULUS": 5, "TAKEREALPART": True, "GPLMR_NEV": 1, "GPLMR_NS": 5, "GPLMR_MAXIT": 100,
    "GPLMR_TOL": 1e-12, "GPLMR_SHIFT": complex(-999.0, 0.0), "HERMOPT": False, "CSHIFTOPT": False,
}


def _value(text: str):
    t = text.strip().rstrip(",").strip()
    if not t:
        return None
    low = t.lower()
    if low in (".true.", "t", ".t."):
        return True
    if low in (".false.", "f", ".f."):
        return False
    if t[0] in "'\"":
        return t.strip("'\"")
    m = re.fullmatch(r"\(\s*([^,]+),\s*([^)]+)\)", t)
    if m:
        return complex(float(_fnum(m.group(1))), float(_fnum(m.group(2))))
    try:
        return int(t)
    except ValueError:
        return float(_fnum(t))


def _fnum(t: str) -> str:
    return t.strip().lower().replace("d", "e")


def parse_namelist(text: str) -> dict:
    """Fortran namelist groups -> {KEY: value} (keys upper-cased; the last
    assignment of a key wins; '!' starts a comment)."""
    out = {}
    in_group = False
    for raw in text.splitlines():
        line = raw.split("!", 1)[0].strip()
        if not line:
            continue
        if line.startswith("&"):
            in_group = True
            rest = line[1:].split(None, 1)
            line = rest[1] if len(rest) > 1 else ""
            if not line:
                continue
        if line.startswith("/"):
            in_group = False
            continue
        if not in_group:
            continue
        for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*=\s*(\([^)]*\)|'[^']*'|\"[^\"]*\"|[^,\s]+)", line):
            out[m.group(1).upper()] = _value(m.group(2))
    return out


def apply_args(cfg: dict, args: list[str]) -> dict:
    """get_args (input.F90:264-362)."""
    cfg = dict(cfg)
    for a in args:
        def val(prefix):
            return _value(a[len(prefix):])
        if a.startswith("N="):
            cfg["N"] = int(val("N="))
        elif a.startswith("Shift="):
            v = val("Shift=")
            cfg["GPLMR_SHIFT"] = complex(v)
        elif a.startswith("Outfile="):
            cfg["OUTVECFILE"] = a[len("Outfile="):]
        elif a.startswith("Infile="):
            cfg["INVECFILE"] = a[len("Infile="):]
        elif a.startswith("NoPetsc"):
            cfg["GM_USEPETSCFLAG"] = 0
        elif a.startswith("Petsc"):
            cfg["GM_USEPETSCFLAG"] = 1
        elif a.startswith("Theta="):
            cfg["THETA"] = float(val("Theta="))
        elif a.startswith("R="):
            cfg["AUTODELTA"] = True
            cfg["RADIUS"] = float(val("R="))
        elif a.startswith("Delta="):
            cfg["AUTODELTA"] = False
            cfg["DELTA"] = float(val("Delta="))
        elif a.startswith("Nev="):
            cfg["GPLMR_NEV"] = int(val("Nev="))
        elif a.startswith("Norbs="):
            cfg["NUMORBS"] = int(val("Norbs="))
        elif a.startswith("Gptol="):
            cfg["GPLMR_TOL"] = float(val("Gptol="))
        elif a.startswith("Skip="):
            cfg["NUMSKIP"] = int(val("Skip="))
        elif a.startswith("Gmtol="):
            cfg["GM_TOL"] = float(val("Gmtol="))
        elif a.startswith("Iter="):
            cfg["GM_MAX_INNER_ITER"] = int(val("Iter="))
        elif a.startswith("Read"):
            cfg["NUMSKIP"] = 0
            cfg["READVECS"] = True
        elif a.startswith("Takereal"):
            cfg["TAKEREALPART"] = True
        elif a.startswith("Notakereal"):
            cfg["TAKEREALPART"] = False
        elif a.startswith("Sing"):
            cfg["SINGLET"] = True
        elif a.startswith("Trip"):
            cfg["SINGLET"] = False
    return cfg


def load_config(argv: list[str]) -> dict:
    inp = "Input.Inp"
    for a in argv:
        if a.startswith("Inp="):
            inp = a[4:]
    cfg = dict(DEFAULTS)
    if Path(inp).exists():
        cfg.update(parse_namelist(Path(inp).read_text()))
    cfg = apply_args(cfg, [a for a in argv if not a.startswith("Inp=")])
    if cfg["AUTODELTA"]:
        cfg["DELTA"] = cfg["RADIUS"] / cfg["N"]  # INIT_STRUCTS (input.F90:392)
    return cfg


def run(cfg: dict, out=sys.stdout, workdir: str | Path = ".") -> dict:
    """The driver's hot path; returns {"e": eigenvalues, "residuals": ..., "shift": ...}."""
    import torch

    from .eigvecs import read_eigenvectors, write_eigenvectors
    from .hamiltonian import HeliumGrid
    from .orbitals import heplus_orbitals

    if cfg["NUCLEI"] != 1 or not cfg["NUC_ON_GRID"]:
        raise SystemExit("only one nucleus on a grid point is supported")
    if cfg["STRETCH"] or cfg["MODBASIS"] or cfg["VQUAD"]:
        raise SystemExit("STRETCH / MODBASIS / VQUAD are not supported")
    if cfg["CSHIFTOPT"]:
        raise SystemExit("CSHIFTOPT (c-norm shift) is not supported")
    n, delta, theta = int(cfg["N"]), float(cfg["DELTA"]), float(cfg["THETA"])
    nev, ns = int(cfg["GPLMR_NEV"]), int(cfg["GPLMR_NS"])
    singlet = bool(cfg["SINGLET"])
    sym = 1 if singlet else -1
    inner = int(cfg["GM_MAX_INNER_ITER"])
    if inner > 1 and int(cfg["GM_USEPETSCFLAG"]) == 0:
        print("gm_usepetscflag 0 (PETSc) is not built: the in-house inner solve runs instead", file=out)
    print(f"GRID POINTS PER DIMENSION: {2 * n + 1:8d}", file=out)
    print(f"GRID SPACING: {delta:8.2f}", file=out)
    print(f"THETA: {theta:8.4f}   {'SINGLET' if singlet else 'TRIPLET'}", file=out)
    bindir = Path(workdir) / "Bin"
    with HeliumGrid(n, delta, theta, charge=int(cfg["CHARGE"]), nsmall=int(cfg["NSMALL"]),
                    tfactor=int(cfg["TFACTOR"])) as g:
        if cfg["READVECS"]:
            vecs = read_eigenvectors(bindir / str(cfg["INVECFILE"]), m=g.m, num=nev)
            x = torch.from_numpy(vecs.reshape(-1).copy()).to("cuda")
            ex = np.zeros(nev, dtype=np.complex128)
            for j in range(nev):
                xj = x[j * g.d:(j + 1) * g.d]
                y = torch.empty_like(xj)
                g.matvec_device(xj, y, 1, symmetry=sym)
                ex[j] = complex(torch.vdot(xj, y).item() / torch.vdot(xj, xj).item())
        else:
            numskip = int(cfg["NUMSKIP"])
            # DIAGONALIZE -> BLOCKLANCZOS (doublyexcited.F90:138, utils.F90:219) on the device,
            # to the deck's LANTOL at most (the product start vectors need no more)
            print("CALLING LANCZOS SOLVER...", file=out)
            _, orb = heplus_orbitals(n, delta, theta, int(cfg["NUMORBS"]), charge=int(cfg["CHARGE"]),
                                     nsmall=int(cfg["NSMALL"]), tfactor=int(cfg["TFACTOR"]), method="lanczos",
                                     blocksize=int(cfg["LANBLOCKSIZE"]), kdim=int(cfg["KDIM"]),
                                     check=int(cfg["CHECK"]), lantol=min(float(cfg["LANTOL"]), 1e-8))
            print("COMPUTING INITIAL GUESSES...", file=out)
            psi, e0, ex = g.diagsubspace_device(orb, nev + numskip, complex(cfg["GPLMR_SHIFT"]),
                                                singlet=singlet, hermopt=bool(cfg["HERMOPT"]))
            x = psi[numskip * g.d:].clone()
            ex = ex[numskip:]
            print("SUBSPACE EIGENVALS:", file=out)
            for v in e0[numskip:]:
                print(f"{v.real:20.14f}{v.imag:20.14f}", file=out)
        print("SUBSPACE EXPECTATION VALS:", file=out)
        for v in ex:
            print(f"{v.real:20.14f}{v.imag:20.14f}", file=out)
        shift = complex(np.mean(ex[:nev]))  # doublyexcited.F90:214
        if cfg["TAKEREALPART"]:
            shift = complex(shift.real, 0.0)
        print(f"CHOSEN SHIFT: {shift.real:20.10f}{shift.imag:20.10f}", file=out)
        out.flush()
        r = g.gplmr_eig_device(x, shift, nev, ns=ns, maxit=int(cfg["GPLMR_MAXIT"]), tol=float(cfg["GPLMR_TOL"]),
                               inner_iter=inner, outmodulus=int(cfg["OUTMODULUS"]), verbose=True, symmetry=sym)
        res = g.residuals_device(x, r.e)
        print(" FINAL EIGVALS:", file=out)
        for v in r.e:
            print(f"{v.real:20.12f}{v.imag:20.12f}", file=out)
        if cfg["WRITEVECS"]:
            bindir.mkdir(parents=True, exist_ok=True)
            print(" WRITING eigvects...", file=out)
            write_eigenvectors(bindir / str(cfg["OUTVECFILE"]), x.cpu().numpy().reshape(nev, -1))
    return {"e": r.e, "residuals": res, "shift": shift, "iterations": r.iterations, "converged": r.converged}


def main(argv: list[str] | None = None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    cfg = load_config(argv)
    r = run(cfg)
    return 0 if r["converged"] else 1


if __name__ == "__main__":
    sys.exit(main())
