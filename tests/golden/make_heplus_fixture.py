"""Writes heplus_nn11_d2_t0.2.json: the one-electron (He+) eigenvalues the
reference printed for NN=11, delta=2, theta=0.2, Z=2.

The numbers are copied from the reference's own run logs (Lanczos output,
CSCALE-WORK/Outs/Out.S.N5:51-64, reproduced in CSCALE-WORK/Outs.first/Out.S.N5:51-64),
together with the log's orbital-product subspace eigenvalues and expectation
values (DIAGSUBSPACE / GET_EXPECTS over 14 singlet orbitals, target -0.76;
Out.S.N5:97-112), its chosen GPLMR shift (:113) and the two-electron singlet
resonance GPLMR reached from there (Out.S.N5:147).
Run: python tests/golden/make_heplus_fixture.py /path/to/reference
"""
import json
import re
import sys
from pathlib import Path


def main(ref: str) -> None:
    log = Path(ref) / "CSCALE-WORK" / "Outs" / "Out.S.N5"
    lines = log.read_text().splitlines()
    start = next(i for i, ln in enumerate(lines) if "COMPUTED EIGENVALUES" in ln)
    vals = []
    for ln in lines[start + 1:start + 15]:
        a, b = ln.split()
        vals.append([float(a), float(b)])
    final = next(ln for ln in lines if re.match(r"\s+-0\.7131\d+\s+0\.0352\d+\s*$", ln))
    fa, fb = final.split()
    # SUBSPACE EIGENVALS: value (re, im) and |E0 - expectation|; "--xx--" splits
    # the NEV chosen ones from the rest
    i0 = next(i for i, ln in enumerate(lines) if "SUBSPACE EIGENVALS" in ln)
    sub = []
    for ln in lines[i0 + 1:]:
        if "SUBSPACE EXPECTATION" in ln:
            break
        if "--xx--" in ln:
            continue
        a, b, _ = ln.split()
        sub.append([float(a), float(b)])
    i1 = next(i for i, ln in enumerate(lines) if "SUBSPACE EXPECTATION VALS" in ln)
    ea, eb = lines[i1 + 1].split()
    shift = next(ln for ln in lines if "CHOSEN SHIFT" in ln).split()[-2:]
    out = {
        "source": "CSCALE-WORK/Outs/Out.S.N5:51-64 (He+ Lanczos), :147 (singlet GPLMR, residual 7.99e-8)",
        "params": {"N": 5, "NN": 11, "delta": 2.0, "theta": 0.2, "charge": 2, "nsmall": 22, "tfactor": 4},
        "heplus_eigenvalues": vals,
        "singlet_resonance": [float(fa), float(fb)],
        "singlet_resonance_residual": 7.99e-08,
        "subspace": {"numorbs": 14, "singlet": True, "hermopt": False, "etarget": [-0.76, 0.0],
                     "eigenvalues": sub, "first_expectation": [float(ea), float(eb)],
                     "chosen_shift": [float(shift[0]), float(shift[1])],
                     "gplmr": {"nev": 1, "ns": 5, "inner": "PETSc GMRES, gm_max_inner_iter 5, gm_usepetscflag 0",
                               "iterations": 29}},
    }
    dst = Path(__file__).resolve().parent / "heplus_nn11_d2_t0.2.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
