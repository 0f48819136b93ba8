"""Writes heplus_nn11_d2_t0.2.json: the one-electron (He+) eigenvalues the
reference printed for NN=11, delta=2, theta=0.2, Z=2.

The numbers are copied from the reference's own run logs (Lanczos output,
CSCALE-WORK/Outs/Out.S.N5:51-64, reproduced in CSCALE-WORK/Outs.first/Out.S.N5:51-64),
together with the log's two-electron singlet resonance (Out.S.N5:147) and
the reference driver's quoted converged-limit energies (doublyexcited.F90:9-12).
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
    out = {
        "source": "CSCALE-WORK/Outs/Out.S.N5:51-64 (He+ Lanczos), :147 (singlet GPLMR, residual 7.99e-8)",
        "params": {"N": 5, "NN": 11, "delta": 2.0, "theta": 0.2, "charge": 2, "nsmall": 22, "tfactor": 4},
        "heplus_eigenvalues": vals,
        "singlet_resonance": [float(fa), float(fb)],
        "singlet_resonance_residual": 7.99e-08,
    }
    dst = Path(__file__).resolve().parent / "heplus_nn11_d2_t0.2.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
