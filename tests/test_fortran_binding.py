"""The Fortran ISO_C_BINDING interface of the binding notes compiles with
flang, links against libdex.so and calls host-only entry points (no GPU)."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle.hamiltonian import getinverse, kematrix

ROOT = Path(__file__).resolve().parent.parent
FLANG = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"


@pytest.mark.skipif(not Path(FLANG).exists(), reason="flang not available")
def test_fortran_binding_compiles_links_and_runs(tmp_path):
    src = ROOT / "tests" / "fortran" / "dex_binding_check.f90"
    exe = tmp_path / "dex_binding_check"
    libdir = ROOT / "doublyexcited_amd" / "lib"
    subprocess.run([FLANG, "-O1", "-J", str(tmp_path), str(src), "-o", str(exe), f"-L{libdir}", "-ldex",
                    f"-Wl,-rpath,{libdir}"], check=True, cwd=tmp_path, timeout=300)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout.split()
    assert out[:3] == ["2", "22", "0"]  # dex_params_default N, NSMALL; dex_solve_opts_default fp64_restarts
    t = kematrix(5, 0.5)
    assert float(out[3]) == pytest.approx(t[0, 0], rel=1e-15)
    assert float(out[4]) == pytest.approx(t[1, 0], rel=1e-15)
    assert float(out[5]) == pytest.approx(getinverse(5, 0.5)[8, 5, 5], rel=1e-12)
