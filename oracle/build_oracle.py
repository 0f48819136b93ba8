"""Builds the C oracle ``oracle/_build/liboracle.so`` (gcc, OpenMP).  Test
infrastructure: used only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build" / "liboracle.so"


def build(verbose: bool = True) -> Path:
    OUT.parent.mkdir(exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-std=c11",
           str(HERE / "hammult.c"), "-o", str(tmp), "-lm"]
    subprocess.run(cmd, check=True)
    tmp.replace(OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
