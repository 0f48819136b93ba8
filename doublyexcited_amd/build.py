"""Build the native library ``lib/libdex.so`` for gfx950 with hipcc (in-tree).

Usage: ``python -m doublyexcited_amd.build [--jobs N]``.  Every call rebuilds.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
BUILDDIR = PKG / "_build"
ROOT = PKG.parent
SOURCES = ["dex_setup.cpp", "dex_smallla.cpp", "dex_api.cpp", "dex_eigvecs.cpp", "dex_lanczos.cpp",
           "dex_kernels.hip"]
ARCH = os.environ.get("DEX_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def common_flags() -> list[str]:
    return [
        "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
        # keep the compiler from emitting packed FP32 instructions
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{ROOT / 'include'}", f"-I{CSRC}",
        "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    ]


def _compile(src: Path, obj: Path) -> None:
    cmd = [hipcc(), *common_flags(), "-x", "hip", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    warn = [ln for ln in res.stderr.splitlines() if "warning" in ln and "packed-fp32-ops" not in ln]
    if warn:
        sys.stderr.write("\n".join(warn[:20]) + "\n")


def build(jobs: int = 4, verbose: bool = True) -> Path:
    BUILDDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = []
        for s in SOURCES:
            src = CSRC / s
            obj = BUILDDIR / (Path(s).stem + ".o")
            objs.append(obj)
            futs.append(ex.submit(_compile, src, obj))
        for f in futs:
            f.result()
    out = LIBDIR / "libdex.so"
    tmp = LIBDIR / "libdex.so.tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *[str(o) for o in objs]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    args = ap.parse_args()
    build(args.jobs)
