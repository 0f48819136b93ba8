"""HBM traffic per kernel launch from two rocprofv3 counter passes.

Usage::

    python -m doublyexcited_amd.traffic FETCH_DIR WRITE_DIR STATS_JSON [OUT_JSON]

``FETCH_DIR`` and ``WRITE_DIR`` hold the ``*counter_collection.csv`` of a
``rocprofv3 --pmc FETCH_SIZE`` and a ``rocprofv3 --pmc WRITE_SIZE`` run of the
same program (``bench.py --traffic-run STATS_JSON ...``), which first runs a
calibration kernel (an elementwise multiply that reads and writes 1 GiB) and
then the profiled solve, whose per-class algorithmic bytes and launch counts it
writes to ``STATS_JSON``.  The calibration dispatch scales each counter to
bytes (FETCH_SIZE reports half of the bytes read on this part); per kernel
class the scaled traffic per dispatch is divided by the algorithmic bytes per
launch.  Prints a Markdown table and writes ``{"traffic_over_algorithmic":
{class: ratio}}`` to ``OUT_JSON`` when given.
"""
from __future__ import annotations

import csv
import json
import sys
from pathlib import Path

# kernel-name prefixes of the library's classes (dex_get_kernel_stats order)
CLASSES = {
    "pass1": ("k_pass1",),
    "pass2": ("k_pass2",),
    "symadd": ("k_symadd",),
    "gram": ("k_gram_mfma", "k_gram_pk", "k_reduce_partials"),
    "combine": ("k_combine",),
    "coldots": ("k_coldots", "k_resid", "k_reduce_real"),
    "generic": ("k_axis", "k_potential"),
}
# partial-sum reductions that finish a launch of their class (not launches themselves)
HELPERS = ("k_reduce_partials", "k_reduce_real")
CAL_BYTES = 1 << 30  # the calibration kernel reads and writes 1 GiB


def _base(name: str) -> str:
    """The kernel's own name: no return type, namespaces, template or
    parameter list."""
    n = name.replace("(anonymous namespace)::", "")
    return n.split("<")[0].split("(")[0].replace("void ", "").strip().split("::")[-1]


def _rows(d: Path, counter: str):
    files = sorted(Path(d).rglob("*counter_collection.csv"))
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    out = []
    for f in files:
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    out.append((r.get("Kernel_Name", ""), int(r.get("Dispatch_Id", 0)), float(r["Counter_Value"])))
    return out


def _per_dispatch(rows):
    acc = {}
    for name, disp, val in rows:
        k = (name, disp)
        acc[k] = acc.get(k, 0.0) + val
    return acc


def _cls(name: str):
    base = _base(name)
    for c, prefixes in CLASSES.items():
        if any(base.startswith(p) for p in prefixes):
            return c
    return None


def analyse(fetch_dir, write_dir, stats_json):
    fetch = _per_dispatch(_rows(fetch_dir, "FETCH_SIZE"))
    write = _per_dispatch(_rows(write_dir, "WRITE_SIZE"))
    # calibration: the elementwise multiply, the largest counter value among torch's kernels
    # (the fill before it writes as much and reads nothing)
    def cal(acc):
        cands = [v for (n, d), v in acc.items() if "at::native" in n and v > 0]
        return max(cands) if cands else None
    cf, cw = cal(fetch), cal(write)
    f_scale = CAL_BYTES / (cf * 1024.0)
    w_scale = CAL_BYTES / (cw * 1024.0)
    stats = json.loads(Path(stats_json).read_text())
    totals = stats.get("totals", stats)
    table = []
    ratios = {}
    for c in CLASSES:
        fv = [v for (n, _), v in fetch.items() if _cls(n) == c and not _base(n).startswith(HELPERS)]
        wv = [v for (n, _), v in write.items() if _cls(n) == c and not _base(n).startswith(HELPERS)]
        if not fv or c not in totals:
            continue
        launches, alg_bytes = totals[c][0], totals[c][1]
        if launches <= 0:
            continue
        # the class's small helper kernels (partial-sum reductions) add their own traffic
        fh = [v for (n, _), v in fetch.items() if _cls(n) == c]
        wh = [v for (n, _), v in write.items() if _cls(n) == c]
        traffic = (sum(fh) * f_scale + sum(wh) * w_scale) * 1024.0 / len(fv)
        alg = alg_bytes / launches
        ratios[c] = traffic / alg
        table.append((c, len(fv), sum(fv) / len(fv), sum(wv) / len(wv), traffic / 1e9, alg / 1e9, traffic / alg))
    return cf, cw, f_scale, w_scale, table, ratios


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 3:
        print(__doc__)
        return 2
    cf, cw, fs, ws, table, ratios = analyse(argv[0], argv[1], argv[2])
    print(f"Calibration (1 GiB read + 1 GiB written): FETCH_SIZE {cf:.1f} KiB, WRITE_SIZE {cw:.1f} KiB; "
          f"scale FETCH x {fs:.4f}, WRITE x {ws:.4f}\n")
    print("| class | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | scaled traffic GB/launch "
          "| algorithmic GB/launch | traffic / algorithmic |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for c, n, f, w, t, a, r in table:
        print(f"| {c} | {n} | {f:.1f} | {w:.1f} | {t:.4f} | {a:.4f} | {r:.3f} |")
    if len(argv) > 3:
        Path(argv[3]).write_text(json.dumps({"source": argv[2], "traffic_over_algorithmic": ratios}, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
