"""Markdown summary of a rocprofv3 ``--kernel-trace --stats`` run.

Usage::

    python -m doublyexcited_amd.profsummary STATS_CSV [TITLE] [COMMAND]

``STATS_CSV`` is the ``*kernel_stats.csv`` rocprofv3 writes with ``-f csv``.
Prints the total kernel time and, per kernel, calls, total and average
duration and share, largest first; kernel names are shortened to the
library's template names.
"""
from __future__ import annotations

import csv
import re
import sys


def short(name: str) -> str:
    n = re.sub(r"dex::\(anonymous namespace\)::", "", name)
    n = re.sub(r"HIP_vector_type<(\w+), \d+u>", r"\1x", n)
    n = re.sub(r"\(.*\)$", "", n).strip()
    return n


def summarize(path: str, title: str = "", command: str = "") -> str:
    rows = list(csv.DictReader(open(path, newline="")))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    calls = sum(int(r["Calls"]) for r in rows)
    out = []
    if title:
        out.append(f"# {title}\n")
    if command:
        out.append("```\n" + command + "\n```\n")
    out.append(f"Total kernel time {tot / 1e9:.3f} s over {calls} dispatches.\n")
    out.append("| kernel | calls | total ms | avg us | % |")
    out.append("|---|---:|---:|---:|---:|")
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        t = float(r["TotalDurationNs"])
        if t < 0.001 * tot:
            continue
        out.append(f"| `{short(r['Name'])}` | {r['Calls']} | {t / 1e6:.1f} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{100.0 * t / tot:.2f} |")
    return "\n".join(out) + "\n"


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print(__doc__)
        return 2
    print(summarize(argv[0], argv[1] if len(argv) > 1 else "", argv[2] if len(argv) > 2 else ""))
    return 0


if __name__ == "__main__":
    sys.exit(main())
