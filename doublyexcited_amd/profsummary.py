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


# kernel -> the library's kernel class (the classes of its launch statistics):
# (class, whether a dispatch of it counts as one launch of the class).  A Gram
# launch is the MFMA partial kernel plus its partial-sum reduction, a residual
# step is its kernel plus a reduction, both counted once.
CLASSES = (
    ("k_gram_mfma", "gram", True), ("k_gram_pk", "gram", True), ("k_reduce_partials", "gram", False),
    ("k_combine", "combine", True),
    ("k_pass1", "pass1", True), ("k_symadd", "symadd", True), ("k_pass2", "pass2", True),
    ("k_coldots", "coldots", True), ("k_resid", "coldots", True), ("k_reduce_real", "coldots", False),
)


def kernel_class(name: str):
    n = short(name)
    for prefix, cls, counts in CLASSES:
        if n.split("<")[0].split()[-1] == prefix or n.split("<")[0].split()[-1].startswith(prefix + "_"):
            return cls, counts
    return None, False


def class_table(rows) -> list[str]:
    tot = {}
    for r in rows:
        cls, counts = kernel_class(r["Name"])
        if cls is None:
            continue
        t, n = tot.get(cls, (0.0, 0))
        tot[cls] = (t + float(r["TotalDurationNs"]), n + (int(r["Calls"]) if counts else 0))
    out = ["| class | launches | total ms | avg us per launch |", "|---|---:|---:|---:|"]
    for cls, (t, n) in sorted(tot.items(), key=lambda kv: -kv[1][0]):
        out.append(f"| {cls} | {n} | {t / 1e6:.1f} | {t / n / 1e3 if n else 0.0:.1f} |")
    return out


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
    out.append("")
    out.append("Per kernel class (the library's launch-statistics classes; a Gram launch includes its "
               "partial-sum reduction):\n")
    out.extend(class_table(rows))
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
