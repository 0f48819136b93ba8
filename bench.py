#!/usr/bin/env python3
"""Benchmark: converged resonance eigenpairs/s and solve latency of the GPLMR /
harmonic-Ritz solver on one MI355X, FP64 beside the mixed form.

FP64 form: every matvec and block operation in FP64 (the inner solve too).
Mixed form: the outer matvec in FP32 (complex128 vectors, complex64
intermediate) until the residual reaches 1e-5, then FP64 restarts to the
tolerance (with --mixed-inner-fp32 also the inner solve's Arnoldi basis in
complex64 with FP32 matvecs).  Orthogonalisation, the projected eigenproblem
and the final residuals stay FP64 in both forms.

A step is one complete resonance solve of the selected workload: a seeded
complex-Gaussian singlet start block taken by GPLMR (with the in-house inner
minimal-residual solve) to the residual tolerance.  The Hamiltonian tables and
every start block are resident in HBM before the timed region; up to
``--concurrency`` solves are in flight at once, each on its own context and
HIP stream (a small-grid solve alone leaves the GPU mostly idle).  ``value`` is
the whole job's converged eigenpairs (final FP64 residual recomputed from
scratch <= tol) per second of the timed region.

Also measured in the same run:
  roofline      the dominant kernel class of the timed solves, timed live with
                HIP events on the library streams; algorithmic bytes (or flops)
                per launch are counted on the host at each launch;
  cpu_baseline  the C oracle's HAMMULT (oracle/hammult.c, OpenMP on the host
                cores) timed on a bounded sample and scaled to the GPU solves'
                matvec counts (block linear algebra not counted: an upper bound
                on the CPU's eigenpairs/s).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                       [--concurrency C] [--no-mixed] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
RIDGE = FP64_PEAK_TFS * 1e12 / (HBM_PEAK_GBS * 1e9)  # flop/B where FP64 compute and HBM balance
# kernel timing in the timed solves: HIP events on a random ~1 in TIMING_SAMPLE
# launches (events on every launch cost about 10% of the solve time)
TIMING_SAMPLE = 8

WORKLOADS = {
    # the plain CPU-runnable configuration of the reference method
    "he-nn11-block8": dict(n=5, delta=0.5, theta=0.3, nev=8, sigma=complex(-0.78, -0.002), tol=1e-8,
                           seed=1001, ns=4, inner=8, inner_block=0, guard=2, maxit=800, concurrency=1,
                           desc="He, NN=11 (D=1.77M), delta 0.5, theta 0.3, 8 eigenpairs nearest "
                                "sigma -0.78-0.002i (block 8 + 2 guard columns), tol 1e-8, "
                                "NS 4, inner solve order 8"),
    # the drop-in FP64 GPU configuration (a solve takes many minutes)
    "he-nn21-block16": dict(n=10, delta=0.4, theta=0.35, nev=16, sigma=complex(-0.70, -0.001), tol=1e-9,
                            seed=1002, ns=1, inner=4, inner_block=4, guard=0, maxit=600, concurrency=1,
                            desc="He, NN=21 (D=85.8M), delta 0.4, theta 0.35, block 16, "
                                 "sigma -0.70-0.001i, tol 1e-9, inner solve order 4 on 4-column groups"),
    # theta-trajectory scan: theta evenly spaced over [0.20, 0.50], NN cycling 11, 13, 15, 17;
    # each point warm-started from the previous one (SIXD_FIX across grids); a step is one point
    "he-theta-scan": dict(n=5, delta=0.5, theta=0.3, nev=8, sigma=complex(-0.78, -0.002), tol=1e-8,
                          seed=1004, ns=3, inner=8, inner_block=0, guard=2, maxit=800, concurrency=1,
                          scan=dict(points=64, theta=(0.20, 0.50), nn=(11, 13, 15, 17)),
                          desc="He theta scan: 64 thetas evenly spaced over [0.20, 0.50], NN cycling "
                               "11/13/15/17, delta 0.5, 8 pairs nearest -0.78-0.002i (block 8 + 2 guard), "
                               "tol 1e-8, NS 3, inner order 8, warm starts"),
    # the mixed-precision configuration (FP64 basis: 2 restarts after the FP32 phase)
    "he-nn25-block8": dict(n=12, delta=0.35, theta=0.3, nev=8, sigma=complex(-0.78, -0.002), tol=1e-7,
                           seed=1003, ns=1, inner=4, inner_block=4, guard=0, maxit=600, concurrency=1,
                           desc="He, NN=25 (D=244M), delta 0.35, theta 0.3, block 8, "
                                "sigma -0.78-0.002i, tol 1e-7, inner solve order 4 on 4-column groups"),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


class Slot:
    """One context (Hamiltonian tables + stream) reused by successive solves."""

    def __init__(self, dex, w: dict, precision: str):
        self.g = dex.HeliumGrid(w["n"], w["delta"], w["theta"], precision=precision)
        self.lock = threading.Lock()


def solve(torch, slot: Slot, w: dict, x, timing: int, inner_fp32: bool = False) -> dict:
    import numpy as np

    g = slot.g
    with slot.lock:
        g.set_kernel_timing(timing)
        if timing:
            g.reset_kernel_stats()
        t0 = time.time()
        r = g.gplmr_eig_device(x, w["sigma"], w["nev"], ns=w["ns"], maxit=w["maxit"], tol=w["tol"],
                               inner_iter=w["inner"], inner_block=w["inner_block"], guard=w["guard"],
                               symmetry=1, fp32_tol=max(w["tol"], 1e-5), inner_fp32=inner_fp32)
        lat = time.time() - t0
        stats = g.kernel_stats() if timing else {}
        g.set_kernel_timing(False)
        res = g.residuals_device(x, r.e)
    conv = int(np.sum(res <= w["tol"]))
    return dict(e=r.e, res=res, conv=conv, lat=lat, iters=r.iterations, matvecs=r.matvecs,
                t_mv=r.t_matvec, t_prec=r.t_prec, t_gram=r.t_gram, t_upd=r.t_update,
                restarts=r.fp64_restarts_used, stats=stats)


def run_form(dex, torch, w: dict, precision: str, seeds, concurrency: int, timing: int, slots=None,
             inner_fp32: bool = False):
    """Solves one start block per seed, up to `concurrency` at once; returns
    (results in seed order, wall seconds of the solves, slots)."""
    if slots is None:
        slots = [Slot(dex, w, precision) for _ in range(concurrency)]
    d = slots[0].g.d
    blocks = []
    for i, sd in enumerate(seeds):
        x = torch.empty(d * w["nev"], dtype=torch.complex128, device="cuda")
        slots[i % len(slots)].g.fill_gaussian_device(x, w["nev"], seed=sd, symmetry=1)
        blocks.append(x)
    torch.cuda.synchronize()
    out = [None] * len(seeds)

    def job(i):
        out[i] = solve(torch, slots[i % len(slots)], w, blocks[i], timing, inner_fp32=inner_fp32)
        r = out[i]
        log(f"  {precision} solve seed {seeds[i]}: {r['lat']:.2f}s iters {r['iters']} matvecs {r['matvecs']} "
            f"converged {r['conv']}/{w['nev']} max res {r['res'].max():.2e}")

    torch.cuda.synchronize()
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=len(slots)) as ex:
        list(ex.map(job, range(len(seeds))))
    torch.cuda.synchronize()
    wall = time.time() - t0
    del blocks
    return out, wall, slots


def algorithmic_roofline(results, sample: int) -> dict:
    """Dominant kernel class by summed device time over the timed solves'
    event-timed launches (a random ~1 in `sample` of them)."""
    tot = {}
    for r in results:
        for k, (ms, n, by, fl) in r["stats"].items():
            a = tot.setdefault(k, [0.0, 0, 0.0, 0.0])
            a[0] += ms
            a[1] += n
            a[2] += by
            a[3] += fl
    tot = {k: v for k, v in tot.items() if v[1] > 0 and v[0] > 0}
    if not tot:
        return {"bound": None, "achieved": None, "peak": None, "unit": None, "frac": None, "traffic": None}
    dom = max(tot, key=lambda k: tot[k][0])
    ms, n, by, fl = tot[dom]
    intensity = fl / by if by > 0 else 0.0
    secs = ms * 1e-3
    if intensity > RIDGE:
        bound, ach, peak, unit = "mfma", fl / secs / 1e12, FP64_PEAK_TFS, "TFLOP/s"
    else:
        bound, ach, peak, unit = "hbm", by / secs / 1e9, HBM_PEAK_GBS, "GB/s"
    allk = {k: {"device_ms": v[0], "launches": v[1], "avg_us": v[0] / v[1] * 1e3,
                "GB/s": v[2] / (v[0] * 1e-3) / 1e9, "TFLOP/s": v[3] / (v[0] * 1e-3) / 1e12,
                "flop_per_byte": v[3] / v[2] if v[2] else None} for k, v in tot.items()}
    return {"kernel": dom, "bound": bound, "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak,
            "traffic": None, "avg_launch_us": ms / n * 1e3, "alg_bytes_per_launch": by / n,
            "alg_flops_per_launch": fl / n,
            "timed_launches": f"HIP events on a random ~1 in {sample} launches of every class" if sample > 1
            else "HIP events on every launch", "kernels": allk}


def cpu_baseline(w: dict, matvecs_per_solve: float, pairs_per_solve: float, budget_s: float = 15.0) -> dict:
    """C oracle HAMMULT on the host cores over a bounded sample, scaled to
    eigenpairs/s with the GPU solves' matvec count (block LA not counted)."""
    import numpy as np

    from oracle.c_oracle import CHammult
    from oracle.hamiltonian import Hamiltonian

    h = Hamiltonian(w["n"], w["delta"], w["theta"])
    c = CHammult(h)
    rng = np.random.default_rng(0)
    x = np.asfortranarray(rng.standard_normal((h.d, 1)) + 1j * rng.standard_normal((h.d, 1)))
    c(x)
    t0 = time.time()
    n = 0
    while time.time() - t0 < budget_s and n < 400:
        c(x)
        n += 1
    tmv = (time.time() - t0) / n
    t_solve = matvecs_per_solve * tmv
    return {
        "value": pairs_per_solve / t_solve if t_solve > 0 else None,
        "unit": "eigenpairs/s",
        "cores": CHammult.max_threads(),
        "kind": "port",
        "sample": (f"{n} C-oracle HAMMULT applications at NN={h.nn} ({tmv * 1e3:.2f} ms each, OpenMP), "
                   f"scaled by the GPU solves' mean {matvecs_per_solve:.0f} matvecs per solve; the CPU's "
                   f"block linear algebra is not counted (an upper bound on CPU eigenpairs/s)"),
        "ms_per_matvec": tmv * 1e3,
    }


def traffic_run(dex, torch, w: dict, out_path: str) -> None:
    """The program the HBM-traffic counter passes profile: a calibration kernel
    of known size (an elementwise multiply reading and writing 1 GiB), then one
    solve of the workload with the library's per-class launch statistics
    (algorithmic bytes per launch) written to ``out_path``."""
    n = 64 * 1024 * 1024
    a = torch.full((n,), 1.0 + 0.5j, dtype=torch.complex128, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    del a, b
    torch.cuda.empty_cache()
    slot = Slot(dex, w, "fp64")
    x = torch.empty(slot.g.d * w["nev"], dtype=torch.complex128, device="cuda")
    slot.g.fill_gaussian_device(x, w["nev"], seed=w["seed"], symmetry=1)
    r = solve(torch, slot, w, x, True)
    with open(out_path, "w") as fh:
        json.dump({k: list(v) for k, v in r["stats"].items()}, fh)
    slot.g.close()
    log(f"traffic run: {r['iters']} iterations, {r['conv']} pairs")


def traffic_from_profile(w_name: str, kernel: str, alg_bytes_per_launch: float):
    """HBM bytes per launch of `kernel` from the committed counter measurement
    of this workload (traffic / algorithmic ratio x this run's algorithmic bytes)."""
    path = ROOT / "profiles" / f"hbm_traffic_{w_name.replace('he-', '').replace('-', '_')}.json"
    if not path.exists():
        return None
    ratio = json.loads(path.read_text())["traffic_over_algorithmic"].get(kernel)
    return None if ratio is None else ratio * alg_bytes_per_launch


def run_scan(w: dict, args) -> int:
    """The theta-scan workload: the first --steps points of the scan, in order
    (warm starts), timed as a whole."""
    import numpy as np

    from doublyexcited_amd.scan import theta_scan

    sc = w["scan"]
    th = np.linspace(sc["theta"][0], sc["theta"][1], sc["points"])
    pts = [((sc["nn"][i % len(sc["nn"])] - 1) // 2, w["delta"], float(th[i])) for i in range(sc["points"])]
    pts = pts[:args.steps]
    kw = dict(sigma=w["sigma"], nev=w["nev"], tol=w["tol"], ns=w["ns"], inner_iter=w["inner"],
              guard=w["guard"], maxit=w["maxit"], seed=w["seed"])
    if args.warmup:
        theta_scan(pts[:1], warm=False, **kw)
    def report(p):
        log(f"  NN={2 * p.n + 1} theta={p.theta:.4f}: {p.seconds:.2f}s iters {p.iterations} "
            f"converged {p.converged_pairs}/{w['nev']} warm {p.warm}")

    t0 = time.time()
    res = theta_scan(pts, warm=True, on_point=report, **kw)
    wall = time.time() - t0
    pairs = sum(p.converged_pairs for p in res)
    lat = [p.seconds for p in res]
    print(json.dumps({
        "metric": "resonance eigenpairs/sec + p99 solve latency; mixed vs FP64 GPU at 1e-7 au",
        "value": pairs / wall if wall > 0 else 0.0, "unit": "eigenpairs/s", "n_gpus": 1,
        "steps": len(pts), "warmup": args.warmup, "ms_per_step": wall / len(pts) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "complex128",
        "data": "synthetic: seeded Philox complex-Gaussian singlet start block for the first point",
        "config": {"workload": args.workload, "desc": w["desc"], "points": len(pts), "parallelism": "1 GPU"},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "converged_pairs": pairs, "iterations": [p.iterations for p in res],
    }))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=os.environ.get("DEX_BENCH_WORKLOAD", "he-nn11-block8"),
                    choices=sorted(WORKLOADS))
    ap.add_argument("--concurrency", type=int, default=0, help="solves in flight (0: workload default)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-inner-fp32", action="store_true",
                    help="mixed form: complex64 Arnoldi basis in the inner solve as well")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-run", metavar="STATS_JSON",
                    help="only run the program the HBM-traffic counter passes profile")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: independent solves "
              "would shard across GPUs with no communication, and this build runs on one GPU")
        return 2
    import numpy as np
    import torch

    import doublyexcited_amd as dex

    w = WORKLOADS[args.workload]
    if args.traffic_run:
        torch.cuda.set_device(0)
        traffic_run(dex, torch, w, args.traffic_run)
        return 0
    if "scan" in w:
        torch.cuda.set_device(0)
        return run_scan(w, args)
    conc = args.concurrency or w["concurrency"]
    conc = max(1, min(conc, args.steps))
    torch.cuda.set_device(0)
    seeds = [w["seed"] + i for i in range(args.steps)]
    slots = [Slot(dex, w, "fp64") for _ in range(conc)]
    if args.warmup:
        log(f"warmup: {args.warmup} fp64 solve(s)")
        run_form(dex, torch, w, "fp64", [w["seed"] + 10000 + i for i in range(args.warmup)], conc, False, slots)
    log(f"timed: {args.steps} fp64 solves, {conc} in flight")
    res64, wall, _ = run_form(dex, torch, w, "fp64", seeds, conc, TIMING_SAMPLE, slots)
    for s in slots:
        s.g.close()
    pairs = sum(r["conv"] for r in res64)
    lat = [r["lat"] for r in res64]
    result = {
        "metric": "resonance eigenpairs/sec + p99 solve latency; mixed vs FP64 GPU at 1e-7 au",
        "value": pairs / wall if wall > 0 else 0.0,
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic: seeded Philox complex-Gaussian singlet start blocks, He Hamiltonian built in-process",
        "config": {"workload": args.workload, "desc": w["desc"], "precision": "fp64",
                   "solves_in_flight": conc, "parallelism": "1 GPU"},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "max": float(max(lat))},
        "converged_pairs": pairs,
        "iterations_mean": float(np.mean([r["iters"] for r in res64])),
        "matvecs_per_s": sum(r["matvecs"] for r in res64) / wall,
        "max_residual": float(max(r["res"].max() for r in res64)),
    }
    if not args.no_mixed:
        log(f"mixed form: {args.steps} solves, {conc} in flight")
        resmx, wall_mx, mslots = run_form(dex, torch, w, "mixed", seeds, conc, False,
                                          inner_fp32=args.mixed_inner_fp32)
        for s in mslots:
            s.g.close()
        de = []
        for a, b in zip(res64, resmx):
            de.append(max(np.min(np.abs(np.asarray(a["e"]) - z)) for z in b["e"]))
        lat_mx = [r["lat"] for r in resmx]
        pm = sum(r["conv"] for r in resmx)
        result["mixed_form"] = {
            "value": pm / wall_mx if wall_mx > 0 else 0.0, "unit": "eigenpairs/s", "wall_s": wall_mx,
            "converged_pairs": pm, "latency_s": {"p50": float(np.percentile(lat_mx, 50)),
                                                  "p99": float(np.percentile(lat_mx, 99))},
            "fp64_restarts_mean": float(np.mean([r["restarts"] for r in resmx])),
            "max_abs_dE_vs_fp64": float(max(de)),
            "ratio_vs_fp64": (pm / wall_mx) / result["value"] if result["value"] > 0 and wall_mx > 0 else None,
        }
    result["roofline"] = algorithmic_roofline(res64, TIMING_SAMPLE)
    if result["roofline"].get("kernel"):
        rf = result["roofline"]
        rf["traffic"] = traffic_from_profile(args.workload, rf["kernel"], rf["alg_bytes_per_launch"])
        rf["traffic_unit"] = "bytes per launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)"
    if not args.no_cpu_baseline:
        try:
            result["cpu_baseline"] = cpu_baseline(w, float(np.mean([r["matvecs"] for r in res64])),
                                                  pairs / args.steps)
        except Exception as exc:  # reported, not fatal
            result["cpu_baseline"] = {"value": None, "error": repr(exc)}
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
