#!/usr/bin/env python3
"""Benchmark: converged resonance eigenpairs/s and p50/p99 solve latency of the
GPLMR / harmonic-Ritz solver on one MI355X, mixed precision beside FP64.

Default workload ``he-served-stream``: the served request stream.  K requests
(``--steps``) arrive as a seeded Poisson process (0.5/s); each draws NN
uniformly from 11..21 (even draws snapped up to the next odd grid, the reference
has no even-grid nucleus), spacing U[0.35, 0.6] bohr, theta U[0.2, 0.5] rad,
Re sigma from {-0.78, -0.69, -0.60} au with Im sigma -0.002 au, block 4, 8 or
16, and tolerance 1e-7, with an iteration budget of 40.  The server admits
requests by device footprint under a 240 GB cap as slots (4) and memory free
up, the smallest waiting footprint first; each admitted solve has its own
context and stream, so solves overlap on the device.  A step is one request of
the stream.  Each form
is served for a fixed horizon (120 s): the timed region is the horizon plus the
last iteration of the solves it stops; requests still waiting at the horizon
are turned away, solves still running return at their next iteration boundary
with the pairs converged by then.  Both forms serve the same requests:

  fp64   every matvec and block operation in FP64 (complex128 vectors);
  mixed  X and AX complex128 (AX = H X in FP64, refreshed every 10 iterations,
         when the residual first falls below 1e-5 and to verify convergence),
         the search blocks W, S, P, their images and the inner solve's basis
         complex64 with FP32 matvecs, every Gram / combination / Cholesky /
         projected pencil in FP64.

Both forms solve each request with the same settings (NS, inner solve, guard
columns: the FP64 form's plan for the 240 GB cap) and the server reserves the
FP64 footprint for each in both, so both forms admit the stream alike and
differ in precision alone; ``--own-plan`` lets the mixed form plan and reserve
for its own, smaller footprint (wider projections, more solves in flight).

``value`` is the mixed form's converged eigenpairs (residual recomputed in FP64
from the returned vector <= tol) per second of its timed region; latency runs
from arrival to completion over the requests that completed within the
horizon.  The FP64 form's figures and the ratio are beside it, with max |dE|
between the forms' converged eigenvalues.

Also measured in the same run:
  roofline      the dominant kernel class of the headline form, from the
                stream's first request on its largest grid served alone for 5
                iterations with every launch timed by HIP events on the
                context's stream; algorithmic bytes (or flops) per launch are
                counted on the host at each launch; the matvec's fraction of
                its bound per grid of the stream;
  cpu_baseline  the oracle's FP64 path on the host cores: the C restatement of
                HAMMULT timed per grid and numpy's block products timed at a
                representative shape, applied to the FP64 solves' own counts
                (matvecs per grid, block-kernel bytes), i.e. timed components of
                a bounded sample scaled to the stream.  ``--cpu-solve-sample
                RID`` (host only, no GPU) times instead the oracle's whole GPLMR
                solve of the stream's request RID for ``--cpu-solve-iterations``
                iterations: the oracle orthogonalises the inner solve column by
                column as the reference does, which the blocked model does not
                price.

The other workloads (``--workload``) are the single-configuration solves
(NN=11 block 8, NN=21 block 16, NN=25 block 8) and the theta scan.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                       [--concurrency C] [--horizon S] [--no-mixed] [--no-cpu-baseline]
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
    # the drop-in FP64 GPU configuration (a solve takes about 10 minutes: --steps 1 --warmup 0)
    "he-nn21-block16": dict(n=10, delta=0.4, theta=0.35, nev=16, sigma=complex(-0.70, -0.001), tol=1e-9,
                            seed=1002, ns=2, inner=4, inner_block=4, guard=0, lock=True, maxit=800,
                            concurrency=1,
                            desc="He, NN=21 (D=85.8M), delta 0.4, theta 0.35, block 16, "
                                 "sigma -0.70-0.001i, tol 1e-9, NS 2, inner solve order 4 on 4-column groups, "
                                 "converged columns locked"),
    # the same in the mixed form, with 2 guard columns (179 GB; FP64 blocks would need 306 GB)
    "he-nn21-block16-mixed": dict(n=10, delta=0.4, theta=0.35, nev=16, sigma=complex(-0.70, -0.001), tol=1e-9,
                                  seed=1002, ns=2, inner=4, inner_block=4, guard=2, lock=True, maxit=700,
                                  concurrency=1, mixed_only=True,
                                  desc="He, NN=21 (D=85.8M), delta 0.4, theta 0.35, block 16 + 2 guard columns, "
                                       "sigma -0.70-0.001i, tol 1e-9, NS 2, inner solve order 4 on 4-column "
                                       "groups, converged columns locked, mixed form"),
    # theta-trajectory scan: theta evenly spaced over [0.20, 0.50], NN cycling 11, 13, 15, 17;
    # each point warm-started from the previous one (SIXD_FIX across grids); a step is one point
    "he-theta-scan": dict(n=5, delta=0.5, theta=0.3, nev=8, sigma=complex(-0.78, -0.002), tol=1e-8,
                          seed=1004, ns=3, inner=8, inner_block=0, guard=2, maxit=800, concurrency=1,
                          scan=dict(points=64, theta=(0.20, 0.50), nn=(11, 13, 15, 17)),
                          desc="He theta scan: 64 thetas evenly spaced over [0.20, 0.50], NN cycling "
                               "11/13/15/17, delta 0.5, 8 pairs nearest -0.78-0.002i (block 8 + 2 guard), "
                               "tol 1e-8, NS 3, inner order 8, warm starts"),
    # the mixed-precision configuration (FP64 basis: 2 restarts after the FP32 phase)
    # FP64 blocks at NN=25 need 285 GB and its FP64 matvec has no exchange-symmetric kernels: mixed form only
    "he-nn25-block8": dict(n=12, delta=0.35, theta=0.3, nev=8, sigma=complex(-0.78, -0.002), tol=1e-7,
                           seed=1003, ns=1, inner=4, inner_block=4, guard=0, lock=True, maxit=600, concurrency=1,
                           mixed_only=True,
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
                               lock=w.get("lock", False), symmetry=1, fp32_tol=max(w["tol"], 1e-5),
                               inner_fp32=inner_fp32)
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


# iterations of the solo request whose every launch is timed for the roofline
SOLO_ITERATIONS = 5

# the served request stream (the metric's configuration)
# the served request stream (the metric's configuration).  Each form is served
# for a fixed horizon: the stream's solves outlast any few-minute run (a block-16
# request on the 21-point grid alone needs about 1.5 s per iteration), so the
# timed region is the horizon, requests still waiting then are turned away and
# solves still running are stopped at their next iteration boundary.
SERVED = dict(seed=1005, rate=0.5, cap=240e9, slots=4, tol=1e-7, maxit=40, horizon=120.0, policy="smallest",
              desc="served request stream: Poisson arrivals (seed 1005); NN uniform over 11..21 "
                   "(even snapped up), delta U[0.35, 0.6], theta U[0.2, 0.5], Re sigma from "
                   "{-0.78, -0.69, -0.60}, Im sigma -0.002, block from {4, 8, 16}, tol 1e-7; "
                   "admission by device footprint under a 240 GB cap, up to 4 solves in flight, "
                   "smallest waiting footprint first; each form served for a fixed horizon")


def matvec_fraction_by_grid(grids, delta=0.5, theta=0.3, reps=5) -> dict:
    """Per grid of the stream: the exchange-symmetric matvec on one column
    (the solver's form), FP64 on complex128 and FP32 on complex64 columns,
    each pass timed with HIP events (every launch) over ``reps`` calls, against
    the bound max(F / peak, B / BW) of one column with F = (24 NN + 14) NN^6
    flops and B = 32 NN^6 bytes (FP64, 78.6 TFLOP/s) or 16 NN^6 (complex64,
    157.3 TFLOP/s FP32)."""
    import torch

    import doublyexcited_amd as dex

    out = {}
    for nn in sorted(set(grids)):
        n = (nn - 1) // 2
        d = float(nn) ** 6
        row = {}
        for prec, dtype, peak, bpe in (("fp64", torch.complex128, FP64_PEAK_TFS, 32.0),
                                       ("fp32", torch.complex64, 157.3, 16.0)):
            with dex.HeliumGrid(n, delta, theta, precision=prec) as g:
                x = torch.randn(g.m, g.m, dtype=dtype, device="cuda")
                x = (0.5 * (x + x.t())).contiguous().view(-1)
                y = torch.empty_like(x)
                g.matvec_device(x, y, 1, single=prec == "fp32", symmetry=1)  # warm
                g.set_kernel_timing(1)
                t1 = t2 = 0.0
                for _ in range(reps):
                    # events around each pass of this call (resolved by kernel_timing)
                    g.matvec_device(x, y, 1, single=prec == "fp32", symmetry=1)
                    a, b = g.kernel_timing()
                    t1 += a
                    t2 += b
                g.set_kernel_timing(False)
                del x, y
            per = (t1 + t2) * 1e-3 / reps
            bound = max((24.0 * nn + 14.0) * d / (peak * 1e12), bpe * d / (HBM_PEAK_GBS * 1e9))
            row[prec] = {"us_per_column": per * 1e6, "bound_us": bound * 1e6,
                         "frac": bound / per if per > 0 else None,
                         "pass1_us": t1 / reps * 1e3, "symadd_us": t2 / reps * 1e3}
        out[str(nn)] = row
    return out


def _roofline_from_results(results, sample: int) -> dict:
    rows = [{"stats": r.kernel_stats} for r in results if r.ok and r.kernel_stats]
    return algorithmic_roofline(rows, sample)


def oracle_solve_sample(q, iterations: int = 1, seed: int = 0) -> dict:
    """Times the oracle's FP64 solve on the host for request q: GPLMR_EIG as
    the oracle restates it in numpy (gplmr.F90:9-386), with the C HAMMULT
    (hammod.F90:19-67) as MATBLOCK, on q's grid with q's block, guard columns,
    NS and inner order (PRECBLOCK_BASIC on the whole block, whose Gram-Schmidt
    runs column by column as harmonic_ritz_small.F90:28-81 does), from a
    seeded exchange-symmetric Gaussian start block, for the given number of
    iterations (fewer if it converges).  Returns the seconds of those
    iterations and the part spent in the matvecs; setting up the grid's tables
    is timed apart."""
    import numpy as np

    from oracle.c_oracle import CHammult
    from oracle.gplmr import gplmr_eig
    from oracle.hamiltonian import Hamiltonian

    if q.inner_block not in (0, q.nev + q.guard):
        raise ValueError("the oracle's inner solve works on the whole block")
    t0 = time.time()
    h = Hamiltonian(q.n, q.delta, q.theta)
    hm = CHammult(h, c_potential=True)
    nb = q.nev + q.guard
    rng = np.random.default_rng(seed)
    x0 = np.empty((h.d, nb), dtype=np.complex128, order="F")
    for j in range(nb):
        g = (rng.standard_normal((h.m, h.m)) + 1j * rng.standard_normal((h.m, h.m))) / np.sqrt(2.0)
        x0[:, j] = ((g + g.T) * 0.5).reshape(-1)
        del g
    t_setup = time.time() - t0
    mv = [0, 0.0]

    def amul(b):
        t = time.time()
        y = hm(b)
        mv[0] += b.shape[1]
        mv[1] += time.time() - t
        return y

    t1 = time.time()
    r = gplmr_eig(amul, x0, q.sigma, nb, ns=q.ns, maxit=iterations, tol=q.tol, inner_iter=q.inner_iter,
                  nconv=q.nev)
    return {"rid": q.rid, "nn": 2 * q.n + 1, "block": q.nev, "guard": q.guard, "ns": q.ns,
            "inner": q.inner_iter, "iterations": r.iters, "matvecs": mv[0], "seconds": time.time() - t1,
            "matvec_seconds": mv[1], "setup_seconds": t_setup, "converged": bool(r.converged),
            "threads": CHammult.max_threads()}


def run_cpu_solve_sample(rid: int, iterations: int) -> int:
    """Host only: the stream's request rid (the FP64 form's plan) solved by the
    oracle for a few iterations; one JSON line."""
    from doublyexcited_amd.serve import poisson_stream

    sv = SERVED
    reqs = poisson_stream(sv["rate"], max(rid + 1, 20), sv["seed"], precision="fp64", plan=True, cap=sv["cap"],
                          maxit=sv["maxit"], tol=sv["tol"])
    print(json.dumps(oracle_solve_sample(reqs[rid], iterations)))
    return 0


def served_cpu_baseline(res64, budget_s: float = 25.0) -> dict:
    """The oracle's FP64 path on the host: the C HAMMULT timed at every grid of
    the stream and numpy's block products (Gram X^H Y and update X C, the
    operations of the block kernels) timed at NN=15 with 16 columns; each FP64
    request's CPU time is its matvec count x the HAMMULT time of its grid plus
    its block kernels' algorithmic bytes / the numpy block-product bandwidth.
    This prices a CPU program that runs the library's blocked algorithm; the
    oracle's own solve, whose inner solve orthogonalises column by column as
    the reference does, takes longer (``oracle_solve_sample``)."""
    import numpy as np

    from oracle.c_oracle import CHammult
    from oracle.hamiltonian import Hamiltonian

    t_start = time.time()
    grids = sorted({r.nn for r in res64 if r.ok})
    t_mv = {}
    for nn in grids:
        h = Hamiltonian((nn - 1) // 2, 0.5, 0.3)
        c = CHammult(h, c_potential=True)
        x = np.ones(h.d, dtype=np.complex128)
        c(x)
        n, t0 = 0, time.time()
        while n < 3 and (n == 0 or time.time() - t0 < 2.0):
            c(x)
            n += 1
        t_mv[nn] = (time.time() - t0) / n
        del c, x
    # block products at NN=15 (D = 11.4M), 16 columns: bytes moved per second
    d, k = 15 ** 6, 16
    rng = np.random.default_rng(0)
    xb = (rng.standard_normal((d, k)) + 1j * rng.standard_normal((d, k))).astype(np.complex128)
    yb = xb[:, ::-1].copy()
    cm = rng.standard_normal((k, k)) + 0j
    xb.conj().T @ yb
    n, t0 = 0, time.time()
    while n < 2 or (time.time() - t0 < 3.0 and n < 10):
        xb.conj().T @ yb   # reads 2 k columns
        xb @ cm            # reads k, writes k columns
        n += 1
    la_bw = n * 4 * k * d * 16 / (time.time() - t0)
    del xb, yb
    cpu_s, pairs = 0.0, 0
    for r in res64:
        if not r.ok:
            continue
        la_bytes = sum(r.kernel_totals[k][1] for k in ("gram", "combine", "coldots"))
        cpu_s += r.matvecs * t_mv[r.nn] + la_bytes / la_bw
        pairs += r.converged_pairs
    threads = CHammult.max_threads()
    try:
        import subprocess

        lscpu = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True).stdout
        phys = len({ln for ln in lscpu.splitlines() if ln and not ln.startswith("#")})
    except Exception:
        phys = None
    return {
        "value": pairs / cpu_s if cpu_s > 0 else None,
        "unit": "eigenpairs/s",
        "cores": threads,
        "physical_cores": phys,
        "kind": "port",
        "sample": (f"C-oracle HAMMULT timed once per grid of the stream "
                   f"({', '.join(f'NN={nn}: {t * 1e3:.0f} ms' for nn, t in t_mv.items())}; OpenMP, "
                   f"{threads} threads) and numpy block products at NN=15 x 16 columns "
                   f"({la_bw / 1e9:.1f} GB/s); applied to the FP64 solves' matvec counts and block-kernel "
                   f"bytes: {cpu_s:.0f} s of CPU time for {pairs} pairs "
                   f"(the sample took {time.time() - t_start:.0f} s)"),
        "cpu_seconds_for_stream": cpu_s,
    }


def roofline_request(steps: int, form: str, same_plan: bool = False):
    """The request the roofline pass serves alone: the stream's first request
    on its largest grid (widest block on ties), with the solver settings of
    ``form`` and SOLO_ITERATIONS iterations.  Returns (index, request)."""
    from doublyexcited_amd.serve import poisson_stream

    sv = SERVED
    reqs = poisson_stream(sv["rate"], steps, sv["seed"], precision=form, plan=True, cap=sv["cap"],
                          maxit=SOLO_ITERATIONS, tol=sv["tol"], plan_precision="fp64" if same_plan else None)
    big = max(range(len(reqs)), key=lambda i: (2 * reqs[i].n + 1, reqs[i].nev, -i))
    req = reqs[big]
    req.arrival = 0.0
    return big, req


def roofline_pass(steps: int, form: str, same_plan: bool = False) -> dict:
    """The roofline of the dominant kernel class: the roofline request
    (``roofline_request``) served alone in ``form``, every launch bracketed by
    HIP events on its context's stream (concurrent solves would share the
    device inside the events)."""
    from doublyexcited_amd.serve import Server

    big, req = roofline_request(steps, form, same_plan)
    log(f"roofline pass: request {big} (NN={2 * req.n + 1}, block {req.nev}, {form}) alone, "
        f"{SOLO_ITERATIONS} iterations, every launch timed")
    solo = Server(SERVED["cap"], 1, timing=1).run([req], realtime=False)[0]
    rf = _roofline_from_results([solo], 1)
    rf["workload"] = (f"request {big} of the stream (NN={solo.nn}, block {solo.nev}, {form} form) served alone, "
                      f"{solo.iterations} iterations")
    if rf.get("kernel"):
        rf["traffic"] = traffic_from_profile("served", rf["kernel"], rf["alg_bytes_per_launch"])
        rf["traffic_unit"] = "bytes per launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)"
    return rf


def traffic_run_served(out_path: str, steps: int, form: str, same_plan: bool = True) -> int:
    """The program the HBM-traffic counter passes profile for the served
    stream: a calibration kernel of known size (an elementwise multiply reading
    and writing 1 GiB), then the roofline pass's request (``roofline_request``)
    served alone in ``form``; every launch's algorithmic bytes per kernel class
    (the library's totals) go to ``out_path``."""
    import torch

    from doublyexcited_amd.serve import Server

    torch.cuda.set_device(0)
    n = 64 * 1024 * 1024
    a = torch.full((n,), 1.0 + 0.5j, dtype=torch.complex128, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    del a, b
    torch.cuda.empty_cache()
    sv = SERVED
    rid, req = roofline_request(steps, form, same_plan)
    r = Server(sv["cap"], 1).run([req], realtime=False)[0]
    with open(out_path, "w") as fh:
        json.dump({"request": rid, "form": form, "nn": r.nn, "nev": r.nev, "iterations": r.iterations,
                   "totals": {k: list(v) for k, v in (r.kernel_totals or {}).items()}}, fh)
    log(f"traffic run: request {rid} ({form}) NN={r.nn} block {r.nev}: {r.iterations} iterations, "
        f"{r.converged_pairs} pairs")
    return 0


def match_forms(res64, resmx, tol: float, window: float = 1e-4):
    """Eigenvalues of the same request in both forms: each converged mixed pair
    (residual <= tol) against the nearest converged FP64 pair of that request.
    A pair with none within ``window`` au (its level did not converge in the
    FP64 form) is counted, not compared.  Returns (|dE| list, unmatched count)."""
    import numpy as np

    de, unmatched = [], 0
    for a, b in zip(res64, resmx):
        if not (a.ok and b.ok) or a.residuals is None or b.residuals is None:
            continue
        e64 = np.asarray(a.e)[np.asarray(a.residuals) <= tol]
        for z in np.asarray(b.e)[np.asarray(b.residuals) <= tol]:
            dz = float(np.min(np.abs(e64 - z))) if len(e64) else np.inf
            if dz <= window:
                de.append(dz)
            else:
                unmatched += 1
    return de, unmatched


def latency_stats(lat) -> dict | None:
    import numpy as np

    if not lat:
        return None
    return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)), "max": float(max(lat)),
            "n": len(lat)}


def run_served(args) -> int:
    import numpy as np
    import torch

    from doublyexcited_amd.serve import Request, Server, poisson_stream

    torch.cuda.set_device(0)
    sv = SERVED
    forms = ["fp64"] if args.no_mixed else [f for f in args.forms.split(",") if f]
    if "fp64" not in forms and not args.only_forms:
        forms.insert(0, "fp64")  # the FP64 form is the reference the mixed one is compared with
    slots = args.concurrency or sv["slots"]
    horizon = args.horizon or sv["horizon"]
    out = {}
    for form in forms:
        if args.warmup:
            warm = [Request(arrival=0.0, rid=i, n=5, delta=0.5, theta=0.3, sigma=-0.78 - 0.002j, nev=4,
                            tol=1e-7, precision=form, seed=777 + i, inner_iter=8, guard=2, ns=4, maxit=400)
                    for i in range(args.warmup)]
            Server(sv["cap"], slots).run(warm, realtime=False)
        reqs = poisson_stream(sv["rate"], args.steps, sv["seed"], precision=form, plan=True, cap=sv["cap"],
                              maxit=args.maxit or sv["maxit"], tol=sv["tol"],
                              plan_precision="fp64" if args.same_plan else None)
        if args.same_plan:
            for q in reqs:
                q.reserve_precision = "fp64"  # admitted as the FP64 form's requests are
        log(f"timed: {len(reqs)} {form} requests, rate {sv['rate']}/s, {slots} slots")
        srv = Server(sv["cap"], slots, progress=10, policy=sv["policy"])
        torch.cuda.synchronize()
        t0 = time.time()
        res = srv.run(reqs, realtime=True, horizon=horizon, on_result=lambda r, q: log(
            f"  {form} request {r.rid}: NN={r.nn} block {r.nev} ns {q.ns} inner {q.inner_iter}/{q.inner_block} "
            f"guard {q.guard}: {r.converged_pairs}/{r.nev} pairs, {r.iterations} iterations, "
            f"latency {r.latency:.1f}s service {r.service:.1f}s" + (" (stopped at the horizon)" if r.stopped else "")
            + (f" ERROR {r.error}" if not r.ok else "")))
        torch.cuda.synchronize()
        wall = time.time() - t0
        out[form] = (reqs, res, wall)
    if "fp64" not in out:  # a probe of the mixed form alone (--only-forms)
        res, wall = out[forms[0]][1], out[forms[0]][2]
        print(json.dumps({"form": forms[0], "wall_s": wall, "converged_pairs": sum(r.converged_pairs for r in res),
                          "latency_s": [r.latency for r in res], "iterations": [r.iterations for r in res],
                          "requests": [{"nn": r.nn, "nev": r.nev, "iterations": r.iterations,
                                        "pairs": r.converged_pairs, "service_s": r.service} for r in res]}))
        return 0
    def form_summary(reqs, res, wall):
        done = [r for r in res if r.ok and not r.stopped]
        lat = [r.latency for r in done]
        pairs = sum(r.converged_pairs for r in res)
        return {
            "value": pairs / wall if wall > 0 else 0.0, "unit": "eigenpairs/s", "wall_s": wall,
            "converged_pairs": pairs, "requested_pairs": sum(q.nev for q in reqs),
            "latency_s": latency_stats(lat),
            "completed_requests": len(done),
            "stopped_at_horizon": sum(r.ok and r.stopped for r in res),
            "stopped_pairs": sum(r.converged_pairs for r in res if r.ok and r.stopped),
            "turned_away": sum((not r.ok) and "window" in (r.error or "") for r in res),
            "failed_requests": sum((not r.ok) and "window" not in (r.error or "") for r in res),
            "iterations": [r.iterations for r in res],
            "matvecs": sum(r.matvecs for r in res),
        }

    res64 = out["fp64"][1]
    wall64 = out["fp64"][2]
    summary = {"fp64": form_summary(*out["fp64"])}
    pairs64 = summary["fp64"]["converged_pairs"]
    head = "fp64"
    if "mixed" in out:
        resmx, wallmx = out["mixed"][1], out["mixed"][2]
        pairsmx = sum(r.converged_pairs for r in resmx)
        de, unmatched = match_forms(res64, resmx, sv["tol"])
        summary["mixed"] = form_summary(*out["mixed"])
        summary["mixed"].update({
            "fp64_refreshes_mean": float(np.mean([r.fp64_restarts for r in resmx if r.ok])),
            "max_abs_dE_vs_fp64": float(max(de)) if de else None,
            "pairs_compared_with_fp64": len(de),
            "pairs_without_converged_fp64_counterpart": unmatched,
            "ratio_vs_fp64": (pairsmx / wallmx) / (pairs64 / wall64) if pairs64 > 0 else None,
        })
        head = "mixed"
    reqs, res, wall = out[head]
    lat = [r.latency for r in res if r.ok and not r.stopped]
    pairs = sum(r.converged_pairs for r in res)
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
        "dtype": ("complex64 search blocks (FP32 matvec) + complex128 X/AX (FP64)" if head == "mixed"
                  else "complex128"),
        "data": "synthetic: seeded request parameters and Philox complex-Gaussian singlet start blocks, "
                "He Hamiltonian built per request",
        "config": {"workload": "he-served-stream", "desc": SERVED["desc"], "form": head,
                   "rate_per_s": SERVED["rate"], "requests": args.steps,
                   "iteration_budget_per_request": args.maxit or SERVED["maxit"],
                   "horizon_s": horizon, "admission": SERVED["policy"],
                   "solver_settings": "the FP64 form's plan and reservation in both forms" if args.same_plan
                   else "each form's own plan and footprint", "parallelism": "1 GPU"},
        "latency_s": latency_stats(lat),
        "latency_note": "arrival to completion of the requests that completed within the horizon "
                        "(converged or at their iteration budget); requests stopped at the horizon "
                        "or turned away are counted in forms.*",
        "converged_pairs": pairs,
        "forms": summary,
        "requests": [{"rid": q.rid, "nn": 2 * q.n + 1, "nev": q.nev, "ns": q.ns, "inner": q.inner_iter,
                      "inner_block": q.inner_block, "guard": q.guard} for q in reqs],
    }
    result["roofline"] = roofline_pass(args.steps, head, args.same_plan)
    result["matvec_roofline_by_grid"] = matvec_fraction_by_grid([2 * q.n + 1 for q in reqs])
    if not args.no_cpu_baseline:
        try:
            result["cpu_baseline"] = served_cpu_baseline(res64)
        except Exception as exc:  # reported, not fatal
            result["cpu_baseline"] = {"value": None, "error": repr(exc)}
    print(json.dumps(result))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=os.environ.get("DEX_BENCH_WORKLOAD", "he-served-stream"),
                    choices=sorted(WORKLOADS) + ["he-served-stream"])
    ap.add_argument("--concurrency", type=int, default=0, help="solves in flight (0: workload default)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed", help="served stream: the forms to serve")
    ap.add_argument("--only-forms", action="store_true",
                    help="served stream: serve only --forms (no FP64 reference; a probe, not a bench line)")
    ap.add_argument("--horizon", type=float, default=0.0,
                    help="served stream: seconds each form is served (0: the workload's, 120)")
    ap.add_argument("--own-plan", dest="same_plan", action="store_false",
                    help="served stream: the mixed form plans its own solver settings (NS, inner solve) "
                         "for its smaller footprint; default: both forms use the FP64 form's settings")
    ap.add_argument("--maxit", type=int, default=0,
                    help="served stream: iteration budget per request (0: the workload's, 40)")
    ap.add_argument("--mixed-inner-fp32", action="store_true",
                    help="mixed form: complex64 Arnoldi basis in the inner solve as well")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--roofline-only", action="store_true",
                    help="served stream: only the roofline pass of the last of --forms (the program the "
                         "kernel-statistics profile runs); prints its roofline object")
    ap.add_argument("--traffic-run", metavar="STATS_JSON",
                    help="only run the program the HBM-traffic counter passes profile")
    ap.add_argument("--cpu-solve-sample", type=int, metavar="RID", default=-1,
                    help="served stream, host only: time the oracle's GPLMR solve of request RID")
    ap.add_argument("--cpu-solve-iterations", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: independent solves "
              "would shard across GPUs with no communication, and this build runs on one GPU")
        return 2
    if args.workload == "he-served-stream":
        if args.cpu_solve_sample >= 0:
            return run_cpu_solve_sample(args.cpu_solve_sample, args.cpu_solve_iterations)
        if args.traffic_run:
            return traffic_run_served(args.traffic_run, args.steps, args.forms.split(",")[-1], args.same_plan)
        if args.roofline_only:
            import torch

            torch.cuda.set_device(0)
            print(json.dumps(roofline_pass(args.steps, args.forms.split(",")[-1], args.same_plan)))
            return 0
        return run_served(args)
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
    # a workload whose settings fit the device only with complex64 search blocks runs the mixed form alone
    head = "mixed" if w.get("mixed_only") else "fp64"
    slots = [Slot(dex, w, head) for _ in range(conc)]
    if args.warmup:
        log(f"warmup: {args.warmup} {head} solve(s)")
        run_form(dex, torch, w, head, [w["seed"] + 10000 + i for i in range(args.warmup)], conc, False, slots)
    log(f"timed: {args.steps} {head} solves, {conc} in flight")
    res64, wall, _ = run_form(dex, torch, w, head, seeds, conc, TIMING_SAMPLE, slots)
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
        "dtype": "complex128" if head == "fp64" else "complex64 search blocks (FP32 matvec) + complex128 X/AX (FP64)",
        "data": "synthetic: seeded Philox complex-Gaussian singlet start blocks, He Hamiltonian built in-process",
        "config": {"workload": args.workload, "desc": w["desc"], "precision": head,
                   "solves_in_flight": conc, "parallelism": "1 GPU"},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "max": float(max(lat))},
        "converged_pairs": pairs,
        "iterations_mean": float(np.mean([r["iters"] for r in res64])),
        "matvecs_per_s": sum(r["matvecs"] for r in res64) / wall,
        "max_residual": float(max(r["res"].max() for r in res64)),
    }
    if not args.no_mixed and head == "fp64":
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
    if not args.no_cpu_baseline and head == "fp64":
        try:
            result["cpu_baseline"] = cpu_baseline(w, float(np.mean([r["matvecs"] for r in res64])),
                                                  pairs / args.steps)
        except Exception as exc:  # reported, not fatal
            result["cpu_baseline"] = {"value": None, "error": repr(exc)}
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
