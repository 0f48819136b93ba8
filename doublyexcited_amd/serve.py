"""Continuous-batching service for resonance solves on one GPU.

Requests (grid N, spacing, scaling angle theta, complex shift sigma, block
size, tolerance, precision) arrive over time.  The scheduler admits a request
as soon as (a) a worker slot is free and (b) its device footprint fits, together
with every solve already on the GPU, under the HBM cap.  Each admitted solve owns
a context (its own HIP stream and buffers), so solves overlap on the device.
Results carry arrival->completion latency.  A request that fails, whether it is
rejected at admission or errors in the solver, yields an error result; it never
stops the service.

The reference runs one solve per process (`SAVE`d scratch, hammod.F90:153-159
and :872-896); per-request contexts replace that restriction.
"""
from __future__ import annotations

import dataclasses
import heapq
import threading
import time

import numpy as np


@dataclasses.dataclass(order=True)
class Request:
    arrival: float
    rid: int
    n: int = dataclasses.field(compare=False)          # half-width N, NN = 2N + 1
    delta: float = dataclasses.field(compare=False)
    theta: float = dataclasses.field(compare=False)
    sigma: complex = dataclasses.field(compare=False)
    nev: int = dataclasses.field(compare=False)
    tol: float = dataclasses.field(compare=False)
    precision: str = dataclasses.field(compare=False, default="fp64")
    maxit: int = dataclasses.field(compare=False, default=3000)
    seed: int = dataclasses.field(compare=False, default=0)
    inner_iter: int = dataclasses.field(compare=False, default=0)   # GM_MAX_INNER_ITER
    inner_block: int = dataclasses.field(compare=False, default=0)
    guard: int = dataclasses.field(compare=False, default=0)
    ns: int = dataclasses.field(compare=False, default=1)            # GPLMR_NS
    # admission reserves the footprint of this precision's storage (None: the request's own), so
    # that two forms of a stream can be admitted identically
    reserve_precision: str | None = dataclasses.field(compare=False, default=None)

    def need(self) -> int:
        """Device bytes the server reserves for this request."""
        return footprint_bytes(self.n, self.nev, self.reserve_precision or self.precision, self.ns,
                               self.inner_iter, self.inner_block, self.guard)


@dataclasses.dataclass
class Result:
    rid: int
    ok: bool
    e: np.ndarray | None
    residuals: np.ndarray | None
    converged_pairs: int
    iterations: int
    latency: float           # completion - arrival (s)
    service: float           # completion - admission (s)
    error: str | None = None
    matvecs: int = 0
    fp64_restarts: int = 0
    nn: int = 0
    nev: int = 0
    kernel_stats: dict | None = None    # sampled event timings per kernel class
    kernel_totals: dict | None = None   # every launch: (count, algorithmic bytes, flops) per class
    stopped: bool = False               # ended unconverged by the service's horizon


def snap_nn(nn: int) -> int:
    """Grid points per axis -> the reference's half-width N.  Even NN has no
    reference counterpart (the nucleus would sit between grid points, a path
    that STOPs at hammod.F90:310); it is snapped up to NN + 1."""
    if nn % 2 == 0:
        nn += 1
    return (nn - 1) // 2


def footprint_bytes(n: int, nev: int, precision: str, ns: int = 1, inner_iter: int = 0,
                    inner_block: int = 0, guard: int = 0, inner_fp32: bool = False) -> int:
    """Device bytes a solve holds, the library's own count (dex_solve_footprint,
    no device needed): the caller's block X and the guard columns, AX, the
    (5 + 2 NS) (NEV + guard) search-block vectors (complex64 in the mixed form),
    the inner solve's basis for an order m >= 2 on groups of g columns
    ((m + 2) g vectors) and the context's fixed buffers (tables, block-kernel
    scratch, matvec intermediates)."""
    import ctypes as C

    from . import _lib
    from .hamiltonian import HeliumGrid

    lib = _lib.load()
    p = _lib.DexParams()
    lib.dex_params_default(C.byref(p))
    p.n = n
    p.precision = {"fp64": _lib.DEX_FP64, "fp32": _lib.DEX_FP32, "mixed": _lib.DEX_FP32}[precision]
    o = HeliumGrid._opts(False, False, 0, 2, 1e-5, 0, False, inner_block, guard, inner_fp32)
    b = int(lib.dex_solve_footprint(C.byref(p), nev, ns, inner_iter, C.cast(C.byref(o), C.c_void_p)))
    if b < 0:
        raise ValueError("invalid solve parameters")
    return b


def plan_solve(n: int, nev: int, precision: str, cap: float = 240e9) -> dict:
    """Solver settings for a request: the widest GPLMR projection whose device
    footprint fits the cap.

    Guard columns (2 for blocks of 8 or fewer) let a block whose last wanted
    level is degenerate hold the whole level; the inner solve (order 8 up to
    NN=13, 6 above) is the reference's PRECBLOCK_BASIC; NS is as large as one
    launch of the block kernels takes (NS (NEV + guard) <= 48, (3 + NS)
    (NEV + guard) <= 96, at most 4).  Where that does not fit the cap, NS
    drops, then the inner solve works on 4-column groups, then the guard
    columns go, then the inner order falls.  The complex64 search blocks of the
    mixed form fit wider projections than FP64 at the same grid."""
    nn = 2 * n + 1
    guard = 2 if nev <= 8 else 0
    inner = 8 if nn <= 13 else 6
    iblk = 0 if nn <= 15 else 4

    def ns_max(g):
        nb = nev + g
        return max(1, min(4, 48 // nb, 96 // nb - 3))

    ns = ns_max(guard)
    while footprint_bytes(n, nev, precision, ns, inner, iblk, guard) > cap:
        if ns > 1:
            ns -= 1
        elif iblk == 0 or iblk > 4:
            iblk = 4
        elif guard:
            guard = 0
            ns = ns_max(0)
        elif inner > 2:
            inner -= 2
        else:
            break
    return dict(ns=ns, inner_iter=inner, inner_block=iblk, guard=guard)


def poisson_stream(rate: float, count: int, seed: int, nn_choices=range(11, 22),
                   delta_range=(0.35, 0.6), theta_range=(0.2, 0.5),
                   sigma_re=(-0.78, -0.69, -0.60), sigma_im=-0.002, nev_choices=(4, 8, 16),
                   tol=1e-7, precision="fp64", plan: bool = False, cap: float = 240e9, maxit: int = 3000,
                   plan_precision: str | None = None):
    """Seeded request stream with Poisson arrivals (rate per second).  With
    ``plan`` each request carries the solver settings of ``plan_solve`` for
    its precision (or ``plan_precision``) and the cap."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        nn = int(rng.choice(list(nn_choices)))
        r = Request(arrival=t, rid=i, n=snap_nn(nn), delta=float(rng.uniform(*delta_range)),
                    theta=float(rng.uniform(*theta_range)),
                    sigma=complex(float(rng.choice(sigma_re)), sigma_im),
                    nev=int(rng.choice(nev_choices)), tol=tol, precision=precision,
                    seed=seed * 100003 + i, maxit=maxit)
        if plan:
            for k, v in plan_solve(r.n, r.nev, plan_precision or precision, cap).items():
                setattr(r, k, v)
        out.append(r)
    return out


class Server:
    """``timing`` > 0 brackets a random ~1 in ``timing`` launches of each
    solve's kernels with HIP events (1: every launch) and returns the per-class
    statistics with the result; ``progress`` > 0 has each solve report on stderr
    every ``progress`` iterations."""

    def __init__(self, hbm_cap_bytes: float = 240e9, slots: int = 4, device: int = 0, timing: int = 0,
                 progress: int = 0, policy: str = "fifo"):
        if policy not in ("fifo", "smallest"):
            raise ValueError("policy must be 'fifo' or 'smallest'")
        self.cap = float(hbm_cap_bytes)
        self.slots = slots
        # "fifo": requests are admitted in arrival order (one that does not fit yet holds the
        # queue); "smallest": the waiting request with the smallest footprint that fits
        # goes first (a large request waits while smaller ones keep the device busy)
        self.policy = policy
        self.device = device
        self.timing = int(timing)
        self.progress = int(progress)  # > 0: a progress line to stderr every this many iterations
        self._lock = threading.Lock()
        self._in_use = 0.0
        self._free_slots = slots
        self._cv = threading.Condition(self._lock)
        self._active: dict = {}     # rid -> grid of each solve in flight
        self._stopping = False      # the horizon has passed: solves are asked to stop

    def _register(self, rid: int, g) -> None:
        with self._cv:
            self._active[rid] = g
            if self._stopping:
                g.stop()

    def _stop_all(self) -> None:
        with self._cv:
            self._stopping = True
            for g in self._active.values():
                g.stop()

    def _execute(self, req: Request):
        """The request's solve on the device: (solve result, FP64 residuals of
        the returned pairs, sampled kernel statistics, kernel totals).  The
        grid is registered for ``_stop_all`` while its solve runs."""
        import torch

        import doublyexcited_amd as dex

        torch.cuda.set_device(self.device)
        with dex.HeliumGrid(req.n, req.delta, req.theta, precision=req.precision, device=self.device) as g:
            x = torch.empty(g.d * req.nev, dtype=torch.complex128, device=f"cuda:{self.device}")
            g.fill_gaussian_device(x, req.nev, seed=req.seed, symmetry=1)
            if self.timing:
                g.set_kernel_timing(self.timing)
            self._register(req.rid, g)
            try:
                r = g.gplmr_eig_device(x, req.sigma, req.nev, ns=req.ns, maxit=req.maxit, tol=req.tol,
                                       symmetry=1, fp32_tol=max(req.tol, 1e-5), inner_iter=req.inner_iter,
                                       inner_block=req.inner_block, guard=req.guard,
                                       verbose=2 if self.progress else False, outmodulus=self.progress)
            finally:
                with self._cv:
                    self._active.pop(req.rid, None)
            stats = g.kernel_stats() if self.timing else None
            totals = g.kernel_totals()
            g.set_kernel_timing(False)
            res = g.residuals_device(x, r.e)
            del x
        return r, res, stats, totals

    def _solve(self, req: Request, t_admit: float, t0: float) -> Result:
        """Runs one request; t_admit and the returned times are relative to t0."""
        try:
            r, res, stats, totals = self._execute(req)
            done = time.time() - t0
            conv = int(np.sum(res <= req.tol))
            # ended early, unconverged, after the horizon's stop request
            stopped = self._stopping and not r.converged and r.iterations < req.maxit
            return Result(req.rid, True, r.e, res, conv, r.iterations, done - req.arrival,
                          done - t_admit, matvecs=r.matvecs, fp64_restarts=r.fp64_restarts_used,
                          nn=2 * req.n + 1, nev=req.nev, kernel_stats=stats, kernel_totals=totals,
                          stopped=bool(stopped))
        except Exception as exc:  # isolate the failure to this request
            done = time.time() - t0
            return Result(req.rid, False, None, None, 0, 0, done - req.arrival, done - t_admit,
                          error=repr(exc))

    def run(self, requests: list[Request], realtime: bool = True, on_result=None,
            window: float | None = None, horizon: float | None = None) -> list[Result]:
        """Serves the requests (arrival times relative to the start); returns
        results in request order.  ``on_result(result, request)`` is called as
        each request completes.  With ``window`` (seconds) the service admits
        nothing after that time: requests still waiting then are turned away
        (an error result each) and the solves in flight finish.  With
        ``horizon`` (seconds) the service ends then: admission closes as with
        a window, and every solve still in flight is asked to return at its
        next iteration boundary (its result holds the pairs converged by then
        and ``stopped`` is set)."""
        pending = sorted(requests)
        results: dict[int, Result] = {}
        threads: list[threading.Thread] = []
        self._stopping = False
        t0 = time.time()
        waiting: list = []  # heap of (arrival, rid, req) that arrived but are not admitted
        close = min(v for v in (window, horizon) if v is not None) if (window, horizon) != (None, None) else None
        i = 0
        while i < len(pending) or waiting:
            now = time.time() - t0
            if close is not None and now > close:
                window = close
                closed = [w[2] for w in waiting] + list(pending[i:])
                with self._cv:
                    for req in closed:
                        results[req.rid] = Result(req.rid, False, None, None, 0, 0, now - min(now, req.arrival),
                                                  0.0, error=f"not admitted within the {window:.0f} s window",
                                                  nn=2 * req.n + 1, nev=req.nev)
                waiting, i = [], len(pending)
                break
            while i < len(pending) and (pending[i].arrival <= now or not realtime):
                heapq.heappush(waiting, (pending[i].arrival, pending[i].rid, pending[i]))
                i += 1
            admitted = False
            with self._cv:
                if waiting and self._free_slots > 0:
                    pick = 0
                    if self.policy == "smallest":
                        # the smallest waiting request that fits now (else the smallest, which waits)
                        needs = [w[2].need() for w in waiting]
                        order = sorted(range(len(waiting)), key=lambda k: (needs[k], waiting[k][:2]))
                        fits = [k for k in order if self._in_use + needs[k] <= self.cap or needs[k] > self.cap]
                        pick = fits[0] if fits else order[0]

                    def take(k=pick):
                        if k == 0:
                            heapq.heappop(waiting)
                        else:
                            waiting[k] = waiting[-1]
                            waiting.pop()
                            heapq.heapify(waiting)

                    _, _, req = waiting[pick]
                    need = req.need()
                    if need > self.cap:
                        take()
                        results[req.rid] = Result(req.rid, False, None, None, 0, 0, 0.0, 0.0,
                                                  error=f"admission: {need/1e9:.1f} GB exceeds the "
                                                        f"{self.cap/1e9:.0f} GB cap")
                        admitted = True
                    elif self._in_use + need <= self.cap:
                        take()
                        self._in_use += need
                        self._free_slots -= 1
                        t_admit = time.time() - t0

                        def work(req=req, need=need, t_admit=t_admit):
                            res = self._solve(req, t_admit, t0)
                            if on_result is not None:
                                on_result(res, req)
                            with self._cv:
                                results[req.rid] = res
                                self._in_use -= need
                                self._free_slots += 1
                                self._cv.notify_all()

                        th = threading.Thread(target=work, daemon=True)
                        th.start()
                        threads.append(th)
                        admitted = True
                if not admitted:
                    timeout = 0.05
                    if i < len(pending) and realtime:
                        timeout = max(0.0, min(timeout, pending[i].arrival - (time.time() - t0)))
                    self._cv.wait(timeout=timeout if timeout > 0 else 0.001)
        if horizon is not None:
            with self._cv:
                while any(th.is_alive() for th in threads) and time.time() - t0 < horizon:
                    self._cv.wait(timeout=max(0.001, min(0.5, horizon - (time.time() - t0))))
            self._stop_all()
        for th in threads:
            th.join()
        return [results[r.rid] for r in requests]


def summarize(results: list[Result], wall: float) -> dict:
    lat = [r.latency for r in results if r.ok]
    pairs = sum(r.converged_pairs for r in results)
    return {
        "requests": len(results),
        "ok": sum(r.ok for r in results),
        "converged_pairs": pairs,
        "eigenpairs_per_s": pairs / wall if wall > 0 else 0.0,
        "p50_latency_s": float(np.percentile(lat, 50)) if lat else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if lat else None,
        "wall_s": wall,
    }
