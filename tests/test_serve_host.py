"""Host-side logic of the serving layer (no GPU): request generation, grid
snapping, footprints and admission bookkeeping."""
import numpy as np

from doublyexcited_amd.serve import Request, Server, footprint_bytes, poisson_stream, snap_nn


def test_snap_nn_rounds_even_grids_up():
    assert snap_nn(11) == 5 and snap_nn(12) == 6 and snap_nn(21) == 10 and snap_nn(20) == 10


def test_footprint_counts_blocks_and_context_buffers():
    d = 21 ** 6
    vec = 16 * d
    # the context's own buffers: the matvec intermediate U, the tables and the
    # block kernels' scratch (a 134 MB partial-sum buffer among them)
    fixed = footprint_bytes(10, 1, "fp64") - (2 + 7) * vec
    assert fixed - vec > 134e6
    # X and AX complex128, (5 + 2 NS) NEV search-block vectors: complex128, or
    # complex64 in the mixed form
    assert footprint_bytes(10, 16, "fp64") == (2 + 7) * 16 * vec + fixed
    assert footprint_bytes(10, 16, "mixed") == 2 * 16 * vec + 7 * 16 * (vec // 2) + fixed
    assert footprint_bytes(10, 16, "fp64", ns=2) - footprint_bytes(10, 16, "fp64") == 2 * 16 * vec
    # an inner solve of order 4 on groups of 4 adds (4 + 2) 4 vectors of its format
    assert footprint_bytes(10, 16, "fp64", 1, 4, 4) - footprint_bytes(10, 16, "fp64") == 6 * 4 * vec
    assert footprint_bytes(10, 16, "mixed", 1, 4, 4) - footprint_bytes(10, 16, "mixed") == 6 * 4 * (vec // 2)
    # NN=21, block 16 fits the 240 GB cap alone; two of them do not, two mixed ones do
    assert footprint_bytes(10, 16, "fp64") < 240e9 < 2 * footprint_bytes(10, 16, "fp64")
    assert 2 * footprint_bytes(10, 16, "mixed") > 240e9 > footprint_bytes(10, 16, "mixed") + footprint_bytes(10, 8, "mixed")


def test_poisson_stream_is_seeded_and_in_range():
    a = poisson_stream(0.5, 20, seed=1005)
    b = poisson_stream(0.5, 20, seed=1005)
    assert [(r.arrival, r.n, r.delta, r.theta, r.nev) for r in a] == \
           [(r.arrival, r.n, r.delta, r.theta, r.nev) for r in b]
    assert all(5 <= r.n <= 10 for r in a)
    assert all(0.35 <= r.delta <= 0.6 and 0.2 <= r.theta <= 0.5 for r in a)
    assert {r.nev for r in a} <= {4, 8, 16}
    assert np.all(np.diff([r.arrival for r in a]) > 0)


def test_oversized_request_is_rejected_without_stopping_the_service():
    s = Server(hbm_cap_bytes=1e9, slots=2)
    big = Request(arrival=0.0, rid=0, n=10, delta=0.4, theta=0.3, sigma=-0.7 + 0j, nev=16, tol=1e-7)
    out = s.run([big], realtime=False)
    assert not out[0].ok and "admission" in out[0].error


def test_requests_after_the_admission_window_are_turned_away():
    s = Server(hbm_cap_bytes=240e9, slots=2)
    reqs = [Request(arrival=0.5 * i, rid=i, n=5, delta=0.5, theta=0.3, sigma=-0.78 + 0j, nev=4, tol=1e-7)
            for i in range(3)]
    out = s.run(reqs, realtime=True, window=0.0)
    assert all((not r.ok) and "window" in r.error for r in out)
    assert [r.rid for r in out] == [0, 1, 2]


def test_horizon_stops_solves_in_flight_and_turns_away_the_rest():
    """At the horizon the server asks every solve in flight to stop (the
    library returns at the next iteration boundary) and turns away requests
    still waiting.  The device solve is replaced by one that runs until its
    grid is stopped, so only the server's bookkeeping is exercised."""
    import threading
    import time
    from types import SimpleNamespace

    class FakeGrid:
        def __init__(self):
            self.ev = threading.Event()

        def stop(self, stop=True):
            self.ev.set()

    class FakeServer(Server):
        def _execute(self, req):
            g = FakeGrid()
            self._register(req.rid, g)
            try:
                assert g.ev.wait(10.0), "never stopped"
            finally:
                with self._cv:
                    self._active.pop(req.rid, None)
            r = SimpleNamespace(converged=False, iterations=3, e=np.zeros(req.nev, complex), matvecs=12,
                                fp64_restarts_used=0)
            return r, np.array([1e-9] + [1.0] * (req.nev - 1)), None, None

    s = FakeServer(hbm_cap_bytes=240e9, slots=1)
    reqs = [Request(arrival=0.0, rid=i, n=5, delta=0.5, theta=0.3, sigma=-0.78 + 0j, nev=4, tol=1e-7, maxit=40)
            for i in range(3)]
    t0 = time.time()
    out = s.run(reqs, realtime=True, horizon=0.3)
    assert time.time() - t0 < 5.0
    assert out[0].ok and out[0].stopped and out[0].converged_pairs == 1 and out[0].iterations == 3
    # one slot: the other two were still waiting at the horizon
    assert all((not r.ok) and "window" in r.error for r in out[1:])
    # a request registered after the horizon is stopped at once
    g = FakeGrid()
    s._register(99, g)
    assert g.ev.is_set()


def test_smallest_first_policy_admits_by_footprint():
    """policy="smallest": with one slot and every request waiting, requests run
    in order of their device footprint; "fifo" keeps arrival order."""
    from types import SimpleNamespace

    class FakeServer(Server):
        def _execute(self, req):
            r = SimpleNamespace(converged=True, iterations=1, e=np.zeros(req.nev, complex), matvecs=1,
                                fp64_restarts_used=0)
            return r, np.zeros(req.nev), None, None

    specs = [(10, 16), (5, 4), (7, 8), (5, 16), (6, 4)]
    reqs = [Request(arrival=0.0, rid=i, n=n, delta=0.5, theta=0.3, sigma=-0.78 + 0j, nev=nev, tol=1e-7)
            for i, (n, nev) in enumerate(specs)]
    need = [footprint_bytes(q.n, q.nev, "fp64") for q in reqs]
    for policy, want in (("smallest", sorted(range(5), key=lambda k: need[k])), ("fifo", list(range(5)))):
        seen = []
        out = FakeServer(hbm_cap_bytes=240e9, slots=1, policy=policy).run(
            reqs, realtime=False, on_result=lambda r, q: seen.append(r.rid))
        assert all(r.ok for r in out)
        assert seen == want, (policy, seen)


def test_reserve_precision_sets_the_admitted_footprint():
    r = Request(arrival=0.0, rid=0, n=10, delta=0.4, theta=0.3, sigma=-0.7 + 0j, nev=16, tol=1e-7,
                precision="mixed", ns=1, inner_iter=4, inner_block=4)
    assert r.need() == footprint_bytes(10, 16, "mixed", 1, 4, 4, 0)
    r.reserve_precision = "fp64"
    assert r.need() == footprint_bytes(10, 16, "fp64", 1, 4, 4, 0) > footprint_bytes(10, 16, "mixed", 1, 4, 4, 0)


def test_plans_fit_the_cap_and_widen_where_memory_allows():
    """plan_solve over the stream's grids and blocks: every plan fits the cap
    by the library's own footprint count; NS respects one launch of the block
    kernels (NS (NEV + guard) <= 48, (3 + NS)(NEV + guard) <= 96); the mixed
    form's smaller search blocks never plan narrower than FP64; and the
    NN=21 block-16 FP64 plan is the one its footprint forces: NS 1, no guard
    columns, the inner solve of order 4 on 4-column groups."""
    from doublyexcited_amd.serve import plan_solve

    cap = 240e9
    for n in range(5, 11):
        for nev in (4, 8, 16):
            plans = {}
            for prec in ("fp64", "mixed"):
                p = plan_solve(n, nev, prec, cap)
                nb = nev + p["guard"]
                assert footprint_bytes(n, nev, prec, p["ns"], p["inner_iter"], p["inner_block"],
                                       p["guard"]) <= cap, (n, nev, prec, p)
                assert 1 <= p["ns"] <= 4 and p["ns"] * nb <= 48 and (3 + p["ns"]) * nb <= 96
                plans[prec] = p
            assert plans["mixed"]["ns"] >= plans["fp64"]["ns"] or plans["mixed"]["guard"] > plans["fp64"]["guard"]
    p21 = plan_solve(10, 16, "fp64", cap)
    assert p21 == dict(ns=1, inner_iter=4, inner_block=4, guard=0)
