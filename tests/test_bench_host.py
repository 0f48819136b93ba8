"""Host-side pieces of the bench's served-stream summary (no GPU): the
mixed-vs-FP64 eigenvalue matching and the latency statistics."""
import numpy as np

import bench
from doublyexcited_amd.serve import Result


def _res(rid, e, res, ok=True):
    return Result(rid, ok, np.asarray(e, complex), np.asarray(res, float), int(np.sum(np.asarray(res) <= 1e-7)),
                  10, 1.0, 1.0)


def test_match_forms_compares_converged_pairs_of_the_same_request():
    res64 = [_res(0, [-0.78 - 0.1j, -0.70 - 0.2j], [1e-8, 1e-3]),   # second FP64 pair not converged
             _res(1, [-0.60 - 0.3j], [1e-9]),
             _res(2, [-0.5], [1e-9], ok=False)]
    resmx = [_res(0, [-0.70 - 0.2j + 2e-9, -0.78 - 0.1j + 3e-9], [1e-8, 1e-8]),
             _res(1, [-0.60 - 0.3j - 1e-9], [1e-3]),                  # mixed pair not converged: not compared
             _res(2, [-0.5], [1e-9])]                                 # failed FP64 request: skipped
    de, unmatched = bench.match_forms(res64, resmx, 1e-7)
    # -0.78-0.1i matches within 3e-9; -0.70-0.2i has no converged FP64 counterpart
    assert len(de) == 1 and abs(de[0] - 3e-9) < 1e-15
    assert unmatched == 1


def test_latency_stats():
    assert bench.latency_stats([]) is None
    st = bench.latency_stats([1.0, 2.0, 3.0, 4.0])
    assert st["n"] == 4 and st["max"] == 4.0 and st["p50"] == 2.5 and 3.9 < st["p99"] <= 4.0


def test_oracle_solve_sample_runs_the_requested_iterations():
    """The host-only timing of the oracle's solve: one iteration of a tiny
    request (NN=5, block 2 + 1 guard, NS 2, inner order 3) counts every
    matvec and times them inside the iteration's seconds."""
    import pytest

    from doublyexcited_amd.serve import Request

    q = Request(arrival=0.0, rid=3, n=2, delta=0.5, theta=0.3, sigma=-0.78 - 0.002j, nev=2, tol=1e-7,
                inner_iter=3, guard=1, ns=2, maxit=40)
    s = bench.oracle_solve_sample(q, iterations=1)
    assert s["iterations"] == 1 and s["nn"] == 5 and s["block"] == 2 and s["guard"] == 1
    # the start block (3 columns), then per iteration the inner solve on W and on each of the
    # NS S blocks (2 order multiplications of <= 3 columns each: H V_k and (H - sigma) V_k)
    # and H W, H S_l
    assert 3 < s["matvecs"] <= 3 + (2 + 1) * (2 * 3 + 1) * 3
    assert 0.0 <= s["matvec_seconds"] <= s["seconds"]
    q.inner_block = 2  # 2-column groups: not what the oracle's inner solve restates
    with pytest.raises(ValueError):
        bench.oracle_solve_sample(q, iterations=1)


def test_roofline_request_is_the_first_request_on_the_largest_grid():
    """The request the roofline pass serves alone: among the stream's 20
    requests, the first on the largest grid with the widest block, with the
    FP64 form's plan when both forms share it, and the pass's own iteration
    count (no GPU: plans use the library's footprint count)."""
    from doublyexcited_amd.serve import poisson_stream

    sv = bench.SERVED
    reqs = poisson_stream(sv["rate"], 20, sv["seed"], precision="mixed", plan=True, cap=sv["cap"],
                          plan_precision="fp64")
    i, q = bench.roofline_request(20, "mixed", same_plan=True)
    key = max((2 * r.n + 1, r.nev) for r in reqs)
    assert (2 * q.n + 1, q.nev) == key
    assert all((2 * r.n + 1, r.nev) < key for r in reqs[:i])
    assert q.arrival == 0.0 and q.maxit == bench.SOLO_ITERATIONS and q.precision == "mixed"
    assert (q.ns, q.inner_iter, q.inner_block, q.guard) == (reqs[i].ns, reqs[i].inner_iter,
                                                            reqs[i].inner_block, reqs[i].guard)
